"""Mixed-precision dense-frame path: the likelihood and tomogram contractions as FP16
products on matrix cores (emc_set_precision modes 1 and 2; 4 and 5 with the likelihood
product in BF16) against the FP64 oracle and the plain FP64 engine path on the same dense
integer frames.

Bar: the merged model and the updated model within 1e-3 relative error (max-norm) of FP64
in the two-term mode; the average log-likelihood and mutual information within 1e-6 (two
terms) resp. 1e-3 (one term). The measured errors are printed for the record."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


def _errs(g, o):
    e = {k: P.rel_err(g[k], o[k]) for k in ('model2', 'inter_weight', 'psum_r', 'likelihood', 'info', 'u')
         if k in g and k in o}
    e['max_exp'] = P.rel_err(g['max_exp'], o['max_exp'])
    e['prob'] = P.rel_err(g['prob'], o['prob'])
    for k in ('avg_likelihood', 'mutual_info', 'rescale'):
        e[k] = abs(g[k] - o[k]) / max(abs(o[k]), 1e-300)
    e['rmax_agree'] = float(np.mean(g['rmax'] == o['rmax']))
    return e


def _run(p, mode, flags, model1=None):
    eng = P.engine_for_dense(p)
    if model1 is not None:
        eng.set_model(model1)
    eng.set_precision(mode)
    g = P.run_engine(eng, flags)
    return eng, g


@pytest.fixture(scope='module')
def bright():
    """Tiny geometry, 300 dense frames of ~400 photons."""
    return P.densify(P.build(P.tiny_config(num_data=300, mean_count=400.)))


@pytest.mark.parametrize('mode', [2, 1, 5, 4])
def test_tiny_dense_vs_oracle(bright, mode):
    """Modes 1, 2: FP16 log W (one term, hi + lo); 4, 5: the same in BF16 (8 mantissa bits
    per term, so one BF16 term moves log-likelihoods by ~1e-3 of a unit per photon)."""
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(bright, flags)
    eng, g = _run(bright, mode, flags)
    eng.close()
    e = _errs(g, o)
    print('mode', mode, {k: '%.2e' % v for k, v in e.items()}, flush=True)
    # the expand and merge are FP64 in both paths
    assert e['u'] <= 1e-12 and e['rescale'] <= 1e-12
    if mode != 4:
        assert e['inter_weight'] <= 1e-12
    else:
        # one BF16 term moves some orientations in or out of every frame's keep set, and the
        # weights sum over the orientations some frame keeps (psum_r > 0): compare that set
        # with the FP64 engine's, which the weights then follow
        eng0, g0 = _run(bright, 0, flags)
        eng0.close()
        moved = np.mean((g['psum_r'] > 0) != (g0['psum_r'] > 0))
        print('mode 4: orientations merged in one path only: %.4f' % moved, flush=True)
        assert moved <= 0.01 and e['inter_weight'] <= 0.1
    if mode == 2:
        assert e['model2'] <= 1e-3
        assert e['avg_likelihood'] <= 1e-6 and e['mutual_info'] <= 1e-6
        assert e['rmax_agree'] == 1.
    elif mode == 5:
        assert e['model2'] <= 1e-2
        assert e['avg_likelihood'] <= 1e-4 and e['mutual_info'] <= 1e-3
    elif mode == 1:
        assert e['model2'] <= 2e-2
        assert e['avg_likelihood'] <= 1e-3 and e['mutual_info'] <= 1e-2
    else:
        assert e['model2'] <= 2e-1
        assert e['avg_likelihood'] <= 1e-2 and e['mutual_info'] <= 1e-1


def test_dense_zero_block_drops_same_orientations(bright):
    """Orientations with log W = -inf on a likelihood pixel are dropped for dense frames in
    the mixed path exactly as in FP64."""
    m = bright['model1'].copy()
    c = m.shape[0] // 2
    m[c + 3:c + 6, c - 2:c + 3, :] = 0.
    p = dict(bright, model1=m)
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(p, flags)
    eng, g = _run(p, 2, flags)
    eng.close()
    assert np.array_equal(g['prob'] == 0, o['prob'] == 0)
    assert P.rel_err(g['model2'], o['model2']) <= 1e-3


def test_cfg1_geometry_dense_two_iterations_vs_fp64_engine():
    """151x151 detector, 125^3 model, 3240 orientations, 600 bright dense frames (several
    tiles along every product dimension, partial panels), two iterations chained through the
    model update with alpha = 0.5, mixed (two-term) against the plain FP64 engine. (Without
    blending, outer-shell voxels that no orientation samples become 0 after the first update,
    and a dense frame then drops every orientation whose slice touches one.)"""
    from dragonfly_amd import synth
    cfg = dict(synth.CONFIGS['cfg1'], num_data=600, mean_count=2000.)
    p = P.densify(P.build(cfg, device='cuda', num_ave=600))
    flags1 = dict(iteration=1, need_scaling=0)
    ref, fp = _run(p, 0, flags1)
    mix, mp = _run(p, 2, flags1)
    for it in (1, 2):
        if it == 2:
            flags = dict(iteration=2, need_scaling=0)
            fp = P.run_engine(ref, flags)
            mp = P.run_engine(mix, flags)
        e = _errs(mp, fp)
        rms_f = ref.update_model(0.5)
        rms_m = mix.update_model(0.5)
        e['model1'] = P.rel_err(mix.get('model1'), ref.get('model1'))
        e['rms_change'] = abs(rms_m - rms_f) / rms_f
        print('cfg1 dense it', it, {k: '%.2e' % v for k, v in e.items()}, flush=True)
        assert e['model2'] <= 1e-3 and e['model1'] <= 1e-3
        # iteration 2 starts from the two paths' own updated models (1e-4 apart), so the
        # per-frame sums differ by more than the single-iteration 1e-6
        tol = 1e-6 if it == 1 else 1e-4
        assert e['avg_likelihood'] <= tol and e['mutual_info'] <= tol
        assert e['rmax_agree'] >= 0.99
        assert (fp['rmax'] >= 0).mean() > 0.9   # frames keep valid orientations
    ref.close()
    mix.close()


def test_mixed_path_refuses_sparse_frames():
    from dragonfly_amd._lib import EMCError
    p = P.build(P.tiny_config(num_data=20))
    eng = P.engine_for(p)
    eng.set_precision(2)
    with pytest.raises(EMCError, match='dense frames'):
        eng.maximize(1, 0, 1, 0)
    eng.set_precision(0)
    eng.maximize(1, 0, 1, 0)
    eng.close()


@pytest.mark.parametrize('scaled', [False, True])
def test_fp32_gathers_vs_oracle(scaled):
    """Precision 3 (FP32 log W and posterior gathers, FP64 sums) on sparse frames, unscaled and
    scaled (iteration 2 uses the per-frame scales), in 64-orientation chunks."""
    import os
    cfg = P.tiny_config(num_data=200, mean_count=200., phi=(0.5, 1.5) if scaled else None)
    p = P.build(cfg)
    fp = dict(p['fparams'])
    if scaled:
        fp['scale'] = np.asarray(p['true_scale'], 'f8').copy()
    p = dict(p, fparams=fp)
    flags = dict(iteration=2 if scaled else 1, need_scaling=int(scaled), update_scale=1)
    o = P.run_oracle(p, flags)
    os.environ['EMC_MAX_CHUNK_COLS'] = '64'
    try:
        eng = P.engine_for(p)
        eng.set_precision(3)
        g = P.run_engine(eng, flags)
        eng.close()
    finally:
        del os.environ['EMC_MAX_CHUNK_COLS']
    e = _errs(g, o)
    print('fp32 gathers scaled=%s' % scaled, {k: '%.2e' % v for k, v in e.items()}, flush=True)
    assert g['num_chunks'] > 1
    # W is interpolated from an FP32 copy of the model in FP32 (u sums those W in FP64); the
    # merge weights are FP64
    assert e['u'] <= 1e-6 and e['inter_weight'] <= 1e-12
    assert e['model2'] <= 1e-4
    assert e['avg_likelihood'] <= 1e-6 and e['mutual_info'] <= 1e-5


@pytest.mark.parametrize('variant', ['default', 'one-chunk', 'no-buckets', 'full-parts', 'full-direct',
                                     'small-units'])
def test_fp32_gathers_refinement_subsets(variant, monkeypatch):
    """Precision 3 on the sparse subset path: partial frames (10 of 60 cells), full frames
    (every cell) and an empty set, in 128-slot chunks (or one), against the oracle's refinement
    branch (tolerances of the FP32 gathers above); also with each kernel variant of the sparse
    M-step: the pair tomogram walking its tiles' pairs instead of per-(tile, range) buckets,
    the full frames through partial views or read directly, merge units of 64 entries."""
    import os
    from oracle import core
    env = {'no-buckets': {'EMC_TOMO_BUCKETS': '0'}, 'full-parts': {'EMC_SIG_FULL_ROWS': '0'},
           'full-direct': {'EMC_SIG_FULL_ROWS': str(1 << 20)}, 'small-units': {'EMC_UNIT_ITEMS': '64'}}
    for k, v in env.get(variant, {}).items():
        monkeypatch.setenv(k, v)
    from dragonfly_amd import _lib
    cfg = P.tiny_config(num_data=120, mean_count=200.)
    p = P.build(cfg)
    qc = _lib.quat_gen(1)
    qmap = _lib.voronoi_subset(qc, p['quats'])
    rng = np.random.default_rng(21)
    D = len(p['fparams']['beta'])
    sets = [np.arange(len(qc), dtype='i4') if d % 3 == 0 else
            np.sort(rng.choice(len(qc), 10, replace=False)).astype('i4') for d in range(D)]
    sets[4] = np.zeros(0, 'i4')
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    ref = dict(quat_mapping=qmap, rel_off=off, rel_quat=np.concatenate(sets))
    flags = dict(iteration=1, need_scaling=0)
    fp = dict(p['fparams'], scale=np.ones(D))
    o = core.maximize(p['det'], p['quats'], p['model1'], p['frames'], fp, flags, dense_prob=True, refine=ref)
    o['scale'] = fp['scale']
    if variant != 'one-chunk':
        os.environ['EMC_MAX_CHUNK_COLS'] = '128'
    try:
        eng = P.engine_for(p)
    finally:
        os.environ.pop('EMC_MAX_CHUNK_COLS', None)
    eng.set_precision(3)
    eng.set_refine(len(qc), qmap, off, ref['rel_quat'])
    g = P.run_engine(eng, flags)
    eng.close()
    e = _errs(g, o)
    print('fp32 gathers, refinement subsets', {k: '%.2e' % v for k, v in e.items()}, flush=True)
    assert (g['num_chunks'] == 1) if variant == 'one-chunk' else (g['num_chunks'] > 1)
    assert g['rmax'][4] == -1
    assert e['u'] <= 1e-6 and e['inter_weight'] <= 1e-12
    assert e['model2'] <= 1e-4
    assert e['avg_likelihood'] <= 1e-6 and e['mutual_info'] <= 1e-5
