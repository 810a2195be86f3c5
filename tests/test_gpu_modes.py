"""Several model modes on the GPU engine (emc_set_modes: combined orientations
r = rotation * num_modes + mode, dragonfly/src/max_prob.c:80-84) against the oracle on the
same seeded inputs. Tolerance: 1e-9 relative for the FP64 path, rmax exact; the FP32-gather
mode at the tolerances of the one-mode FP32 test."""
import numpy as np
import pytest

import problems as P
from dragonfly_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-9
FIELDS_F = ('u', 'max_exp', 'p_norm', 'likelihood', 'info', 'model2', 'inter_weight', 'prob',
            'quat_norm')


@pytest.fixture(scope='module')
def tiny():
    return P.build(P.tiny_config(num_data=60))


def _vols(p, M):
    S = p['model1'].shape[0]
    vols = [p['model1']] + [synth.start_model(S, p['fparams']['mean_count'], p['det']['mask'], seed=30 + k)
                            for k in range(M - 1)]
    return np.ascontiguousarray(np.stack(vols))


def _engine(p, M, model1):
    from dragonfly_amd._lib import Engine
    eng = Engine(0)
    eng.set_detector(p['det']['qvals'], p['det']['corr'], p['det']['mask'])
    eng.set_modes(M)
    eng.set_quats(p['quats'])
    eng.set_model(model1)
    eng.set_frames(*[p['frames'][k] for k in P.SPARSE_KEYS])
    fp = p['fparams']
    eng.set_frame_params(fp['scale'], fp['beta'], fp['blacklist'], fp['fcounts'], fp['sum_fact'],
                         fp['mean_count'], fp['num_blacklist'])
    return eng


def _run(eng, flags):
    g = P.run_engine(eng, flags)
    g['quat_norm'] = eng.get('quat_norm')
    return g


def _compare(o, g, label, scaling=False):
    errs = {k: P.rel_err(g[k], o[k]) for k in FIELDS_F}
    for k in ('rescale', 'avg_likelihood', 'mutual_info'):
        errs[k] = abs(g[k] - o[k]) / max(abs(o[k]), 1e-300)
    if scaling:
        errs['scale'] = P.rel_err(g['scale'], o['scale'])
    print(label, {k: '%.1e' % v for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, '%s: relative errors above %g: %s' % (label, TOL, bad)
    assert np.array_equal(g['rmax'], o['rmax']), label + ': rmax differs'


@pytest.mark.parametrize('M,scaled', [(2, False), (3, False), (2, True)])
def test_modes_one_iteration(tiny, M, scaled):
    model1 = _vols(tiny, M)
    fp = dict(tiny['fparams'])
    if scaled:
        fp['scale'] = np.asarray(tiny['true_scale'], 'f8').copy()
        flags = dict(iteration=2, need_scaling=1, update_scale=1)
    else:
        flags = dict(iteration=1, need_scaling=0)
    p = dict(tiny, fparams=fp)
    print('case: modes M=%d scaled=%s' % (M, scaled), flush=True)
    o = P.run_oracle(p, flags, model1=model1)
    eng = _engine(p, M, model1)
    assert eng.num_rot == M * tiny['quats'].shape[0]
    g = _run(eng, flags)
    eng.close()
    assert g['quat_norm'].shape == (60, M) and g['model2'].shape == model1.shape
    _compare(o, g, 'modes M=%d scaled=%s' % (M, scaled), scaling=scaled)


def test_modes_chain_with_model_update(tiny):
    """Two scaled iterations with two modes chained through update_model (each mode
    symmetrised, rms over both volumes) and normalize_scale, which rescales the first mode's
    volume only (iterate.pyx:373-374)."""
    from oracle import core
    M = 2
    model1 = _vols(tiny, M)
    D = len(tiny['fparams']['beta'])
    scale = np.ones(D)
    eng = _engine(tiny, M, model1)
    for it in (1, 2):
        flags = dict(iteration=it, need_scaling=1, update_scale=1)
        p2 = dict(tiny, fparams=dict(tiny['fparams'], scale=scale.copy()))
        print('case: modes chain it=%d' % it, flush=True)
        o = P.run_oracle(p2, flags, model1=model1)
        g = _run(eng, flags)
        _compare(o, g, 'modes chain it=%d' % it, scaling=True)
        m2 = o['model2'].copy()
        rms_o = core.update_model(model1, m2, o['inter_weight'], 0.)
        rms_g = eng.update_model(0.)
        assert abs(rms_g - rms_o) <= TOL * abs(rms_o)
        scale = o['scale'].copy()
        mean = scale.mean()
        model1[0] *= mean
        scale /= mean
        assert abs(eng.normalize_scale() - mean) <= 1e-14 * mean
        assert P.rel_err(eng.get('model1'), model1) <= TOL
        assert P.rel_err(eng.get('scale'), scale) <= TOL
    eng.close()


@pytest.mark.parametrize('mode', ['2', '2-box', '0'])
def test_modes_chunked_tomogram_and_merge_forms(tiny, monkeypatch, mode):
    """Two modes in 64-orientation chunks (the combined orientations of one rotation sit side
    by side, so modes alternate inside every chunk) with the brick merge (frame-partitioned
    tomogram), the voxel-box merge, and the per-pixel tomogram with the voxel-box merge."""
    monkeypatch.setenv('EMC_MAX_CHUNK_COLS', '64')
    if mode == '2-box':
        monkeypatch.setenv('EMC_MERGE_MODE', '0')
        mode = '2'
    monkeypatch.setenv('EMC_TOMO_TILE_MODE', mode)
    model1 = _vols(tiny, 2)
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(tiny, flags, model1=model1)
    eng = _engine(tiny, 2, model1)
    g = _run(eng, flags)
    eng.close()
    assert g['num_chunks'] > 1
    _compare(o, g, 'modes chunked mode=' + mode)


def test_modes_fp32_gathers():
    """Precision 3 (FP32 log W and posterior gathers, FP64 sums) with two modes."""
    cfg = P.tiny_config(num_data=200, mean_count=200.)
    p = P.build(cfg)
    model1 = _vols(p, 2)
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(p, flags, model1=model1)
    eng = _engine(p, 2, model1)
    eng.set_precision(3)
    g = _run(eng, flags)
    eng.close()
    e = {k: P.rel_err(g[k], o[k]) for k in ('u', 'inter_weight', 'model2', 'quat_norm')}
    for k in ('avg_likelihood', 'mutual_info'):
        e[k] = abs(g[k] - o[k]) / abs(o[k])
    print('modes fp32 gathers', {k: '%.2e' % v for k, v in e.items()}, flush=True)
    # W is interpolated from an FP32 copy of the model in FP32 (u sums those W in FP64); the
    # merge weights are FP64
    assert e['u'] <= 1e-6 and e['inter_weight'] <= 1e-12
    assert e['model2'] <= 1e-4 and e['quat_norm'] <= 1e-4
    assert e['avg_likelihood'] <= 1e-6 and e['mutual_info'] <= 1e-5


def test_modes_argument_checks(tiny):
    """Paths that take one mode refuse several, and the model must hold every mode."""
    from dragonfly_amd._lib import EMCError
    model1 = _vols(tiny, 2)
    eng = _engine(tiny, 2, model1)
    with pytest.raises(ValueError):
        eng.set_model(tiny['model1'])
    with pytest.raises(EMCError):
        eng.set_refine(1, np.zeros(eng.num_rot, 'i4'), np.zeros(61, 'i8'), np.zeros(0, 'i4'))
    eng.set_precision(1)
    with pytest.raises(EMCError):
        eng.maximize(1, 0)
    eng.set_precision(0)
    # back to one mode: the model is dropped until it is set again
    eng.set_modes(1)
    assert eng.num_rot == tiny['quats'].shape[0]
    with pytest.raises(EMCError):
        eng.maximize(1, 0)
    eng.set_model(tiny['model1'])
    flags = dict(iteration=1, need_scaling=0)
    g = P.run_engine(eng, flags)
    eng.close()
    o = P.run_oracle(tiny, flags)
    assert P.rel_err(g['prob'], o['prob']) <= TOL
    assert np.array_equal(g['rmax'], o['rmax'])
