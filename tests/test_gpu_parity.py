"""GPU engine vs CPU oracle on identical seeded inputs: one EMC iteration (FP64 path) and
iterations chained through the model update. Tolerance: 1e-9 relative (the plain FP64
path's parity bar); integer outputs (rmax) must match exactly."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-9
FIELDS_F = ('u', 'max_exp', 'p_norm', 'likelihood', 'info', 'model2', 'inter_weight')


def _compare(o, g, label, scaling=False):
    errs = {k: P.rel_err(g[k], o[k]) for k in FIELDS_F}
    errs['prob'] = P.rel_err(g['prob'], o['prob'])
    for k in ('rescale', 'avg_likelihood', 'mutual_info'):
        errs[k] = abs(g[k] - o[k]) / max(abs(o[k]), 1e-300)
    if scaling:
        errs['psum_d'] = P.rel_err(g['psum_d'], o['psum_d'])
        errs['scale'] = P.rel_err(g['scale'], o['scale'])
    print(label, {k: '%.1e' % v for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, '%s: relative errors above %g: %s' % (label, TOL, bad)
    assert np.array_equal(g['rmax'], o['rmax']), label + ': rmax differs'


@pytest.fixture(scope='module')
def tiny():
    return P.build(P.tiny_config())


@pytest.fixture(scope='module')
def tiny_scaled():
    cfg = P.tiny_config(num_data=80, phi=(0.5, 1.5), gap=2)
    return P.build(cfg)


def test_slice_gen_bit_exact(tiny):
    from oracle import core
    eng = P.engine_for(tiny)
    for r in (0, 5, 77, len(tiny['quats']) - 1):
        q = tiny['quats'][r]
        g = eng.slice_gen(q)
        o = core.slice_gen3d(q, tiny['model1'], tiny['det']['qvals'], tiny['det']['corr'])
        assert np.array_equal(g.view('u8'), o.view('u8')), 'slice_gen differs at r=%d' % r
    eng.close()


def test_slice_merge(tiny):
    from oracle import core
    eng = P.engine_for(tiny)
    eng.clear_merge()
    S = tiny['model1'].shape[0]
    m2 = np.zeros((S,) * 3)
    w2 = np.zeros((S,) * 3)
    rng = np.random.default_rng(7)
    for r in (3, 40):
        view = rng.random(len(tiny['det']['corr']))
        eng.slice_merge(tiny['quats'][r], view)
        core.slice_merge3d(tiny['quats'][r], view, m2, w2, tiny['det']['qvals'], tiny['det']['corr'],
                           tiny['det']['mask'])
    assert P.rel_err(eng.get('model2'), m2) < 1e-14
    assert P.rel_err(eng.get('inter_weight'), w2) < 1e-14
    eng.close()


def test_one_iteration_unscaled(tiny):
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(tiny, flags)
    eng = P.engine_for(tiny)
    g = P.run_engine(eng, flags)
    eng.close()
    _compare(o, g, 'tiny unscaled')


def test_one_iteration_scaled_first_and_later(tiny_scaled):
    p = tiny_scaled
    for it in (1, 2):
        flags = dict(iteration=it, need_scaling=1, update_scale=1)
        p2 = dict(p)
        p2['fparams'] = dict(p['fparams'], scale=np.asarray(p['true_scale'], dtype='f8').copy())
        o = P.run_oracle(p2, flags)
        eng = P.engine_for(p2)
        g = P.run_engine(eng, flags)
        eng.close()
        _compare(o, g, 'tiny scaled it=%d' % it, scaling=True)


def test_blacklist_empty_frames_and_beta(tiny):
    p = dict(tiny)
    f = {k: v.copy() for k, v in tiny['frames'].items()}
    # frame 3 loses all photons (empty frame), frames 1 and 7 blacklisted, per-frame beta
    d = 3
    o0, n0 = f['ones_accum'][d], f['ones'][d]
    m0, nm = f['multi_accum'][d], f['multi'][d]
    f['place_ones'] = np.delete(f['place_ones'], np.arange(o0, o0 + n0))
    f['place_multi'] = np.delete(f['place_multi'], np.arange(m0, m0 + nm))
    f['count_multi'] = np.delete(f['count_multi'], np.arange(m0, m0 + nm))
    f['ones'][d] = 0
    f['multi'][d] = 0
    f['ones_accum'] = np.concatenate([[0], np.cumsum(f['ones'])[:-1]]).astype('i8')
    f['multi_accum'] = np.concatenate([[0], np.cumsum(f['multi'])[:-1]]).astype('i8')
    p['frames'] = f
    D = len(f['ones'])
    bl = np.zeros(D, 'u1')
    bl[[1, 7]] = 1
    beta = np.linspace(0.05, 1.0, D)
    fc = tiny['fparams']['fcounts'].copy()
    fc[d] = 0
    sf = tiny['fparams']['sum_fact'].copy()
    sf[d] = 0.
    p['fparams'] = dict(tiny['fparams'], blacklist=bl, beta=beta, num_blacklist=2, fcounts=fc,
                        sum_fact=sf)
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(p, flags)
    eng = P.engine_for(p)
    g = P.run_engine(eng, flags)
    eng.close()
    assert g['rmax'][1] == -1 and g['rmax'][7] == -1
    _compare(o, g, 'blacklist/empty/beta')


def test_zero_model_region_gives_minus_inf(tiny):
    """Zero voxels give log(0) = -inf in the likelihood, as in the reference (no clamp):
    orientations whose slice crosses the zeroed block drop out of frames with photons there."""
    p = dict(tiny)
    m = tiny['model1'].copy()
    S = m.shape[0]
    c = S // 2
    m[c + 3:c + 6, c - 2:c + 3, :] = 0.
    p['model1'] = m
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(p, flags)
    eng = P.engine_for(p)
    g = P.run_engine(eng, flags)
    eng.close()
    n_minus_inf = int((o['prob'] == 0).sum() - (P.run_oracle(tiny, flags)['prob'] == 0).sum())
    assert n_minus_inf > 0, 'zeroed block removed no orientation'
    assert o['avg_likelihood'] != 0.
    _compare(o, g, 'zero block')


@pytest.mark.parametrize('alpha', [0., 0.5])
def test_three_iterations_with_model_update(tiny_scaled, alpha):
    """Three scaled iterations chained through update_model (with alpha > 0 the blend
    model1 = alpha model1 + (1 - alpha) model2, recon.pyx:140-147) and normalize_scale."""
    from oracle import core
    p = tiny_scaled
    eng = P.engine_for(p)
    model1 = p['model1'].copy()
    scale = np.ones(len(p['frames']['ones']))
    for it in (1, 2, 3):
        flags = dict(iteration=it, need_scaling=1, update_scale=1)
        p2 = dict(p, fparams=dict(p['fparams'], scale=scale.copy()))
        o = P.run_oracle(p2, flags, model1=model1)
        g = P.run_engine(eng, flags)
        _compare(o, g, 'chain alpha=%g it=%d' % (alpha, it), scaling=True)
        # host-side reference sequence: update_model then normalize_scale
        m2 = o['model2'].copy()
        rms_o = core.update_model(model1, m2, o['inter_weight'], alpha)
        rms_g = eng.update_model(alpha)
        scale = o['scale'].copy()
        mean = scale.mean()
        model1 *= mean
        scale /= mean
        mean_g = eng.normalize_scale()
        assert abs(rms_g - rms_o) <= TOL * abs(rms_o)
        assert abs(mean_g - mean) <= 1e-14 * mean
        assert P.rel_err(eng.get('model1'), model1) <= TOL
        assert P.rel_err(eng.get('scale'), scale) <= TOL
    eng.close()


def test_cfg1_geometry_two_iterations():
    """The CPU-runnable configuration at its full size: 125^3 model, 151x151 pixels,
    num_div 4 (3240 orientations), all 3000 frames, beta 1, two iterations chained through
    the model update (the oracle takes ~20 s per iteration on one host thread)."""
    from dragonfly_amd import synth
    from oracle import core
    cfg = dict(synth.CONFIGS['cfg1'])
    assert cfg['num_data'] == 3000
    p = P.build(cfg, device='cuda', num_ave=1000)
    eng = P.engine_for(p)
    model1 = p['model1'].copy()
    for it in (1, 2):
        flags = dict(iteration=it, need_scaling=0)
        o = P.run_oracle(p, flags, model1=model1)
        g = P.run_engine(eng, flags)
        _compare(o, g, 'cfg1 geometry it=%d' % it)
        m2 = o['model2'].copy()
        rms_o = core.update_model(model1, m2, o['inter_weight'], 0.)
        rms_g = eng.update_model(0.)
        assert abs(rms_g - rms_o) <= TOL * abs(rms_o)
        assert P.rel_err(eng.get('model1'), model1) <= TOL
    eng.close()


@pytest.mark.parametrize('mode', ['2', '2-seq16', '2-box', '1', '0'])
def test_orientation_chunks_and_tomogram_forms(tiny, monkeypatch, mode):
    """Orientations split into 64-wide chunks (seven chunks for 420 orientations), with each
    tomogram form (frame-partitioned, frame-partitioned with 16 ranges walked in sequence per
    XCD and 16-deep gathers, frame-partitioned with the voxel-box merge instead of the brick
    merge, pixel-tile, per-pixel gather)."""
    monkeypatch.setenv('EMC_MAX_CHUNK_COLS', '64')
    if mode == '2-box':
        monkeypatch.setenv('EMC_MERGE_MODE', '0')
        mode = '2'
    if mode == '2-seq16':
        for k, v in (('EMC_TOMO_SEQ', '1'), ('EMC_TOMO_PARTS', '16'), ('EMC_TOMO_UNROLL', '16'),
                     ('EMC_LIK_UNROLL', '16')):
            monkeypatch.setenv(k, v)
        mode = '2'
    monkeypatch.setenv('EMC_TOMO_TILE_MODE', mode)
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(tiny, flags)
    eng = P.engine_for(tiny)
    g = P.run_engine(eng, flags)
    eng.close()
    assert g['num_chunks'] == 7
    _compare(o, g, 'chunked mode=' + mode)


@pytest.mark.parametrize('zero_block', [False, True])
def test_dense_integer_frames(tiny, zero_block):
    """Dense integer frames through emc_set_frames_dense against the oracle's dense branch."""
    p = P.densify(tiny)
    if zero_block:
        m = tiny['model1'].copy()
        c = m.shape[0] // 2
        m[c + 3:c + 6, c - 2:c + 3, :] = 0.
        p['model1'] = m
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(p, flags)
    eng = P.engine_for_dense(p)
    g = P.run_engine(eng, flags)
    eng.close()
    _compare(o, g, 'dense zero_block=%s' % zero_block)


# kernel-variant options of the subset path's M-step (start values from EMC_<NAME>)
SUBSET_VARIANTS = {
    'no-buckets': {'EMC_TOMO_BUCKETS': '0'},                # pair tomogram walks its tiles' pairs
    'full-parts': {'EMC_SIG_FULL_ROWS': '0'},               # full frames through partial views
    'full-direct': {'EMC_SIG_FULL_ROWS': str(1 << 20)},     # full frames' rows read directly
    'full-parts-no-buckets': {'EMC_SIG_FULL_ROWS': '0', 'EMC_TOMO_BUCKETS': '0'},
    'small-units': {'EMC_UNIT_ITEMS': '64'},                # merge units of 64 list entries
}


@pytest.mark.parametrize('form', ['subset', 'subset-chunked', 'subset-padded', 'masked'] +
                         ['subset-' + k for k in SUBSET_VARIANTS])
def test_refinement_subsets(tiny, monkeypatch, form):
    """emc_set_refine: per-frame coarse cells (Voronoi of quat_gen(1)), one frame with an
    empty set and every fourth frame with every cell (a full frame: the fall-back to the plain
    search, with repeated cells in its list), against the oracle's refinement branch;
    evaluated sparsely over (frame, tile) pairs (in one or in 64-slot chunks; cells packed back
    to back, or each padded to whole tiles), or over every pair with the excluded ones masked;
    and with each kernel variant of the sparse M-step (SUBSET_VARIANTS)."""
    if form == 'subset-chunked':
        monkeypatch.setenv('EMC_MAX_CHUNK_COLS', '64')
    if form == 'subset-padded':
        monkeypatch.setenv('EMC_SLOT_PACK', '0')
    if form == 'masked':
        monkeypatch.setenv('EMC_SUBSET', '0')
    for k, v in SUBSET_VARIANTS.get(form[len('subset-'):], {}).items():
        monkeypatch.setenv(k, v)
    from oracle import core
    from dragonfly_amd import _lib
    qc = _lib.quat_gen(1)
    qmap = _lib.voronoi_subset(qc, tiny['quats'])
    rng = np.random.default_rng(11)
    D = len(tiny['fparams']['beta'])
    sets = [np.sort(rng.choice(len(qc), 10, replace=False)).astype('i4') for _ in range(D)]
    for d in range(1, D, 4):
        sets[d] = np.concatenate([rng.permutation(len(qc)), [3, 3]]).astype('i4')
    sets[2] = np.zeros(0, 'i4')
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    ref = dict(quat_mapping=qmap, rel_off=off, rel_quat=np.concatenate(sets))
    flags = dict(iteration=1, need_scaling=0)
    fp = dict(tiny['fparams'], scale=np.ones(D))
    o = core.maximize(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'], fp, flags,
                      dense_prob=True, refine=ref)
    o['scale'] = fp['scale']
    eng = P.engine_for(tiny)
    eng.set_refine(len(qc), qmap, off, ref['rel_quat'])
    g = P.run_engine(eng, flags)
    # the saved sparse posterior (one batch, then batches of 7 frames) is the cut of the dense one
    sparse = [eng.sparse_prob()]
    eng.set_option('sparse_rows', 7)
    sparse.append(eng.sparse_prob())
    eng.close()
    _compare(o, g, 'refinement ' + form)
    assert g['rmax'][2] == -1
    keep = g['prob'] > 1e-6
    for num, place, prob in sparse:
        assert np.array_equal(num, keep.sum(1))
        assert np.array_equal(place, np.nonzero(keep)[1])
        assert np.array_equal(prob, g['prob'][keep])
    if form == 'subset-chunked':
        assert g['num_chunks'] > 1


def test_refinement_subsets_scaled_blacklisted(tiny_scaled):
    """Sparse subsets with scale factors (iterations 1 and 2) and blacklisted frames; every
    third frame is a full frame (frame 7, blacklisted, among them)."""
    from oracle import core
    from dragonfly_amd import _lib
    p = tiny_scaled
    qc = _lib.quat_gen(1)
    qmap = _lib.voronoi_subset(qc, p['quats'])
    rng = np.random.default_rng(12)
    D = len(p['fparams']['beta'])
    sets = [np.sort(rng.choice(len(qc), 20, replace=False)).astype('i4') for _ in range(D)]
    for d in range(1, D, 3):
        sets[d] = np.arange(len(qc), dtype='i4')
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    ref = dict(quat_mapping=qmap, rel_off=off, rel_quat=np.concatenate(sets))
    bl = np.zeros(D, 'u1')
    bl[[1, 7]] = 1
    fp = dict(p['fparams'], blacklist=bl, num_blacklist=2, scale=np.asarray(p['true_scale'], 'f8').copy())
    eng = P.engine_for(dict(p, fparams=fp))
    eng.set_refine(len(qc), qmap, off, ref['rel_quat'])
    for it in (1, 2):
        flags = dict(iteration=it, need_scaling=1, update_scale=1)
        fo = dict(fp, scale=fp['scale'].copy())
        o = core.maximize(p['det'], p['quats'], p['model1'], p['frames'], fo, flags, dense_prob=True,
                          refine=ref)
        o['scale'] = fo['scale']
        eng.set_scale(fp['scale'])
        g = P.run_engine(eng, flags)
        _compare(o, g, 'refinement scaled it=%d' % it, scaling=True)
    eng.close()


@pytest.mark.parametrize('cells', ['all', 'random'])
def test_refinement_subsets_at_scale_vs_engine(cells):
    """151x151 detector, 125^3 model, fine num_div 4 (3240 orientations), coarse num_div 2
    (420 cells), 1500 frames: the sparse subset path against the plain engine with every cell
    (identical to the full search), and against the masked form with 24 random cells per frame;
    several slot chunks (1024 slots each)."""
    import os
    from dragonfly_amd import _lib, synth
    cfg = dict(synth.CONFIGS['cfg1'], num_data=1500, mean_count=300.)
    p = P.build(cfg, device='cuda', num_ave=1000)
    qc = _lib.quat_gen(2)
    qmap = _lib.voronoi_subset(qc, p['quats'])
    D = len(p['fparams']['beta'])
    if cells == 'all':
        sets = [np.arange(len(qc), dtype='i4')] * D
    else:
        rng = np.random.default_rng(3)
        sets = [np.sort(rng.choice(len(qc), 24, replace=False)).astype('i4') for _ in range(D)]
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    rq = np.concatenate(sets)
    flags = dict(iteration=1, need_scaling=0)
    ref = P.engine_for(p)
    if cells == 'random':
        os.environ['EMC_SUBSET'] = '0'
        try:
            ref.close()
            ref = P.engine_for(p)
        finally:
            del os.environ['EMC_SUBSET']
        ref.set_refine(len(qc), qmap, off, rq)
    o = P.run_engine(ref, flags)
    ref.close()
    os.environ['EMC_MAX_CHUNK_COLS'] = '1024'
    try:
        eng = P.engine_for(p)
    finally:
        del os.environ['EMC_MAX_CHUNK_COLS']
    eng.set_refine(len(qc), qmap, off, rq)
    g = P.run_engine(eng, flags)
    eng.close()
    assert g['num_chunks'] > 1
    _compare(o, g, 'refinement at scale, %s cells' % cells)


def test_refinement_subsets_past_32bit_grid():
    """More (frame, tile) pairs than a 32-bit count of work-items holds: 17000 frames on a
    22x20 detector, fine num_div 16 (204960 orientations); 16000 frames take all but one of
    the 3240 coarse cells (4571 slot tiles), 7.3e7 pairs or 4.7e9 lanes of the pair
    likelihood, and 1000 frames take every cell (full frames). The subset path must agree
    with the plain engine for every frame (a wrapped launch grid drops the last pairs)."""
    from dragonfly_amd import _lib
    cfg = P.tiny_config(num_data=17000, mean_count=40., num_div=16)
    p = P.build(cfg, device='cuda')
    qc = _lib.quat_gen(4)
    qmap = _lib.voronoi_subset(qc, p['quats'])
    D = len(p['fparams']['beta'])
    every = np.arange(len(qc), dtype='i4')
    sets = [every if d % 17 == 5 else np.delete(every, (7 * d) % len(qc)) for d in range(D)]
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    rq = np.concatenate(sets)
    keys = ('u', 'max_exp', 'p_norm', 'rmax', 'likelihood', 'info', 'psum_r', 'model2',
            'inter_weight')
    out = []
    import os
    for sparse in (False, True):
        # reference: the same subsets evaluated over every pair and masked
        if not sparse:
            os.environ['EMC_SUBSET'] = '0'
        try:
            eng = P.engine_for(p)
        finally:
            os.environ.pop('EMC_SUBSET', None)
        eng.set_refine(len(qc), qmap, off, rq)
        eng.maximize(1, 0, 1, 0)
        out.append({k: eng.get(k) for k in keys})
        eng.close()
    o, g = out
    errs = {k: P.rel_err(g[k], o[k]) for k in keys if k != 'rmax'}
    print('refinement past 2^32 lanes', {k: '%.1e' % v for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, bad
    assert np.array_equal(g['rmax'], o['rmax'])


def test_cell_posterior(tiny):
    """emc_cell_posterior: per-frame posterior mass of every coarse cell, from the plain
    path's rows and from the subset rows (partial and full frames), against the dense
    posterior summed per cell (float32 output)."""
    from dragonfly_amd import _lib
    qc = _lib.quat_gen(1)
    qmap = _lib.voronoi_subset(qc, tiny['quats'])
    D = len(tiny['fparams']['beta'])
    rng = np.random.default_rng(5)
    sets = [np.arange(len(qc), dtype='i4') if d % 3 == 0 else
            np.sort(rng.choice(len(qc), 12, replace=False)).astype('i4') for d in range(D)]
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    for refine in (False, True):
        eng = P.engine_for(tiny)
        if refine:
            eng.set_refine(len(qc), qmap, off, np.concatenate(sets))
        g = P.run_engine(eng, dict(iteration=1, need_scaling=0))
        cp = eng.cell_posterior(len(qc), qmap)
        eng.close()
        want = np.zeros((D, len(qc)))
        for r in range(len(qmap)):
            want[:, qmap[r]] += g['prob'][:, r]
        assert cp.shape == (D, len(qc))
        assert np.abs(cp - want).max() <= 1e-6, refine
        ok = g['rmax'] >= 0
        assert np.allclose(cp[ok].sum(1), 1., atol=1e-5)


def test_options_roundtrip_and_variants_agree(tiny):
    """emc_set_option: values read back, unknown names and out-of-range values are refused,
    and variant choices give the same iteration (to the parity tolerance)."""
    from dragonfly_amd import _lib
    eng = P.engine_for(tiny)
    eng.set_option('lik_unroll', 16)
    assert eng.get_option('lik_unroll') == 16
    with pytest.raises(_lib.EMCError):
        eng.set_option('no_such_option', 1)
    with pytest.raises(_lib.EMCError):
        eng.set_option('tomo_parts', 0)
    flags = dict(iteration=1, need_scaling=0)
    g1 = P.run_engine(eng, flags)
    eng.set_option('tomo_parts', 3)
    eng.set_option('lik_tiles', 2)
    eng.set_option('lik_xcd', 0)
    g2 = P.run_engine(eng, flags)
    eng.set_option('lik_split', 4)
    eng.set_option('max_chunk_cols', 128)
    g3 = P.run_engine(eng, flags)
    # XCDs sharing a log W tile in the likelihood's 1-D order (2, 4, 8), chunked
    eng.set_option('lik_split', 1)
    gx = []
    for v, tiles in ((2, 1), (4, 2), (8, 1)):
        eng.set_option('lik_xcd', v)
        eng.set_option('lik_tiles', tiles)
        gx.append(P.run_engine(eng, flags))
    eng.close()
    for g in [g2, g3] + gx:
        for k in FIELDS_F:
            assert P.rel_err(g[k], g1[k]) <= TOL, k
        assert np.array_equal(g1['rmax'], g['rmax'])
    assert g3['num_chunks'] > 1


def test_sparse_prob(tiny):
    """emc_get_sparse_prob (the saved probabilities, max_prob.c:295-326): counts and entries
    compacted on the GPU equal the cut P > 1e-6 of the engine's dense posterior bit for bit,
    orientations ascending, blacklisted frames empty; against the oracle within 1e-9."""
    D = len(tiny['fparams']['beta'])
    bl = np.zeros(D, 'u1')
    bl[[3, 17]] = 1
    p = dict(tiny, fparams=dict(tiny['fparams'], blacklist=bl, num_blacklist=2))
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(p, flags)
    eng = P.engine_for(p)
    P.run_engine(eng, flags)
    num, place, prob = eng.sparse_prob()
    dense = eng.get('prob')
    eng.close()
    keep = dense > 1e-6
    assert np.array_equal(num, keep.sum(1))
    assert num[3] == 0 and num[17] == 0 and num.sum() > D
    assert np.array_equal(place, np.nonzero(keep)[1])
    assert np.array_equal(prob, dense[keep])
    okeep = o['prob'] > 1e-6
    assert np.array_equal(keep, okeep)
    assert P.rel_err(prob, o['prob'][okeep]) <= TOL


def test_estep_only_and_subset_mass(tiny_scaled):
    """The expectation-step-only iteration gives the full iteration's posterior and
    per-frame statistics and leaves model2, inter_weight and the scale factors alone; the
    per-frame mass over listed orientations is the sum of those posterior entries."""
    p = tiny_scaled
    flags = dict(iteration=2, need_scaling=1, update_scale=1)
    p2 = dict(p, fparams=dict(p['fparams'], scale=np.asarray(p['true_scale'], 'f8').copy()))
    full = P.engine_for(p2)
    g = P.run_engine(full, flags)
    full.close()
    eng = P.engine_for(p2)
    eng.set_option('estep_only', 1)
    e = P.run_engine(eng, flags)
    for k in ('max_exp', 'p_norm', 'likelihood', 'info', 'psum_d', 'prob'):
        assert np.array_equal(e[k], g[k]), k
    assert np.array_equal(e['rmax'], g['rmax'])
    assert e['avg_likelihood'] == g['avg_likelihood'] and e['mutual_info'] == g['mutual_info']
    assert not e['model2'].any() and not e['inter_weight'].any()
    assert np.array_equal(e['scale'], p2['fparams']['scale'])
    D, R = g['prob'].shape
    rng = np.random.default_rng(11)
    sets = [np.sort(rng.choice(R, rng.integers(0, 30), replace=False)).astype('i4') for _ in range(D)]
    off = np.concatenate([[0], np.cumsum([len(x) for x in sets])]).astype('i8')
    mass = eng.subset_mass(off, np.concatenate(sets))
    want = np.array([g['prob'][d, s].sum() for d, s in enumerate(sets)])
    assert np.allclose(mass, want, rtol=1e-14, atol=1e-300)
    with pytest.raises(Exception):
        eng.subset_mass(off, np.concatenate(sets) + R)   # orientation out of range
    eng.close()


def _refine_setup(p, seed=5, ncells=10):
    from dragonfly_amd import _lib
    qc = _lib.quat_gen(1)
    qmap = _lib.voronoi_subset(qc, p['quats'])
    rng = np.random.default_rng(seed)
    D = len(p['fparams']['beta'])
    sets = [np.sort(rng.choice(len(qc), ncells, replace=False)).astype('i4') for _ in range(D)]
    for d in range(3, D, 5):
        sets[d] = np.array([-1], 'i4')  # the full search
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    return qc, qmap, off, np.concatenate(sets)


@pytest.mark.parametrize('g32', [False, True])
def test_estep_then_mstep_resume_equals_one_pass(tiny_scaled, g32):
    """On refinement subsets, an estep_only iteration followed by an mstep_only one gives the
    single full iteration's results (bit for bit: scales, posterior, per-frame and
    per-orientation sums; model2 and weights to the order of their atomic additions);
    mstep_only is refused without an estep_only iteration of the
    current state (none run, or a setter called in between)."""
    p = tiny_scaled
    p = dict(p, fparams=dict(p['fparams'], scale=np.asarray(p['true_scale'], 'f8').copy()))
    qc, qmap, off, cells = _refine_setup(p)
    flags = dict(iteration=2, need_scaling=1, update_scale=1)
    out = []
    for two in (False, True):
        eng = P.engine_for(p)
        if g32:
            eng.set_precision(3)
        eng.set_refine(len(qc), qmap, off, cells)
        if two:
            with pytest.raises(Exception):
                eng.set_option('mstep_only', 1)
                eng.maximize(2, 1, 1, 0)
            eng.set_option('mstep_only', 0)
            eng.set_option('estep_only', 1)
            eng.maximize(2, 1, 1, 0)
            eng.set_option('estep_only', 0)
            eng.set_option('mstep_only', 1)
        out.append(P.run_engine(eng, flags))
        if two:
            # a second resume has nothing left to resume; nor has one after a setter
            with pytest.raises(Exception):
                eng.maximize(2, 1, 1, 0)
            eng.set_option('mstep_only', 0)
            eng.set_option('estep_only', 1)
            eng.maximize(2, 1, 1, 0)
            eng.set_option('estep_only', 0)
            eng.set_scale(p['fparams']['scale'])
            eng.set_option('mstep_only', 1)
            with pytest.raises(Exception):
                eng.maximize(2, 1, 1, 0)
        eng.close()
    a, b = out
    for k in ('u', 'max_exp', 'p_norm', 'rmax', 'likelihood', 'info', 'psum_d', 'psum_r', 'scale', 'prob'):
        assert np.array_equal(a[k], b[k]), k
    for k in ('avg_likelihood', 'mutual_info', 'rescale'):
        assert a[k] == b[k], k
    # the merge adds into the volume with atomics, whose order differs from run to run (and
    # with FP32 gathers sums its bricks in FP32)
    for k in ('model2', 'inter_weight'):
        err = P.rel_err(a[k], b[k])
        print('resume %s: %.1e' % (k, err), flush=True)
        assert err <= (1e-5 if g32 else 1e-13), k


@pytest.mark.parametrize('g32', [False, True])
def test_second_pass_reuses_resident_log_w(tiny_scaled, g32):
    """An estep_only pass over one set of refinement lists, then a whole iteration over another
    set on the same slots (the judged iteration's second pass) skips the expand, log W of every
    slot being still in place (one orientation chunk): its results equal those of an engine
    that runs the second lists alone (bit for bit up to the merge's atomic order), and a
    model change in between makes the expand run again."""
    p = tiny_scaled
    p = dict(p, fparams=dict(p['fparams'], scale=np.asarray(p['true_scale'], 'f8').copy()))
    qc, qmap, off1, cells1 = _refine_setup(p, seed=5)
    _, _, off2, cells2 = _refine_setup(p, seed=9, ncells=12)
    flags = dict(iteration=2, need_scaling=1, update_scale=1)
    out = []
    for two in (False, True):
        eng = P.engine_for(p)
        if g32:
            eng.set_precision(3)
        if two:
            eng.set_refine(len(qc), qmap, off1, cells1)
            eng.set_option('estep_only', 1)
            st1 = eng.maximize(2, 1, 1, 0)
            eng.set_option('estep_only', 0)
            assert st1['stage_ms']['expand'] > 0.
            eng.set_scale(p['fparams']['scale'])
        eng.set_refine(len(qc), qmap, off2, cells2)
        r = P.run_engine(eng, flags)
        out.append(r)
        if two:
            assert r['stage_ms']['expand'] == 0., 'the second pass ran the expand'
            # after the iteration (views written over log W) and a model change: expand again
            eng.set_model(p['model1'])
            eng.set_scale(p['fparams']['scale'])
            eng.set_option('estep_only', 1)
            st3 = eng.maximize(2, 1, 1, 0)
            assert st3['stage_ms']['expand'] > 0.
        else:
            assert r['stage_ms']['expand'] > 0.
        eng.close()
    a, b = out
    for k in ('u', 'max_exp', 'p_norm', 'rmax', 'likelihood', 'info', 'psum_d', 'psum_r', 'scale', 'prob'):
        assert np.array_equal(a[k], b[k]), k
    for k in ('avg_likelihood', 'mutual_info', 'rescale'):
        assert a[k] == b[k], k
    for k in ('model2', 'inter_weight'):
        err = P.rel_err(b[k], a[k])
        print('resident log W %s: %.1e' % (k, err), flush=True)
        assert err <= (1e-5 if g32 else 1e-13), k


def test_copy_model_between_engines(tiny_scaled):
    """emc_copy_model gives the destination engine the source's model1 bit for bit, and an
    iteration from it equals one from the same model set from the host."""
    p = tiny_scaled
    a = P.engine_for(p)
    m = a.get('model1')
    b = P.engine_for(dict(p, model1=np.full_like(p['model1'], 3.)))
    b.copy_model(a)
    assert np.array_equal(b.get('model1'), m)
    ra = P.run_engine(a, dict(iteration=1))
    rb = P.run_engine(b, dict(iteration=1))
    for k in ('u', 'max_exp', 'p_norm', 'likelihood', 'prob'):
        assert np.array_equal(ra[k], rb[k]), k
    a.close()
    b.close()


def test_judge_outside_estimate(tiny_scaled):
    """emc_judge_outside on the tiny problem (fine num_div 2, coarse num_div 1): the estimate
    equals its definition recomputed from the two engines' results (the coarse posterior,
    maxima and normalisers, the fine maxima and normalisers inside the cells, cell sizes and
    weights), full frames get 0; and it follows the true outside mass of a full fine search --
    frames estimated below 1e-3 hold less than 1e-2 outside their cells."""
    from dragonfly_amd import _lib
    p = tiny_scaled
    p = dict(p, fparams=dict(p['fparams'], scale=np.asarray(p['true_scale'], 'f8').copy()))
    qc, qmap, off, cells = _refine_setup(p, seed=9, ncells=24)
    D = len(off) - 1
    beta = p['fparams']['beta']
    full = P.engine_for(p)
    full.maximize(2, 1, 0, 0)
    post = full.cell_posterior(len(qc), qmap).astype('f8')
    full.close()
    coarse = P.engine_for(dict(p, quats=qc))
    coarse.set_option('estep_only', 1)
    coarse.maximize(2, 1, 0, 0)
    Pc = coarse.get('prob')
    Mc, pnc = coarse.get('max_exp'), coarse.get('p_norm')
    fine = P.engine_for(p)
    fine.set_refine(len(qc), qmap, off, cells)
    fine.set_option('estep_only', 1)
    fine.maximize(2, 1, 0, 0)
    Mf, pnf = fine.get('max_exp'), fine.get('p_norm')
    est = fine.judge_outside(coarse)
    tau = 1e-3
    est_w, num_add, added = fine.judge_widen(coarse, tau, 16, target=tau / 4)
    fine.close()
    coarse.close()
    n = np.bincount(qmap, minlength=len(qc)).astype('f8')
    wbar = np.bincount(qmap, weights=p['quats'][:, 4], minlength=len(qc)) / n
    want = np.zeros(D)
    inside = np.zeros((D, len(qc)), bool)
    for d in range(D):
        cs = cells[off[d]:off[d + 1]]
        if len(cs) == 1 and cs[0] == -1:
            inside[d] = True
            continue
        inside[d, cs] = True
        g = n * (wbar / qc[:, 4]) ** beta[d]
        acc = (Pc[d] * g)[~inside[d]].sum()
        if acc > 0:
            zout = pnc[d] * np.exp(beta[d] * (Mc[d] - Mf[d])) * acc
            want[d] = zout / (pnf[d] + zout)
    assert np.allclose(est, want, rtol=1e-10, atol=1e-300), np.abs(est - want).max()
    assert (est[3::5] == 0).all()
    # widening: the same estimate; frames at most tau keep their cells; a rejected frame gets
    # the cells of largest share of the outside sum whose removal leaves at most the remainder
    # allowed by the target tau / 4 (the shortest such prefix of the shares in decreasing
    # order), or -1 past 16
    assert np.array_equal(est_w, est)
    a = 0
    nwide = 0
    for d in range(D):
        k = int(num_add[d])
        if est[d] <= tau:
            assert k == 0, d
            continue
        g = n * (wbar / qc[:, 4]) ** beta[d]
        term = np.where(inside[d], 0., Pc[d] * g)
        allow = (tau / 4) / (1. - tau / 4) * pnf[d] / (pnc[d] * np.exp(beta[d] * (Mc[d] - Mf[d])))
        order = np.argsort(-term, kind='stable')
        rem = term.sum() - np.cumsum(term[order])
        kmin = int(np.argmax(rem <= allow * (1. + 1e-9))) + 1
        if k < 0:
            assert kmin > 16 - 1, (d, kmin)  # more than max_add needed (up to the cut's float rounding)
            continue
        got = added[a:a + k]
        a += k
        nwide += 1
        assert np.all(np.diff(got) > 0) and not inside[d, got].any()
        assert abs(k - kmin) <= 1, (d, k, kmin)
        left = term.sum() - term[got].sum()
        assert left <= allow * (1. + 1e-5), (d, left, allow)
        assert term[got].min() >= np.delete(term, got).max() * (1. - 1e-5)
    assert a == len(added)
    print('judge widen: %d frames widened, %d to the full search' % (nwide, (num_add < 0).sum()), flush=True)
    true_out = np.clip(1. - (post * inside).sum(1), 0., 1.)
    low = est <= 1e-3
    print('judge: frames estimated <= 1e-3: %d of %d; their true outside mass max %.2e; correlation of logs %.3f'
          % (low.sum(), D, true_out[low].max() if low.any() else 0.,
             np.corrcoef(np.log(est[est > 0] + 1e-12), np.log(true_out[est > 0] + 1e-12))[0, 1]), flush=True)
    assert low.any() and (true_out[low] <= 1e-2).all()


@pytest.mark.parametrize('g32', [False, True])
def test_graph_iteration_equals_eager(tiny_scaled, g32):
    """The refinement-subset iteration run as a captured HIP graph (option graph) against the
    same iterations run eagerly on a second engine: three iterations chained through
    update_model and normalize_scale (the third replays the graph of the second, reading the
    updated model and scales, with eager work and synchronisation in between), then new
    refinement lists (the graph is captured again). Per-frame and per-orientation results equal; the merged volumes to the
    order of their atomic additions."""
    p = tiny_scaled
    p = dict(p, fparams=dict(p['fparams'], scale=np.asarray(p['true_scale'], 'f8').copy()))
    qc, qmap, off, cells = _refine_setup(p)
    _, _, off2, cells2 = _refine_setup(p, seed=21, ncells=16)
    engs = [P.engine_for(p), P.engine_for(p)]
    for eng in engs:
        if g32:
            eng.set_precision(3)
        eng.set_refine(len(qc), qmap, off, cells)
    engs[1].set_option('graph', 1)
    tol = 1e-5 if g32 else 1e-13
    for it in (1, 2, 3, 4):
        if it == 4:
            for eng in engs:
                eng.set_refine(len(qc), qmap, off2, cells2)
        flags = dict(iteration=it, need_scaling=1, update_scale=1)
        print('case: graph iteration %d (g32 %s)' % (it, g32), flush=True)
        a, b = (P.run_engine(eng, flags) for eng in engs)
        for k in ('u', 'max_exp', 'p_norm', 'rmax', 'likelihood', 'info', 'psum_d', 'psum_r', 'scale', 'prob'):
            assert np.array_equal(a[k], b[k]), (it, k)
        for k in ('avg_likelihood', 'mutual_info', 'rescale'):
            assert a[k] == b[k], (it, k)
        for k in ('model2', 'inter_weight'):
            assert P.rel_err(a[k], b[k]) <= tol, (it, k, P.rel_err(a[k], b[k]))
        rms = [eng.update_model(0.) for eng in engs]
        assert abs(rms[0] - rms[1]) <= tol * max(rms[0], 1e-300)
        for eng in engs:
            eng.normalize_scale()
        # the same input for the next iteration on both (the merges differ in their last bits)
        engs[1].set_model(engs[0].get('model1'))
        engs[1].set_scale(engs[0].get('scale'))
    # the graph option refuses the two-pass switches
    engs[1].set_option('estep_only', 1)
    with pytest.raises(Exception):
        engs[1].maximize(5, 1, 1, 0)
    for eng in engs:
        eng.close()


@pytest.mark.parametrize('g32, chunked, graph', [(False, False, False), (True, False, False), (False, True, False),
                                                 (True, True, True)])
def test_fused_expand_likelihood_equals_separate(tiny_scaled, monkeypatch, g32, chunked, graph):
    """The persistent expand + pair likelihood kernel (option fused) against the separate
    expand and likelihood launches on the refinement subsets (partial and full frames), in one
    or in 64-slot chunks, eagerly or as a captured graph: the same per-frame and
    per-orientation results bit for bit (the same arithmetic in the same order), the merged
    volumes to the order of their atomic additions."""
    if chunked:
        monkeypatch.setenv('EMC_MAX_CHUNK_COLS', '64')
    p = tiny_scaled
    p = dict(p, fparams=dict(p['fparams'], scale=np.asarray(p['true_scale'], 'f8').copy()))
    qc, qmap, off, cells = _refine_setup(p, seed=3, ncells=12)
    flags = dict(iteration=2, need_scaling=1, update_scale=1)
    out = []
    for fused in (0, 1):
        eng = P.engine_for(p)
        if g32:
            eng.set_precision(3)
        eng.set_refine(len(qc), qmap, off, cells)
        eng.set_option('fused', fused)
        if fused and graph:
            eng.set_option('graph', 1)
        g = P.run_engine(eng, flags)
        if fused and graph:
            # a replay of the captured graph from the same input gives the same again
            eng.set_scale(p['fparams']['scale'])
            g2 = P.run_engine(eng, flags)
            assert np.array_equal(g2['max_exp'], g['max_exp']) and np.array_equal(g2['prob'], g['prob'])
        out.append(g)
        eng.close()
    a, b = out
    assert a['num_chunks'] == b['num_chunks'] and (a['num_chunks'] > 1) == chunked
    for k in ('u', 'max_exp', 'p_norm', 'rmax', 'likelihood', 'info', 'psum_d', 'psum_r', 'scale', 'prob'):
        assert np.array_equal(a[k], b[k]), k
    for k in ('model2', 'inter_weight'):
        assert P.rel_err(a[k], b[k]) <= (1e-5 if g32 else 1e-13), k
