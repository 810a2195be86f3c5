"""The HIP path at the geometries of the benchmark configurations (251^3 model, 260x257
detector as two panels with a 10-pixel gap) rather than the tiny test problems:

* plain FP64 at num_div 10 (50,100 orientations) with scale optimisation against the oracle
  on a frame subset at every orientation, two iterations chained through the model update
  (1e-9 relative, rmax exact), and the full 100,000-frame configuration against
  size-independent identities plus the oracle's per-frame results on its first frames;
* dense frames of ~5000 photons with the FP16 matrix-core products against the FP64 path;
* per-frame orientation subsets at num_div 12 (uneven photon counts) against the masked form;
* proposals judged by a coarse full search at fine num_div 16 against the plain full search.

The oracle runs on up to 16 host threads here (a full orientation set on one thread would
take minutes); its per-thread keep lists then equal the one-thread rule because every frame
keeps every orientation at these betas, which the posterior comparison checks.
"""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-9
FIELDS = ('u', 'max_exp', 'p_norm', 'likelihood', 'info', 'psum_d', 'model2', 'inter_weight', 'prob')


def _errs(g, o, keys):
    out = {k: P.rel_err(g[k], o[k]) for k in keys}
    for k in ('rescale', 'avg_likelihood', 'mutual_info'):
        if k in g and k in o:
            out[k] = abs(g[k] - o[k]) / max(abs(o[k]), 1e-300)
    return out


def _check(label, errs, tol=TOL):
    print(label, {k: '%.1e' % v for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, '%s: relative errors above %g: %s' % (label, tol, bad)


@pytest.fixture(scope='module')
def cfg2_full():
    from dragonfly_amd import synth
    print('case: cfg2 workload, 100000 frames', flush=True)
    return synth.make_problem('cfg2', device='cuda', batch=2048)


def test_cfg2_geometry_subset_vs_oracle(cfg2_full):
    """128 frames of the cfg2 generator at all 50,100 orientations, need_scaling on:
    iteration 1 (rescale) and 2 (scale factors), through update_model and normalize_scale."""
    from dragonfly_amd import synth
    from oracle import core
    prob = dict(cfg2_full, frames=synth.take_frames(cfg2_full['frames'], 128))
    p = P.from_synth(prob)
    assert len(p['quats']) == 50100 and len(p['det']['corr']) == 260 * 257
    eng = P.engine_for(p)
    model1 = p['model1'].copy()
    scale = np.ones(128)
    for it in (1, 2):
        flags = dict(iteration=it, need_scaling=1, update_scale=1)
        print('case: cfg2 subset iteration %d' % it, flush=True)
        q = dict(p, fparams=dict(p['fparams'], scale=scale.copy()))
        o = P.run_oracle(q, flags, model1=model1, num_threads=P.oracle_threads())
        g = P.run_engine(eng, flags)
        assert (o['prob'] > 0).sum() == (g['prob'] > 0).sum()
        errs = _errs(g, o, FIELDS)
        errs['scale'] = P.rel_err(g['scale'], o['scale'])
        _check('cfg2 subset it=%d' % it, errs)
        assert np.array_equal(g['rmax'], o['rmax'])
        m2 = o['model2'].copy()
        rms_o = core.update_model(model1, m2, o['inter_weight'], 0.)
        rms_g = eng.update_model(0.)
        scale = o['scale'].copy()
        mean = scale.mean()
        model1 *= mean
        scale /= mean
        assert abs(eng.normalize_scale() - mean) <= 1e-14 * mean
        assert abs(rms_g - rms_o) <= TOL * rms_o
        assert P.rel_err(eng.get('model1'), model1) <= TOL
    eng.close()


def test_cfg2_full_size_identities(cfg2_full):
    """All 100,000 frames x 50,100 orientations, iteration 1 with scale optimisation:
    sum_r P_dr = 1, sum_r w_r u_r rescale = mean_count (max_prob.c:40-59), the new scale
    times psum_d = fcounts (max_scale.c:15), p_norm >= 1 with the arg-max term, and the
    per-frame results of the first 48 frames equal to the oracle's on those frames (which
    depend on the rest only through mean_count)."""
    p = P.from_synth(cfg2_full)
    D = len(p['frames']['ones'])
    eng = P.engine_for(p)
    flags = dict(iteration=1, need_scaling=1, update_scale=1)
    st = eng.maximize(1, 1, 1, 0)
    mass = eng.cell_posterior(1, np.zeros(eng.num_rot, 'i4'))[:, 0]
    assert np.abs(mass - 1.).max() < 1e-6
    u = eng.get('u')
    w = p['quats'][:, 4]
    mc = p['fparams']['mean_count']
    assert abs((w * u).sum() * st['rescale'] - mc) <= 1e-12 * mc
    scale, psum_d, fc = eng.get('scale'), eng.get('psum_d'), p['fparams']['fcounts']
    assert np.allclose(scale * psum_d, fc, rtol=1e-12, atol=0)
    pn, rmax = eng.get('p_norm'), eng.get('rmax')
    assert (pn >= 1.).all() and ((rmax >= 0) & (rmax < eng.num_rot)).all()
    g = {k: eng.get(k)[:48] for k in ('max_exp', 'p_norm', 'likelihood', 'info', 'psum_d', 'rmax')}
    eng.close()
    from dragonfly_amd import synth
    sub = dict(p, frames=synth.take_frames(p['frames'], 48),
               fparams={k: (v[:48] if isinstance(v, np.ndarray) else v) for k, v in p['fparams'].items()})
    print('case: cfg2 full size, oracle on 48 frames', flush=True)
    o = P.run_oracle(sub, flags, dense_prob=False, num_threads=P.oracle_threads())
    _check('cfg2 full size, first 48 frames', _errs(g, o, ('max_exp', 'p_norm', 'likelihood', 'info', 'psum_d')))
    assert np.array_equal(g['rmax'], o['rmax'])


@pytest.mark.parametrize('mode, bar', [(1, 1e-4), (2, 1e-5), (4, 1e-2), (5, 1e-3)])
def test_cfg3_geometry_mixed_vs_fp64(mode, bar):
    """Dense frames of ~5000 photons (ftype DENSE_INT, 1024 of the cfg3 generator's 20,000) at
    every num_div 10 orientation: the 16-bit products (mode 1: FP16 log W rounded once; mode 2:
    FP16 hi + lo pair; 4, 5: the same in BF16) against the FP64 path, two iterations with alpha = 0.5. The updated model within
    `bar` of the FP64 one relative to its maximum, average log-likelihood within 1e-6 and the
    most likely orientation the same for at least 99.9% of frames (99% for one BF16 term)."""
    from dragonfly_amd import synth
    print('case: cfg3 geometry, mode %d' % mode, flush=True)
    prob = synth.make_problem('cfg3', device='cuda', num_data=1024, batch=512)
    p = P.densify(P.from_synth(prob))
    engs = [P.engine_for_dense(p), P.engine_for_dense(p)]
    engs[1].set_precision(mode)
    out = []
    for eng in engs:
        tr = []
        for it in (1, 2):
            st = eng.maximize(it, 0, 1, 0)
            eng.update_model(0.5)
            tr.append((st['avg_likelihood'], st['mutual_info'], eng.get('rmax'), eng.get('model1')))
        out.append(tr)
        eng.close()
    for it in range(2):
        (l0, i0, r0, m0), (l1, i1, r1, m1) = out[0][it], out[1][it]
        merr = np.abs(m1 - m0).max() / np.abs(m0).max()
        lerr = abs(l1 - l0) / abs(l0)
        print('cfg3 mode %d it=%d: model %.2e, loglik %.2e, info %.2e, rmax agree %.4f'
              % (mode, it + 1, merr, lerr, abs(i1 - i0) / abs(i0), np.mean(r0 == r1)), flush=True)
        assert merr <= bar and lerr <= 1e-6
        # an arg-max may move between orientations whose log-likelihoods tie to within the
        # products' rounding (one BF16 term: 8 mantissa bits, ~1 unit of log-likelihood over
        # 5000 photons)
        assert np.mean(r0 == r1) >= (0.99 if mode == 4 else 0.999)


def test_cfg4_geometry_subsets_vs_masked():
    """Uneven frames (lognormal photon counts, median 150, 20-20,000) of the cfg4 generator with
    per-frame subsets of 50-5000 of the num_div 12 orientations (whole num_div 3 cells around a
    prior): the sparse subset path against the masked form (every pair evaluated, the excluded
    ones set to -inf), scaled, two iterations through the model update."""
    from dragonfly_amd import _lib, synth
    print('case: cfg4 geometry, 3000 frames', flush=True)
    prob = synth.make_problem('cfg4', device='cuda', num_data=3000, batch=1024)
    cfg = prob['cfg']
    p = P.from_synth(prob)
    qc = _lib.quat_gen(cfg['prior_div'])
    off, rq = synth.prior_subsets(prob['frames']['true_quats'], qc, len(p['quats']), cfg['subset_range'],
                                  cfg['prior_error_deg'])
    qmap = _lib.voronoi_subset(qc, p['quats'])
    res = []
    for subset in (1, 0):
        eng = P.engine_for(p)
        eng.set_option('subset', subset)
        eng.set_refine(len(qc), qmap, off, rq)
        tr = []
        for it in (1, 2):
            g = P.run_engine(eng, dict(iteration=it, need_scaling=1, update_scale=1))
            eng.update_model(0.)
            g['model1'] = eng.get('model1')
            tr.append(g)
        res.append(tr)
        eng.close()
    for it in range(2):
        errs = _errs(res[0][it], res[1][it], ('max_exp', 'p_norm', 'likelihood', 'info', 'psum_d', 'model2',
                                              'inter_weight', 'scale', 'model1', 'prob'))
        _check('cfg4 sparse vs masked it=%d' % (it + 1), errs)
        assert np.array_equal(res[0][it]['rmax'], res[1][it]['rmax'])


def test_cfg4_geometry_fused_graph_equals_eager():
    """The cfg4 batching pattern at its geometry (3000 frames of lognormal photon counts, per-
    frame subsets of 50-5000 of the num_div 12 orientations, FP32 gathers): the iteration run
    with the persistent expand + likelihood kernel and captured as a HIP graph against the
    separate launches run eagerly, two iterations through the model update (the second replays
    the graph on the updated model and scales). Per-frame results equal; the merged and updated
    volumes to the FP32 brick sums' order."""
    from dragonfly_amd import _lib, synth
    print('case: cfg4 geometry, fused + graph, 3000 frames', flush=True)
    prob = synth.make_problem('cfg4', device='cuda', num_data=3000, batch=1024)
    cfg = prob['cfg']
    p = P.from_synth(prob)
    qc = _lib.quat_gen(cfg['prior_div'])
    off, rq = synth.prior_subsets(prob['frames']['true_quats'], qc, len(p['quats']), cfg['subset_range'],
                                  cfg['prior_error_deg'])
    qmap = _lib.voronoi_subset(qc, p['quats'])
    engs = []
    for fast in (0, 1):
        eng = P.engine_for(p)
        eng.set_precision(3)
        eng.set_refine(len(qc), qmap, off, rq)
        eng.set_option('fused', fast)
        eng.set_option('graph', fast)
        engs.append(eng)
    for it in (1, 2, 3):
        a, b = (P.run_engine(eng, dict(iteration=it, need_scaling=1, update_scale=1)) for eng in engs)
        for k in ('max_exp', 'p_norm', 'rmax', 'likelihood', 'info', 'psum_d', 'scale', 'u'):
            assert np.array_equal(a[k], b[k]), (it, k)
        errs = {k: P.rel_err(a[k], b[k]) for k in ('model2', 'inter_weight')}
        print('cfg4 fused+graph it=%d' % it, {k: '%.1e' % v for k, v in errs.items()}, flush=True)
        assert max(errs.values()) <= 1e-5
        for eng in engs:
            eng.update_model(0.)
            eng.normalize_scale()
        engs[1].set_model(engs[0].get('model1'))
        engs[1].set_scale(engs[0].get('scale'))
    for eng in engs:
        eng.close()


def test_cfg5_geometry_judged_proposals():
    """20,000 frames of the cfg5 generator at fine num_div 16 (204,960 orientations) with beta
    at 4x beta_start, from the true intensity at photon scale: after three plain iterations,
    iteration 4 on proposals (each frame's 16 most likely cells of iteration 3's posterior joined
    with the 4 most likely of a coarse full search over the 3240 coarse orientations), judged by
    the likelihood (a frame falls back to the full search where the estimated posterior mass
    outside its cells exceeds 1e-3), against the plain full search from the same model and
    scales. FP64 on both sides, so the difference is the proposals' truncation alone: the
    updated model within 1e-3 of the plain one relative to its maximum, frames that fall back
    give the plain arg-max, and the accepted frames' true outside mass (from the plain
    iteration's cell posterior) is small."""
    from dragonfly_amd import _lib, synth, proposal
    nd = 20000
    print('case: cfg5 geometry, %d frames' % nd, flush=True)
    prob = synth.make_problem('cfg5', device='cuda', num_data=nd, batch=2048)
    p = P.from_synth(prob)
    p['fparams']['beta'] = np.minimum(p['fparams']['beta'] * 4., 1.)
    p['model1'] = np.ascontiguousarray(prob['intens'] * prob['frames']['rescale'])
    qc = _lib.quat_gen(prob['cfg']['coarse_div'])
    Nc = len(qc)
    qmap = _lib.voronoi_subset(qc, p['quats'])
    plain = P.engine_for(p)
    for it in (1, 2, 3):
        plain.maximize(it, 1, 1, 0)
        if it == 3:
            top, _ = plain.cell_topk(Nc, qmap, 16)
        plain.update_model(0.)
        plain.normalize_scale()
    model_k, scale_k = plain.get('model1'), plain.get('scale')
    plain.maximize(4, 1, 1, 0)
    post4 = plain.cell_posterior(Nc, qmap).astype('f8')
    plain.update_model(0.)
    model_plain, rmax_plain = plain.get('model1'), plain.get('rmax')
    plain.close()
    judge = P.engine_for(dict(p, quats=qc, model1=model_k))
    judge.set_scale(scale_k)
    judge.set_option('estep_only', 1)
    judge.maximize(4, 1, 0, 0)
    ctop, _ = judge.cell_topk(Nc, np.arange(Nc, dtype='i4'), 4)
    off, cells = proposal.union_cells(top, ctop)
    eng = P.engine_for(dict(p, model1=model_k))
    eng.set_scale(scale_k)
    eng.set_refine(Nc, qmap, off, cells)
    eng.set_option('estep_only', 1)
    eng.maximize(4, 1, 1, 0)
    est = eng.judge_outside(judge)
    _, num_add, added = eng.judge_widen(judge, 1e-3, 256)
    judge.close()
    accept = est <= 1e-3
    rel_off, rel_quat = proposal.refine_lists_csr(off, cells, accept, Nc)
    eng.set_option('estep_only', 0)
    eng.set_refine(Nc, qmap, rel_off, rel_quat)
    eng.maximize(4, 1, 1, 0)
    eng.update_model(0.)
    model_l, rmax_l = eng.get('model1'), eng.get('rmax')
    eng.close()
    inside = np.zeros(post4.shape, bool)
    fr = np.repeat(np.arange(len(off) - 1), np.diff(off))
    inside[fr, cells] = True
    true_out = np.clip(1. - (post4 * inside).sum(1), 0., 1.)
    err = np.abs(model_l - model_plain).max() / np.abs(model_plain).max()
    rms = np.sqrt(np.mean((model_l - model_plain) ** 2) / np.mean(model_plain ** 2))
    print('cfg5 geometry: accepted %.4f, true outside of accepted mean %.2e / max %.2e, model max-abs %.2e, '
          'rms %.2e, rmax agree %.4f' % (accept.mean(), true_out[accept].mean(), true_out[accept].max(), err, rms,
                                         np.mean(rmax_l == rmax_plain)), flush=True)
    assert np.array_equal(rmax_l[~accept], rmax_plain[~accept])
    assert accept.mean() > 0.5
    assert true_out[accept].mean() <= 1e-3 and np.quantile(true_out[accept], 0.99) <= 1e-2
    assert err <= 1e-3
    # the judge widening the rejected frames' cells (at most 256 more) instead of the full search
    assert np.array_equal(num_add == 0, accept)
    wide = num_add > 0
    w_off, w_cells = proposal.widen_lists_csr(off, cells, num_add, added)
    eng = P.engine_for(dict(p, model1=model_k))
    eng.set_scale(scale_k)
    eng.set_refine(Nc, qmap, w_off, w_cells)
    eng.maximize(4, 1, 1, 0)
    eng.update_model(0.)
    model_w, rmax_w = eng.get('model1'), eng.get('rmax')
    eng.close()
    inside_w = np.zeros(post4.shape, bool)
    fr = np.repeat(np.arange(len(w_off) - 1), np.diff(w_off))
    sel = w_cells >= 0
    inside_w[fr[sel], w_cells[sel]] = True
    inside_w[num_add < 0] = True
    true_w = np.clip(1. - (post4 * inside_w).sum(1), 0., 1.)
    err_w = np.abs(model_w - model_plain).max() / np.abs(model_plain).max()
    print('cfg5 geometry, widened: %.4f of the frames widened, %.4f full; true outside of widened mean %.2e / '
          'q99 %.2e; model max-abs %.2e, rmax agree %.4f' % (wide.mean(), (num_add < 0).mean(),
                                                              true_w[wide].mean() if wide.any() else 0.,
                                                              np.quantile(true_w[wide], 0.99) if wide.any() else 0.,
                                                              err_w, np.mean(rmax_w == rmax_plain)), flush=True)
    assert wide.any()
    assert true_w[wide].mean() <= 1e-3 and np.quantile(true_w[wide], 0.99) <= 1e-2
    assert err_w <= 1e-3
