"""The oracle's full iteration (orc_maximize, restating dragonfly/src/maximize.c:8-56 with
max_prob.c, max_update.c, max_scale.c) pinned without the reference's MPI/GSL build:

* against oracle/brute.py, an independent dense formulation (no keep lists, no C), on tiny
  inputs, unscaled and scaled, first and later iterations, with blacklisted frames;
* against closed forms: a single orientation, the posterior summing to one, the rescale
  identity sum_r w_r u_r rescale = mean_count, the scale update psum_d * scale = fcounts;
* planted mistakes: the brute-force comparison must fail when the oracle's inputs emulate a
  dropped term or a wrong mask threshold, so it is sensitive to those errors.
"""
import numpy as np
import pytest

import problems as P
from oracle import brute

TOL = 1e-11
KEYS = ('u', 'max_exp', 'p_norm', 'likelihood', 'info', 'model2', 'inter_weight', 'prob')


@pytest.fixture(scope='module')
def tiny():
    return P.build(P.tiny_config(num_data=50))


@pytest.fixture(scope='module')
def tiny_scaled():
    return P.build(P.tiny_config(num_data=50, phi=(0.5, 1.5), gap=2))


def _diff(o, b):
    errs = {k: P.rel_err(o[k], b[k]) for k in KEYS}
    for k in ('rescale', 'avg_likelihood', 'mutual_info'):
        errs[k] = abs(o[k] - b[k]) / max(abs(b[k]), 1e-300)
    return errs


def _check(o, b):
    errs = _diff(o, b)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, bad
    assert np.array_equal(o['rmax'], b['rmax'])


def test_unscaled_vs_brute(tiny):
    flags = dict(iteration=1, need_scaling=0)
    _check(P.run_oracle(tiny, flags), brute.iterate(tiny['det'], tiny['quats'], tiny['model1'],
                                                    tiny['frames'], tiny['fparams'], flags))


@pytest.mark.parametrize('iteration', [1, 2])
def test_scaled_vs_brute(tiny_scaled, iteration):
    p = tiny_scaled
    fp = dict(p['fparams'], scale=np.asarray(p['true_scale'], 'f8').copy())
    flags = dict(iteration=iteration, need_scaling=1, update_scale=1)
    b = brute.iterate(p['det'], p['quats'], p['model1'], p['frames'], fp, flags)
    o = P.run_oracle(dict(p, fparams=fp), flags)
    _check(o, b)
    assert P.rel_err(o['psum_d'], b['psum_d']) <= TOL
    assert P.rel_err(o['scale'], b['scale']) <= TOL


def test_blacklist_and_beta_vs_brute(tiny):
    D = len(tiny['frames']['ones'])
    bl = np.zeros(D, 'u1')
    bl[[0, 5, 9]] = 1
    fp = dict(tiny['fparams'], blacklist=bl, num_blacklist=3, beta=np.linspace(0.1, 1., D))
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(dict(tiny, fparams=fp), flags)
    _check(o, brute.iterate(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'], fp, flags))
    assert (o['rmax'][[0, 5, 9]] == -1).all()
    assert (o['prob'][[0, 5, 9]] == 0).all()


def test_blacklist_denominator_is_every_frame(tiny):
    """The reference's Cython path never sets iter->num_blacklist, so combine_information
    divides the summed likelihood and information of the frames that are not blacklisted
    by tot_num_data (max_prob.c:373-374). Both oracles follow the field; with it at 0 the
    averages are the per-frame sums over N_d, not over the frames kept."""
    D = len(tiny['frames']['ones'])
    bl = (np.arange(D) % 2).astype('u1')
    fp = dict(tiny['fparams'], blacklist=bl, num_blacklist=0)
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(dict(tiny, fparams=fp), flags)
    _check(o, brute.iterate(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'], fp, flags))
    keep = bl == 0
    assert abs(o['avg_likelihood'] - o['likelihood'][keep].sum() / D) <= 1e-13 * abs(o['avg_likelihood'])
    assert abs(o['mutual_info'] - o['info'][keep].sum() / D) <= 1e-13 * abs(o['mutual_info'])
    # with the field set to the true count the average is over the kept frames instead
    o2 = P.run_oracle(dict(tiny, fparams=dict(fp, num_blacklist=int(bl.sum()))), flags)
    assert abs(o2['avg_likelihood'] * keep.sum() - o['avg_likelihood'] * D) <= 1e-12 * abs(o['avg_likelihood'] * D)


@pytest.mark.parametrize('size', [9, 10])
@pytest.mark.parametrize('alpha', [0., 0.5])
def test_update_model_vs_brute(size, alpha):
    """orc_update_model (recon.pyx:112-149 with symmetrize_friedel, model.c:470-490) against
    the vectorised numpy restatement: Friedel average of the weighted sums before the
    divide, zero weights, rms change and the alpha blend, odd and even grids."""
    from oracle import core
    rng = np.random.default_rng(size)
    m1 = rng.random((size,) * 3)
    m2 = rng.random((size,) * 3) * 3.
    w = rng.random((size,) * 3) * 2.
    w[rng.random(w.shape) < 0.2] = 0.
    w[0, 0, 0] = w[-1, -1, -1] = 0.  # a Friedel pair with no weight on either side
    b1, b2, brms = brute.update_model(m1.copy(), m2.copy(), w, alpha)
    o1, o2 = m1.copy(), m2.copy()
    rms = core.update_model(o1, o2, w.copy(), alpha)
    assert P.rel_err(o2, b2) <= 1e-14 and P.rel_err(o1, b1) <= 1e-14
    assert abs(rms - brms) <= 1e-14 * brms
    # a planted mistake in the oracle's inputs (dividing before the Friedel average, the
    # legacy order) is detected
    d1, d2 = m1.copy(), np.where(w > 0, m2 / np.where(w > 0, w, 1.), m2)
    core.update_model(d1, d2, (w > 0).astype('f8'), alpha)
    assert P.rel_err(d2, b2) > 1e-6


def test_single_orientation_closed_form(tiny):
    """One orientation of weight 1: every frame has P = 1, info = 0 and
    likelihood = R - sum_fact with R = -u rescale + sum_t K log W."""
    from oracle import core
    q = np.array([[1., 0., 0., 0., 1.]])
    p = dict(tiny, quats=q)
    o = P.run_oracle(p, dict(iteration=1, need_scaling=0))
    det = tiny['det']
    W = core.slice_gen3d(q[0], tiny['model1'], det['qvals'], det['corr'])
    good = det['mask'] < 1
    u = W[good].sum()
    rescale = tiny['fparams']['mean_count'] / u
    K = brute.dense_counts(tiny['frames'], len(W))
    R = -u * rescale + (K[:, good] * np.log(W[good])).sum(1)
    assert np.allclose(o['prob'][:, 0], 1., rtol=1e-14)
    assert np.allclose(o['info'], 0., atol=1e-15)
    assert np.allclose(o['likelihood'], R - tiny['fparams']['sum_fact'], rtol=1e-12)
    assert abs(o['rescale'] - rescale) <= 1e-14 * rescale
    # tomogram of P = 1 everywhere is the photon sum over frames / number of frames,
    # merged once with inter_weight of the one orientation
    S = tiny['model1'].shape[0]
    m2 = np.zeros((S,) * 3)
    w2 = np.zeros((S,) * 3)
    view = np.where(det['mask'] < 2, K.sum(0), 0) / K.shape[0]
    core.slice_merge3d(q[0], view, m2, w2, det['qvals'], det['corr'], det['mask'])
    assert P.rel_err(o['model2'], m2) < 1e-13 and P.rel_err(o['inter_weight'], w2) < 1e-13


def test_invariants(tiny_scaled):
    p = tiny_scaled
    fp = dict(p['fparams'], scale=np.ones(len(p['frames']['ones'])))
    o = P.run_oracle(dict(p, fparams=fp), dict(iteration=1, need_scaling=1, update_scale=1))
    assert np.allclose(o['prob'].sum(1), 1., rtol=1e-13)
    w = p['quats'][:, 4]
    assert abs((w * o['u']).sum() * o['rescale'] - fp['mean_count']) < 1e-12 * fp['mean_count']
    assert np.allclose(o['scale'] * o['psum_d'], fp['fcounts'], rtol=1e-13)
    # the kept set is exactly {r : R > M - 14/beta}: P of dropped orientations is 0 and the
    # smallest kept P is above exp(-14) / p_norm
    kept = o['prob'] > 0
    assert (o['prob'][kept].reshape(-1) >= 0).all()
    pmin = np.where(kept, o['prob'], np.inf).min(1)
    assert (pmin * o['p_norm'] >= np.exp(-14.) * (1 - 1e-9)).all()


@pytest.mark.parametrize('mistake', ['dropped_sum_fact', 'mask1_in_likelihood', 'no_log_weight',
                                     'scale_instead_of_rescale'])
def test_planted_mistakes_are_detected(tiny, mistake):
    """Emulate a plausible oracle mistake through its inputs and check that the
    brute-force comparison flags it."""
    flags = dict(iteration=1, need_scaling=0)
    p = dict(tiny)
    fp = dict(tiny['fparams'])
    det = dict(tiny['det'])
    if mistake == 'dropped_sum_fact':
        fp['sum_fact'] = np.zeros_like(fp['sum_fact'])
    elif mistake == 'mask1_in_likelihood':
        m = det['mask'].copy()
        m[m == 1] = 0
        det['mask'] = m
    elif mistake == 'no_log_weight':
        q = tiny['quats'].copy()
        q[:, 4] = 1.
        p['quats'] = q
    elif mistake == 'scale_instead_of_rescale':
        flags = dict(iteration=2, need_scaling=1, update_scale=0)
        fp['scale'] = np.ones(len(fp['scale']))
    o = P.run_oracle(dict(p, fparams=fp, det=det), flags)
    b = brute.iterate(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'], tiny['fparams'],
                      dict(iteration=1, need_scaling=0))
    errs = _diff(o, b)
    assert max(errs.values()) > 1e-6, (mistake, errs)


@pytest.mark.parametrize('zero_block', [False, True])
def test_dense_frames_vs_brute(tiny, zero_block):
    """Dense integer frames (ftype DENSE_INT): with a zeroed model block, orientations whose
    slice has log W = -inf on a likelihood pixel drop out of every dense frame."""
    p = P.densify(tiny)
    if zero_block:
        m = tiny['model1'].copy()
        c = m.shape[0] // 2
        m[c + 3:c + 6, c - 2:c + 3, :] = 0.
        p['model1'] = m
    flags = dict(iteration=1, need_scaling=0)
    o = P.run_oracle(p, flags)
    b = brute.iterate(p['det'], p['quats'], p['model1'], p['frames'], p['fparams'], flags)
    _check(o, b)
    if zero_block:
        sparse = P.run_oracle(dict(tiny, model1=p['model1']), flags)
        assert (o['prob'] == 0).sum() > (sparse['prob'] == 0).sum()


def _refine_sets(p, coarse_div, per_frame, seed):
    """Voronoi map of the problem's orientations onto quat_gen(coarse_div) and random
    per-frame coarse sets of `per_frame` cells."""
    from oracle import core
    qc = core.quat_gen(coarse_div)
    qmap = core.voronoi_subset(qc, p['quats'])
    rng = np.random.default_rng(seed)
    D = len(p['fparams']['beta'])
    sets = [np.sort(rng.choice(len(qc), per_frame, replace=False)).astype('i4') for _ in range(D)]
    sets[3] = np.zeros(0, 'i4')          # a frame with no orientation at all
    off = np.zeros(D + 1, 'i8')
    np.cumsum([len(x) for x in sets], out=off[1:])
    return dict(quat_mapping=qmap, rel_off=off, rel_quat=np.concatenate(sets)), len(qc)


def test_refinement_vs_brute(tiny):
    """Refinement (legacy src/max_emc.c:264-274): frames evaluated only on their coarse cells'
    orientations; oracle vs brute force, and the frame with an empty set gets no orientation."""
    ref, _ = _refine_sets(tiny, 1, 12, 5)
    flags = dict(iteration=1, need_scaling=0)
    from oracle import core
    o = core.maximize(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'],
                      dict(tiny['fparams'], scale=np.ones(len(tiny['fparams']['beta']))), flags,
                      dense_prob=True, refine=ref)
    b = brute.iterate(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'], tiny['fparams'],
                      flags, refine=ref)
    _check(o, b)
    assert o['rmax'][3] == -1 and (o['prob'][3] == 0).all()
    allowed = np.isin(ref['quat_mapping'], ref['rel_quat'][ref['rel_off'][0]:ref['rel_off'][1]])
    assert (o['prob'][0][~allowed] == 0).all() and o['prob'][0][allowed].sum() > 0.999999


def test_refinement_with_every_cell_is_the_full_search(tiny):
    from oracle import core
    qc = core.quat_gen(1)
    D = len(tiny['fparams']['beta'])
    ref = dict(quat_mapping=core.voronoi_subset(qc, tiny['quats']),
               rel_off=np.arange(D + 1, dtype='i8') * len(qc),
               rel_quat=np.tile(np.arange(len(qc), dtype='i4'), D))
    flags = dict(iteration=1, need_scaling=0)
    fp = dict(tiny['fparams'], scale=np.ones(D))
    a = core.maximize(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'], fp, flags,
                      dense_prob=True, refine=ref)
    b = core.maximize(tiny['det'], tiny['quats'], tiny['model1'], tiny['frames'], dict(fp), flags,
                      dense_prob=True)
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
