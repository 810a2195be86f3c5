"""EMCRecon end to end on the GPU (dragonfly/recon.pyx:48-279 mirrored by
dragonfly_amd/recon.py): one iteration with two model modes and save_prob, its output file
checked against the oracle on the same inputs (1e-9 relative; orientation lists exact)."""
import numpy as np
import pytest

import dragonfly_amd as df
import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-9


def test_recon_two_modes_save_prob(tmp_path):
    from dragonfly_amd import synth
    from oracle import core
    p = P.build(P.tiny_config(num_data=40))
    det = df.CDetector.from_arrays(p['det']['qvals'], p['det']['corr'], p['det']['mask'])
    f = p['frames']
    itr = df.Iterate()
    itr.dets = [det]
    itr.set_data(df.CDataset.from_arrays(det, f['ones'], f['multi'], f['place_ones'], f['place_multi'],
                                         f['count_multi']))
    itr.set_quat(df.Quaternion(2))
    par = itr.params
    par.num_modes, par.save_prob = 2, 1
    par.beta_factor = 1.  # the config default; EMCParams() starts at -1 as the reference's
    par.output_folder, par.log_fname = str(tmp_path), str(tmp_path / 'EMC.log')
    S = p['model1'].shape[0]
    vols = np.stack([p['model1'], synth.start_model(S, p['fparams']['mean_count'], p['det']['mask'], seed=9)])
    model = df.Model(S, 2)
    model.model1 = vols.copy()
    model.model2 = np.zeros_like(vols)
    model.inter_weight = np.zeros_like(vols)
    itr.set_model(model)
    itr.calc_beta(0.8)
    rec = df.EMCRecon()
    rec.set_iterate(itr)
    par.iteration = 1
    print('case: recon two modes save_prob', flush=True)
    rec.run_iteration()
    out = np.load(tmp_path / 'output_001.npz')
    rec.free()

    fp = dict(beta=itr.beta.copy(), scale=np.ones(40), blacklist=np.zeros(40, 'u1'), fcounts=itr.fcounts,
              sum_fact=itr.sum_fact, mean_count=itr.mean_count[0], num_blacklist=0)
    o = core.maximize(p['det'], itr.quat.quats, vols, f, fp, dict(iteration=1, need_scaling=0),
                      dense_prob=True)
    m1, m2 = vols.copy(), o['model2'].copy()
    rms = core.update_model(m1, m2, o['inter_weight'], 0.)
    assert out['intens'].shape == (2, S, S, S)
    assert P.rel_err(out['intens'], m1) <= TOL
    assert P.rel_err(out['inter_weight'], o['inter_weight']) <= TOL
    assert P.rel_err(out['occupancies'], o['quat_norm']) <= TOL
    assert np.array_equal(out['orientations'], o['rmax'])
    assert abs(itr.rms_change - rms) <= TOL * rms
    keep = o['prob'] > 1e-6
    assert np.array_equal(out['probabilities/num_prob'], keep.sum(1))
    assert np.array_equal(out['probabilities/place'], np.nonzero(keep)[1])
    assert P.rel_err(out['probabilities/prob'], o['prob'][keep]) <= TOL
    assert int(out['probabilities/num_rot'][0]) == itr.quat.num_rot


def test_recon_from_config_odd_only(tmp_path):
    """Iterate('config.ini') -> EMCRecon.run_iteration on files written to disk, with
    selection = odd_only: every other frame blacklisted (iterate.pyx:317-359). The reference's
    Cython path never sets iter->num_blacklist, so the logged log-likelihood and mutual
    information are sums over the kept frames divided by all N_d frames
    (max_prob.c:373-374); the oracle is run with the same rule and the log line is checked
    against it."""
    from oracle import core
    p = P.build(P.tiny_config(num_data=50))
    cfg = P.write_config(p, str(tmp_path), selection='odd_only', beta=0.7, num_div=2)
    itr = df.Iterate(cfg)
    assert list(itr.blacklist[:4]) == [0, 1, 0, 1] and itr.num_blacklist == 0
    rec = df.EMCRecon()
    rec.set_iterate(itr)
    itr.params.iteration = 1
    print('case: recon from config, odd_only', flush=True)
    rec.run_iteration()
    out = np.load(tmp_path / 'out' / 'output_001.npz')
    lines = (tmp_path / 'EMC.log').read_text().splitlines()
    rec.free()
    det = itr.dets[0]
    d = dict(qvals=det.qvals, corr=det.corr, mask=det.raw_mask)
    D = itr.tot_num_data
    fp = dict(beta=np.full(D, 0.7), scale=np.ones(D), blacklist=itr.blacklist.copy(), fcounts=itr.fcounts,
              sum_fact=itr.sum_fact, mean_count=itr.mean_count[0], num_blacklist=0)
    o = core.maximize(d, itr.quat.quats, p['model1'], p['frames'], fp, dict(iteration=1, need_scaling=0),
                      dense_prob=True)
    keep = itr.blacklist == 0
    assert abs(o['avg_likelihood'] - o['likelihood'][keep].sum() / D) <= 1e-12 * abs(o['avg_likelihood'])
    assert abs(itr.likelihood - o['avg_likelihood']) <= TOL * abs(o['avg_likelihood'])
    assert abs(itr.mutual_info - o['mutual_info']) <= TOL * abs(o['mutual_info'])
    assert np.array_equal(out['orientations'], o['rmax'])
    assert (out['orientations'][~keep] == -1).all()
    m1, m2 = p['model1'].copy(), o['model2'].copy()
    core.update_model(m1, m2, o['inter_weight'], 0.)
    assert P.rel_err(out['intens'][0], m1) <= TOL
    # EMC.log: header without a blacklist count (num_blacklist == 0), then the iteration line
    # with the mutual information, the average log-likelihood and the mean beta over N_d
    assert any(line.strip() == 'num_data = %d' % D for line in lines)
    it_line = lines[-1].split('\t')
    assert int(it_line[0]) == 1
    assert abs(float(it_line[3]) - o['mutual_info']) <= 1e-6  # written with %f
    assert abs(float(it_line[4]) - o['avg_likelihood']) <= 1e-6 * abs(o['avg_likelihood'])
    assert abs(float(it_line[6]) - 0.7 * keep.sum() / D) <= 1e-6


def test_model_slice_gen_merge_reference_signature():
    """Model.slice_gen(quat, det, mode=0, view=None) and Model.slice_merge(quat, view, det,
    mode=0) with the reference's signatures (model.pyx:105-150): the view of a chosen mode
    bit for bit as the oracle's slice_gen3d, written into `view` when given, and a merge
    that adds to that mode's model2 / inter_weight only."""
    from oracle import core
    p = P.build(P.tiny_config())
    det = df.CDetector.from_arrays(p['det']['qvals'], p['det']['corr'], p['det']['mask'])
    S = p['model1'].shape[0]
    m = df.Model(S, 2)
    m.allocate('')
    m.model1 = np.stack([p['model1'], p['model1'][::-1].copy()])
    q = p['quats'][17]
    for mode in (0, 1):
        want = core.slice_gen3d(q, m.model1[mode], det.qvals, det.corr)
        got = m.slice_gen(q, det, mode=mode)
        assert np.array_equal(got.view('u8'), want.view('u8')), mode
    buf = np.zeros(det.num_pix)
    assert m.slice_gen(q, det, 1, buf) is buf and np.array_equal(buf, got)
    view = np.random.default_rng(5).random(det.num_pix)
    m2, w2 = np.zeros((S,) * 3), np.zeros((S,) * 3)
    core.slice_merge3d(q, view, m2, w2, det.qvals, det.corr, det.raw_mask)
    m.slice_merge(q, view, det, mode=1)
    assert P.rel_err(m.model2[1], m2) < 1e-14 and P.rel_err(m.inter_weight[1], w2) < 1e-14
    assert not m.model2[0].any() and not m.inter_weight[0].any()
    with pytest.raises(ValueError):
        m.slice_gen(q, det, mode=2)
