"""Orientation-proposal network on the GPU (csrc/emc_nn.hip through the C ABI) against the
numpy oracle (oracle/mlp.py), and what the trained network gains over the search without it.

Tolerances: BF16 activations and gradients (8-bit mantissa) with FP32 accumulation, against
float64 on the same BF16-rounded weights and inputs: relative Frobenius error <= 3e-2 for
gradients and logits, 1e-2 for the loss."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


def _bf(t):
    return t.float().cpu().numpy().astype('f8')


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope='module')
def small_net():
    from dragonfly_amd.proposal import OrientationProposal
    return OrientationProposal(300, nbins=512, width=256, depth=2, seed=3)


def test_forward_and_gradients_vs_oracle(small_net):
    import torch
    from oracle import mlp
    net = small_net
    rng = np.random.default_rng(1)
    B, rows = 256, 200
    X = torch.zeros((B, net.pdims[0]), dtype=torch.bfloat16, device='cuda')
    X[:rows, :net.nbins] = torch.tensor(rng.random((rows, net.nbins)) * 2., dtype=torch.bfloat16)
    Yn = rng.random((rows, net.num_out)) ** 4
    Yn /= Yn.sum(1, keepdims=True)
    Y = torch.zeros((B, net.num_out), dtype=torch.float32, device='cuda')
    Y[:rows] = torch.tensor(Yn, dtype=torch.float32)
    loss = net.loss_and_grads(X, Y, rows)
    Ws = [_bf(w)[:net.dims[l + 1], :net.dims[l]] for l, w in enumerate(net.Wb)]
    bs = [b.cpu().numpy()[:net.dims[l + 1]].astype('f8') for l, b in enumerate(net.b)]
    Xn = _bf(X)[:rows, :net.nbins]
    acts = mlp.forward(Xn, Ws, bs)
    logits = net._acts[-1][:rows, :net.num_out].cpu().numpy()
    ref_loss, gW, gb = mlp.loss_and_grads(Xn, Yn, Ws, bs)
    errs = {'logits': _rel(logits, acts[-1]), 'loss': abs(loss - ref_loss) / ref_loss}
    for l in range(len(Ws)):
        errs['gW%d' % l] = _rel(net.gW[l].cpu().numpy()[:net.dims[l + 1], :net.dims[l]], gW[l])
        errs['gb%d' % l] = _rel(net.gb[l].cpu().numpy()[0, :net.dims[l + 1]], gb[l])
        # padding rows and columns of the gradient stay zero
        assert not net.gW[l].cpu().numpy()[net.dims[l + 1]:].any()
    print({k: '%.2e' % v for k, v in errs.items()}, flush=True)
    assert errs['loss'] <= 1e-2
    assert all(v <= 3e-2 for k, v in errs.items()), errs


def test_adam_step_vs_oracle(small_net):
    from oracle import mlp
    net = small_net
    w0 = net.W32[0].cpu().numpy().astype('f8')
    g = net.gW[0].cpu().numpy().astype('f8')
    m0 = net.mW[0].cpu().numpy().astype('f8')
    v0 = net.vW[0].cpu().numpy().astype('f8')
    step = net.step + 1
    net.apply_adam(lr=1e-3)
    w1, _, _ = mlp.adam(w0, g, m0, v0, step, 1e-3)
    assert np.allclose(net.W32[0].cpu().numpy(), w1, rtol=1e-5, atol=1e-7)
    assert np.array_equal(_bf(net.Wb[0]), _bf(net.Wt[0]).T)


def test_topk_and_polar_histogram():
    import torch
    from dragonfly_amd import proposal
    net = proposal.OrientationProposal(300, nbins=64, width=256, depth=1, seed=0)
    rng = np.random.default_rng(5)
    rows = 100
    Z = torch.tensor(rng.permutation(rows * 300).reshape(rows, 300).astype('f4'), device='cuda')
    Zp = torch.zeros((256, 512), dtype=torch.float32, device='cuda')
    Zp[:rows, :300] = Z
    import ctypes as C
    from dragonfly_amd import _lib
    idx = torch.empty((rows, 7), dtype=torch.int32, device='cuda')
    mass = torch.empty(rows, dtype=torch.float32, device='cuda')
    Zs = Z / 3000.
    Zp[:rows, :300] = Zs
    assert _lib.lib().emc_nn_topk(proposal._stream(), proposal._ptr(Zp), 512, rows, 300, 7,
                                  proposal._ptr(idx), proposal._ptr(mass)) == 0
    zn = Zs.cpu().numpy().astype('f8')
    want = np.argsort(-zn, axis=1)[:, :7]
    assert np.array_equal(idx.cpu().numpy(), want)
    pz = np.exp(zn - zn.max(1, keepdims=True))
    pz /= pz.sum(1, keepdims=True)
    assert np.allclose(mass.cpu().numpy(), np.take_along_axis(pz, want, 1).sum(1), rtol=1e-4)
    # histogram: 3 frames over 10 pixels in 4 bins (pixel 9 masked)
    bin_of = np.array([0, 0, 1, 1, 2, 2, 3, 3, 3, -1], 'i4')
    off = np.array([0, 3, 3, 6], 'i8')
    pix = np.array([0, 2, 9, 4, 5, 8], 'i4')
    cnt = np.array([1., 3., 5., 2., 2., 4.])
    X = net.features(off, pix, cnt, bin_of)
    got = X.float().cpu().numpy()[:3, :4]
    hist = np.zeros((3, 4))
    for d in range(3):
        for k in range(off[d], off[d + 1]):
            if bin_of[pix[k]] >= 0:
                hist[d, bin_of[pix[k]]] += cnt[k]
    tot = hist.sum(1, keepdims=True)
    want = np.log1p(1000. * np.divide(hist, tot, out=np.zeros_like(hist), where=tot > 0))
    assert np.allclose(got, want, rtol=1e-2, atol=1e-2)


def test_trained_proposal_captures_posterior_mass():
    """cfg1 geometry, 6000 frames of ~1000 photons at random orientations; targets are the
    posteriors over the 60 num_div-1 orientations from one FP64 iteration at the true model.
    Trained on 5000 frames, the network's top-4 cells capture far more of the held-out
    frames' posterior mass than the best frame-independent choice of 4 cells (the method
    without the network: a fixed subset, or the full search)."""
    import torch
    from dragonfly_amd import _lib, synth
    from dragonfly_amd.proposal import OrientationProposal, polar_bins
    cfg = dict(synth.CONFIGS['cfg1'], num_data=6000, mean_count=1000., num_div=1)
    p = P.build(cfg, device='cuda', num_ave=1000)
    p['model1'] = p['intens'] * (p['fparams']['mean_count'] / 1000.)
    eng = P.engine_for(p)
    g = P.run_engine(eng, dict(iteration=1, need_scaling=0))
    eng.close()
    post = g['prob']                               # (6000, 60)
    f = p['frames']
    # photon lists (mask < 1) per frame for the histogram
    mask = p['det']['mask']
    offs, pixs, cnts = [0], [], []
    for d in range(len(f['ones'])):
        po = f['place_ones'][f['ones_accum'][d]:f['ones_accum'][d] + f['ones'][d]]
        pm = f['place_multi'][f['multi_accum'][d]:f['multi_accum'][d] + f['multi'][d]]
        cm = f['count_multi'][f['multi_accum'][d]:f['multi_accum'][d] + f['multi'][d]]
        pixs += [po, pm]
        cnts += [np.ones(len(po)), cm.astype('f8')]
        offs.append(offs[-1] + len(po) + len(pm))
    net = OrientationProposal(60, nbins=4096, width=512, depth=2, seed=1)
    X = net.features(np.array(offs, 'i8'), np.concatenate(pixs), np.concatenate(cnts),
                     polar_bins(p['det']['qvals'], mask))
    Y = torch.tensor(post, dtype=torch.float32, device='cuda')
    ntr = 5000
    hist = net.fit(X, Y, ntr, epochs=30, batch=512, lr=1e-3)
    k = 4
    Xte = X[ntr:].contiguous()
    top = net.propose(Xte, 1000, k)
    test = post[ntr:]
    mass_net = np.take_along_axis(test, top, 1).sum(1).mean()
    fixed = np.argsort(-post[:ntr].sum(0))[:k]
    mass_fixed = test[:, fixed].sum(1).mean()
    print('loss %.3f -> %.3f; held-out mass captured by top-%d: network %.3f, fixed subset %.3f'
          % (hist[0], hist[-1], k, mass_net, mass_fixed), flush=True)
    assert hist[-1] < hist[0]
    assert mass_net > 3 * mass_fixed
