"""Learned orientation proposal: a BF16 multilayer perceptron that maps a polar photon
histogram of a frame to logits over a coarse orientation set, trained on the GPU on
(pattern, EMC posterior) pairs that earlier iterations produce.

Every arithmetic step runs in this package's own HIP kernels (csrc/emc_nn.hip: matrix-core
products with FP32 accumulation, softmax cross entropy, column sums, Adam, histogram,
top-k) through the C ABI; torch only allocates the device buffers and provides the stream.

The proposals feed refinement: the top-k coarse cells of each frame become its coarse set
(Iterate.setup_refine / Engine.set_refine), and the fine search is restricted to the Voronoi
cells of those orientations.
"""
import ctypes as C
import math

import numpy as np

from . import _lib

TILE = 256      # product tile (rows of both operands, output columns)
KSTEP = 32      # product K step


def _round(x, m):
    return (x + m - 1) // m * m


def polar_bins(qvals, mask, num_r=64, num_a=64):
    """Bin id of every pixel: radial bin of |q| over [0, max |q|] times angular bin of
    atan2(qy, qx); -1 for pixels with mask >= 1."""
    q = np.asarray(qvals, dtype='f8')
    r = np.sqrt((q ** 2).sum(1))
    rb = np.minimum((r / max(r.max(), 1e-30) * num_r).astype(int), num_r - 1)
    ab = np.minimum(((np.arctan2(q[:, 1], q[:, 0]) + math.pi) / (2 * math.pi) * num_a).astype(int), num_a - 1)
    ids = (rb * num_a + ab).astype('i4')
    ids[np.asarray(mask) >= 1] = -1
    return ids


def photon_lists(frames, mask):
    """Per-frame photon lists of mask < 1 pixels from sparse frames (ones / multi arrays):
    (off int64 (num_frames + 1), pix int32, cnt float64), frames in order."""
    ones, multi = np.asarray(frames['ones']), np.asarray(frames['multi'])
    D = len(ones)
    fr = np.concatenate([np.repeat(np.arange(D), ones), np.repeat(np.arange(D), multi)])
    pix = np.concatenate([frames['place_ones'], frames['place_multi']]).astype('i4')
    cnt = np.concatenate([np.ones(len(frames['place_ones'])), np.asarray(frames['count_multi'], 'f8')])
    keep = np.asarray(mask)[pix] < 1
    fr, pix, cnt = fr[keep], pix[keep], cnt[keep]
    order = np.argsort(fr, kind='stable')
    off = np.zeros(D + 1, 'i8')
    np.cumsum(np.bincount(fr, minlength=D), out=off[1:])
    return off, pix[order], cnt[order]


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ok(rc, what):
    if rc != 0:
        raise _lib.EMCError('%s: bad shape or argument (%d)' % (what, rc))


class OrientationProposal:
    """MLP nbins -> width (x depth, ReLU) -> num_out logits, BF16 weights and activations,
    FP32 master weights with Adam."""

    def __init__(self, num_out, nbins=4096, width=1024, depth=3, device=0, seed=0):
        import torch
        self.torch = torch
        self.dev = torch.device('cuda', device)
        self.num_out = int(num_out)
        self.nbins = int(nbins)
        dims = [self.nbins] + [int(width)] * int(depth) + [self.num_out]
        self.dims = dims
        self.pdims = [_round(d, TILE) for d in dims]
        rng = np.random.default_rng(seed)
        self.W32, self.Wb, self.Wt, self.b = [], [], [], []
        self.mW, self.vW, self.mb, self.vb, self.gW, self.gb = [], [], [], [], [], []
        for l in range(len(dims) - 1):
            fin, fout = dims[l], dims[l + 1]
            pin, pout = self.pdims[l], self.pdims[l + 1]
            w = np.zeros((pout, pin), 'f4')
            w[:fout, :fin] = rng.normal(0., math.sqrt(2. / fin), (fout, fin))
            W32 = torch.tensor(w, device=self.dev)
            self.W32.append(W32)
            self.Wb.append(W32.to(torch.bfloat16))
            self.Wt.append(W32.t().contiguous().to(torch.bfloat16))
            self.b.append(torch.zeros(pout, dtype=torch.float32, device=self.dev))
            for lst, shape in ((self.mW, (pout, pin)), (self.vW, (pout, pin)), (self.gW, (pout, pin)),
                               (self.mb, (1, pout)), (self.vb, (1, pout)), (self.gb, (1, pout))):
                lst.append(torch.zeros(shape, dtype=torch.float32, device=self.dev))
        self.step = 0
        self._acts = None

    # ---- features ------------------------------------------------------------------------
    def features(self, off, pix, cnt, bin_of):
        """Polar histograms of all frames: off (num_frames + 1) int64, pix int32, cnt float64
        photon lists (mask < 1 photons), bin_of int32 per pixel (polar_bins). Returns a BF16
        (num_frames rounded up to the tile, nbins) device tensor, rows past num_frames zero."""
        torch = self.torch
        n = len(off) - 1
        X = torch.zeros((_round(max(n, 1), TILE), self.pdims[0]), dtype=torch.bfloat16, device=self.dev)
        t_off = torch.as_tensor(np.asarray(off, 'i8'), device=self.dev)
        t_pix = torch.as_tensor(np.asarray(pix, 'i4') if len(pix) else np.zeros(1, 'i4'), device=self.dev)
        t_cnt = torch.as_tensor(np.asarray(cnt, 'f8') if len(cnt) else np.zeros(1), device=self.dev)
        t_bin = torch.as_tensor(np.asarray(bin_of, 'i4'), device=self.dev)
        _ok(_lib.lib().emc_nn_polar_hist(_stream(), _ptr(t_off), _ptr(t_pix), _ptr(t_cnt), _ptr(t_bin),
                                         self.nbins, n, _ptr(X), X.shape[1]), 'polar histogram')
        return X

    # ---- forward / backward ------------------------------------------------------------------
    def forward(self, X, rows):
        """Logits (FP32, padded columns) of the first `rows` rows of X (rows padded to TILE)."""
        torch = self.torch
        L = _lib.lib()
        if X.shape[0] % TILE or X.shape[1] != self.pdims[0]:
            Xp = torch.zeros((_round(max(X.shape[0], 1), TILE), self.pdims[0]), dtype=X.dtype, device=self.dev)
            Xp[:X.shape[0], :X.shape[1]] = X
            X = Xp
        B = X.shape[0]
        acts = [X]
        h = X
        for l in range(len(self.dims) - 1):
            last = l == len(self.dims) - 2
            out = torch.zeros((B, self.pdims[l + 1]), device=self.dev,
                              dtype=torch.float32 if last else torch.bfloat16)
            _ok(L.emc_nn_forward(_stream(), _ptr(h), h.shape[1], _ptr(self.Wb[l]), self.Wb[l].shape[1], B,
                                 self.pdims[l + 1], self.pdims[l], rows, self.pdims[l + 1], _ptr(self.b[l]),
                                 0 if last else 1, 1 if last else 0, _ptr(out), out.shape[1]), 'forward')
            acts.append(out)
            h = out
        self._acts = acts
        return h

    def loss_and_grads(self, X, Y, rows):
        """Forward, soft-target cross entropy against Y (FP32 (B, num_out)), and the gradients
        of the mean loss into gW / gb. Returns the mean loss."""
        torch = self.torch
        L = _lib.lib()
        B = X.shape[0]
        Z = self.forward(X, rows)
        dZ = torch.zeros((B, self.pdims[-1]), dtype=torch.bfloat16, device=self.dev)
        loss = torch.zeros(max(rows, 1), dtype=torch.float32, device=self.dev)
        Yc = Y.contiguous()
        _ok(L.emc_nn_softmax_xent(_stream(), _ptr(Z), Z.shape[1], _ptr(Yc), Yc.shape[1], rows, self.num_out,
                                  C.c_float(1. / rows), _ptr(loss), _ptr(dZ), dZ.shape[1]), 'softmax')
        g = dZ
        for l in range(len(self.dims) - 2, -1, -1):
            A = self._acts[l]                         # input of layer l, (B, pin)
            pin, pout = self.pdims[l], self.pdims[l + 1]
            gT = torch.empty((pout, B), dtype=torch.bfloat16, device=self.dev)
            AT = torch.empty((pin, B), dtype=torch.bfloat16, device=self.dev)
            _ok(L.emc_nn_transpose(_stream(), _ptr(g), g.shape[1], B, pout, _ptr(gT), B), 'transpose')
            _ok(L.emc_nn_transpose(_stream(), _ptr(A), A.shape[1], B, pin, _ptr(AT), B), 'transpose')
            _ok(L.emc_nn_backward_weight(_stream(), _ptr(gT), B, _ptr(AT), B, pout, pin, B, pout, pin,
                                         _ptr(self.gW[l]), pin), 'weight gradient')
            _ok(L.emc_nn_colsum(_stream(), _ptr(g), g.shape[1], B, pout, _ptr(self.gb[l])), 'bias gradient')
            if l > 0:
                dX = torch.zeros((B, pin), dtype=torch.bfloat16, device=self.dev)
                _ok(L.emc_nn_backward_data(_stream(), _ptr(g), g.shape[1], _ptr(self.Wt[l]), self.Wt[l].shape[1],
                                           B, pin, pout, rows, pin, _ptr(A), A.shape[1], _ptr(dX), pin),
                    'data gradient')
                g = dX
        return float(loss.cpu().numpy()[:rows].sum()) / max(rows, 1)

    def apply_adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.):
        L = _lib.lib()
        self.step += 1
        for l in range(len(self.dims) - 1):
            W = self.W32[l]
            _ok(L.emc_nn_adam(_stream(), _ptr(W), _ptr(self.gW[l]), _ptr(self.mW[l]), _ptr(self.vW[l]),
                              W.shape[0], W.shape[1], W.shape[1], C.c_float(lr), C.c_float(b1), C.c_float(b2),
                              C.c_float(eps), C.c_float(wd), self.step, _ptr(self.Wb[l]), _ptr(self.Wt[l]),
                              self.Wt[l].shape[1]), 'adam')
            bvec = self.b[l]
            _ok(L.emc_nn_adam(_stream(), _ptr(bvec), _ptr(self.gb[l]), _ptr(self.mb[l]), _ptr(self.vb[l]), 1,
                              bvec.shape[0], bvec.shape[0], C.c_float(lr), C.c_float(b1), C.c_float(b2),
                              C.c_float(eps), C.c_float(0.), self.step, None, None, 0), 'adam')

    # ---- training and inference ----------------------------------------------------------------
    def fit(self, X, Y, num, epochs=5, batch=1024, lr=1e-3, seed=0, log=None, wd=0.):
        """Minibatch Adam (L2 weight decay wd on the weights) on the first `num` rows of X
        (features) and Y (targets, rows summing to 1). Returns the mean loss of every epoch."""
        torch = self.torch
        batch = _round(batch, TILE)
        gen = torch.Generator(device='cpu')
        gen.manual_seed(seed)
        Xb = torch.zeros((batch, X.shape[1]), dtype=X.dtype, device=self.dev)
        Yb = torch.zeros((batch, Y.shape[1]), dtype=Y.dtype, device=self.dev)
        history = []
        for ep in range(epochs):
            perm = torch.randperm(num, generator=gen).to(self.dev)
            tot, cnt = 0., 0
            for b0 in range(0, num, batch):
                idx = perm[b0:b0 + batch]
                rows = idx.numel()
                Xb.zero_()
                Yb.zero_()
                Xb[:rows] = X[idx]
                Yb[:rows] = Y[idx]
                tot += self.loss_and_grads(Xb, Yb, rows) * rows
                cnt += rows
                self.apply_adam(lr, wd=wd)
            history.append(tot / max(cnt, 1))
            if log:
                log('epoch %d loss %.4f' % (ep + 1, history[-1]))
        return history

    def logits(self, X, num):
        return self.forward(X, num)[:num, :self.num_out]

    def propose(self, X, num, k, with_mass=False):
        """Top-k output indices (descending logit) of the first `num` rows: (num, k) int32;
        with_mass also returns the softmax probability each row's k outputs hold."""
        torch = self.torch
        Z = self.forward(X, num)
        idx = torch.empty((max(num, 1), k), dtype=torch.int32, device=self.dev)
        mass = torch.empty(max(num, 1), dtype=torch.float32, device=self.dev)
        _ok(_lib.lib().emc_nn_topk(_stream(), _ptr(Z), Z.shape[1], num, self.num_out, k, _ptr(idx), _ptr(mass)),
            'top-k')
        if with_mass:
            return idx[:num].cpu().numpy(), mass[:num].cpu().numpy()
        return idx[:num].cpu().numpy()


def topk_device(Z, rows, n, k):
    """Top-k column indices (descending, ties to the lower index) of the first `rows` rows of
    an FP32 device matrix Z (row stride Z.shape[1]), with the softmax mass they hold, on the
    GPU (emc_nn_topk). Returns host arrays ((rows, k) int32, (rows,) float32)."""
    import torch
    idx = torch.empty((max(rows, 1), k), dtype=torch.int32, device=Z.device)
    mass = torch.empty(max(rows, 1), dtype=torch.float32, device=Z.device)
    _ok(_lib.lib().emc_nn_topk(_stream(), _ptr(Z), Z.shape[1], rows, n, k, _ptr(idx), _ptr(mass)), 'top-k')
    return idx[:rows].cpu().numpy(), mass[:rows].cpu().numpy()


def union_cells(*tops):
    """Per-frame union of several (num_frames, k_i) cell arrays, as CSR (off int64
    (num_frames + 1), cells int32 ascending within a frame)."""
    allc = np.sort(np.concatenate([np.asarray(t, 'i4') for t in tops], axis=1), axis=1)
    keep = np.ones(allc.shape, bool)
    keep[:, 1:] = allc[:, 1:] != allc[:, :-1]
    off = np.zeros(allc.shape[0] + 1, 'i8')
    np.cumsum(keep.sum(1), out=off[1:])
    return off, allc[keep].astype('i4')


class CoarseJudge:
    """The full-search half of the likelihood judge: a search over the coarse orientations
    quat_gen(coarse_div) (every frame, every coarse orientation -- the coarse orientation c is
    the centre of Voronoi cell c of the fine sampling) stopped after the posterior
    (estep_only), at the same model, scale factors and beta as the fine iteration it judges.
    Its posterior proposes cells of its own (top_cells), and against the fine search's maximum
    inside a frame's proposal it estimates the mass the full fine search would put outside it
    (Engine.judge_outside; judged_iteration)."""

    def __init__(self, itr, coarse_div, device=0, precision=0):
        det = itr.dets[0]
        self.eng = eng = _lib.Engine(device)
        eng.set_precision(precision)
        eng.set_detector(det.qvals, det.corr, det.raw_mask)
        self.quats = _lib.quat_gen(coarse_div)
        self.num_coarse = len(self.quats)
        eng.set_quats(self.quats)
        eng.set_model(itr.model.model1)
        sp = itr.sparse_arrays()
        eng.set_frames(sp['ones'], sp['multi'], sp['ones_accum'], sp['multi_accum'], sp['place_ones'],
                       sp['place_multi'], sp['count_multi'])
        eng.set_frame_params(itr.scale, itr.beta, itr.blacklist, itr.fcounts, itr.sum_fact, itr.mean_count[0],
                             0)
        eng.set_option('estep_only', 1)
        self.stats = None

    def estep(self, model1, scale, beta, iteration, need_scaling):
        """Coarse posterior of every frame at this model / scale factors / beta; model1 is a
        host array or an Engine holding the model (copied on the device)."""
        eng = self.eng
        if isinstance(model1, np.ndarray):
            eng.set_model(model1)
        else:
            eng.copy_model(model1)
        eng.set_scale(scale)
        eng.set_beta(beta)
        self.stats = eng.maximize(iteration, need_scaling, 0, 0)
        return self.stats

    def top_cells(self, k):
        """Each frame's k most likely coarse cells, (num_frames, k) int32."""
        idx, _ = self.eng.cell_topk(self.num_coarse, np.arange(self.num_coarse, dtype='i4'), k)
        return idx

    def argmax_cells(self):
        """Each frame's most likely coarse orientation (cell), (num_frames, 1) int32; a frame
        with no orientation (blacklisted) gets cell 0."""
        return np.maximum(self.eng.get('rmax'), 0).astype('i4')[:, None]

    def mass(self, rel_off, rel_quat):
        """Coarse posterior mass inside each frame's cells rel_quat[rel_off[d]:rel_off[d + 1]]."""
        return self.eng.subset_mass(rel_off, rel_quat)

    def close(self):
        self.eng.close()


def stack_cells(*tops):
    """Per-frame lists of several (num_frames, k_i) cell arrays side by side, as CSR (off int64
    (num_frames + 1), cells int32) of fixed width sum(k_i). A cell may appear more than once in a
    frame's list; emc_set_refine takes a frame's set as the union of its list."""
    X = np.ascontiguousarray(np.concatenate([np.asarray(t, 'i4') for t in tops], axis=1))
    D, k = X.shape
    return np.arange(D + 1, dtype='i8') * k, X.reshape(-1)


def refine_rows(off, cells, accept):
    """refine_lists_csr for lists of one fixed width (stack_cells): kept frames keep their
    lists, the others get the single entry -1 (the full search)."""
    accept = np.asarray(accept, bool)
    D = len(accept)
    k = int(off[1] - off[0]) if D else 0
    X = np.asarray(cells, 'i4').reshape(D, k)
    n = np.where(accept, k, 1)
    rel_off = np.zeros(D + 1, 'i8')
    np.cumsum(n, out=rel_off[1:])
    out = np.full((D, k), -2, 'i4')
    out[accept] = X[accept]
    out[~accept, 0] = -1
    return rel_off, out[out != -2]


def judged_iteration(recon, itr, judge, tops, coarse_div, threshold=0.99, widen=0):
    """One EMC iteration (recon.run_iteration) on proposals judged by the likelihood.

    Pass 1: the expectation step of the fine search restricted to each frame's proposed cells
    (the union of its rows of the (num_frames, k_i) arrays in `tops`), which gives every frame its fine
    maximum and normaliser inside the cells. The judge (a coarse full search already run at the
    same model, scales and beta: CoarseJudge.estep) then estimates, per frame, the posterior
    mass the full fine search would put outside the cells (Engine.judge_outside). Frames
    estimated to hold at least `threshold` inside keep their cells; the rest fall back to the
    full fine search. With `widen` > 0 a rejected frame instead gets the fewest further coarse
    cells (at most `widen`) after which the judge's estimate, an upper bound taken with the
    fine maximum and normaliser of its first cells, is at most (1 - threshold) / 10
    (Engine.judge_widen: the added cells are judged at the coarse sampling only, hence the
    margin), and falls back to the full search only when more are needed. Pass 2:
    when every frame is kept, the M-step resumes from pass 1 (option mstep_only); otherwise the
    iteration runs again with the widened and fall-back frames. Returns a dict: accept (bool
    per frame: kept as proposed), widened (bool per frame), outside (estimates), passes (1 or
    2)."""
    import time
    eng = recon.engine
    par = itr.params
    t = [time.perf_counter()]
    off, cells = stack_cells(*tops)
    itr.set_refine_lists(coarse_div, off, cells)
    eng.set_refine(itr.num_coarse, itr.quat_mapping, itr.rel_off, itr.rel_quat)
    recon.update_beta()  # the beta of this iteration (run_iteration sets the same again)
    t.append(time.perf_counter())
    eng.set_option('estep_only', 1)
    try:
        st1 = eng.maximize(par.iteration, par.need_scaling, par.update_scale, par.known_scale)
    finally:
        eng.set_option('estep_only', 0)
    t.append(time.perf_counter())
    if widen > 0:
        outside, num_add, added = eng.judge_widen(judge.eng, 1. - threshold, widen)
        accept = num_add == 0
    else:
        outside = eng.judge_outside(judge.eng)
        accept = outside <= 1. - threshold
        num_add = np.where(accept, 0, -1)
    if accept.all():
        t.append(time.perf_counter())
        eng.set_option('mstep_only', 1)
        try:
            recon.run_iteration()
        finally:
            eng.set_option('mstep_only', 0)
        passes = 1
    else:
        if widen > 0:
            rel_off, rel_quat = widen_lists_csr(off, cells, num_add, added)
        else:
            rel_off, rel_quat = refine_rows(off, cells, accept)
        itr.set_refine_lists(coarse_div, rel_off, rel_quat)
        eng.set_refine(itr.num_coarse, itr.quat_mapping, itr.rel_off, itr.rel_quat)
        t.append(time.perf_counter())
        recon.run_iteration()
        passes = 2
    t.append(time.perf_counter())
    host_ms = dict(zip(('lists', 'pass1', 'judge_and_lists', 'pass2'), [round((b - a) * 1e3, 1) for a, b in zip(t, t[1:])]))
    # distinct cells per frame, on a sample of frames (the lists may repeat cells)
    X = cells.reshape(len(off) - 1, -1)[:: max(1, (len(off) - 1) // 2000)]
    distinct = (np.diff(np.sort(X, axis=1), axis=1) != 0).sum(1) + 1
    return dict(accept=accept, widened=num_add > 0, outside=outside, passes=passes,
                cells_per_frame=float(distinct.mean()),
                estep_stage_ms=st1['stage_ms'], phase_ms=host_ms)


def refine_lists(top, mass, num_coarse, threshold=0.99, buf=None):
    """refine_sets flattened for Iterate.set_refine_lists: (rel_off, rel_quat, fallback, buf).
    rel_quat is a view of `buf` (reused when large enough, so repeated calls do not allocate
    and fault in a new list of every cell for every fall-back frame)."""
    top = np.asarray(top)
    D, k = top.shape
    fallback = np.asarray(mass) < threshold
    rel_off = np.zeros(D + 1, dtype='i8')
    np.cumsum(np.where(fallback, num_coarse, k), out=rel_off[1:])
    total = int(rel_off[-1])
    if buf is None or buf.size < total:
        buf = np.empty(max(total, 1), dtype='i4')
    rel_quat = buf[:total]
    keep = ~fallback
    if keep.any():
        idx = rel_off[:-1][keep][:, None] + np.arange(k)[None, :]
        rel_quat[idx.ravel()] = np.sort(top[keep], axis=1).astype('i4').ravel()
    every = np.arange(num_coarse, dtype='i4')
    for s in rel_off[:-1][fallback]:
        rel_quat[s:s + num_coarse] = every
    return rel_off, rel_quat, fallback, buf


def _ranges(starts, lens):
    """Concatenation of arange(s, s + n) over (s, n) pairs, int64."""
    lens = np.asarray(lens, 'i8')
    tot = int(lens.sum())
    if tot == 0:
        return np.zeros(0, 'i8')
    first = np.cumsum(lens) - lens
    return np.repeat(np.asarray(starts, 'i8') - first, lens) + np.arange(tot, dtype='i8')


def refine_lists_csr(off, cells, accept, num_coarse=None):
    """Refinement lists from per-frame proposals in CSR form (off, cells): a frame keeps its
    cells where `accept`; elsewhere its list is the single entry -1, which emc_set_refine
    reads as every cell (the full search). Returns (rel_off, rel_quat)."""
    off = np.asarray(off, 'i8')
    accept = np.asarray(accept, bool)
    n = np.diff(off)
    rel_off = np.zeros(len(n) + 1, 'i8')
    np.cumsum(np.where(accept, n, 1), out=rel_off[1:])
    rel_quat = np.full(int(rel_off[-1]), -1, 'i4')
    rel_quat[_ranges(rel_off[:-1][accept], n[accept])] = np.asarray(cells, 'i4')[_ranges(off[:-1][accept], n[accept])]
    return rel_off, rel_quat


def widen_lists_csr(off, cells, num_add, add):
    """Refinement lists after the judge's widening (Engine.judge_widen): a frame with num_add 0
    keeps its cells, one with k > 0 gets its cells followed by its k added cells (`add`, frame
    after frame), one with -1 the single entry -1 (the full search). Returns (rel_off,
    rel_quat)."""
    off = np.asarray(off, 'i8')
    num_add = np.asarray(num_add, 'i8')
    n = np.diff(off)
    full = num_add < 0
    grow = np.maximum(num_add, 0)
    rel_off = np.zeros(len(n) + 1, 'i8')
    np.cumsum(np.where(full, 1, n + grow), out=rel_off[1:])
    rel_quat = np.full(int(rel_off[-1]), -1, 'i4')
    keep = ~full
    rel_quat[_ranges(rel_off[:-1][keep], n[keep])] = np.asarray(cells, 'i4')[_ranges(off[:-1][keep], n[keep])]
    wide = grow > 0
    rel_quat[_ranges(rel_off[:-1][wide] + n[wide], grow[wide])] = np.asarray(add, 'i4')
    return rel_off, rel_quat


def refine_sets(top, mass, num_coarse, threshold=0.99):
    """Per-frame coarse sets for refinement from proposals: the frame's top-k cells, or every
    cell (fall back to the full search) where the proposal holds less than `threshold` of the
    network's probability. Returns (list of int32 arrays, fallback mask)."""
    fallback = np.asarray(mass) < threshold
    every = np.arange(num_coarse, dtype='i4')
    sets = [every if fb else np.sort(np.asarray(t, 'i4')) for t, fb in zip(top, fallback)]
    return sets, fallback
