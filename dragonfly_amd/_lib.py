"""ctypes binding of libemc_hip.so (declared in include/emc_hip.h).

The GPU engine is the only compute path: if the library is missing or no GPU is present,
the calls below raise instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libemc_hip.so')

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_u8p = C.POINTER(C.c_uint8)

# enum emc_field
FIELDS = {
    'u': (0, 'f8', 'rot'), 'max_exp': (1, 'f8', 'data'), 'p_norm': (2, 'f8', 'data'),
    'rmax': (3, 'i4', 'data'), 'likelihood': (4, 'f8', 'data'), 'info': (5, 'f8', 'data'),
    'psum_d': (6, 'f8', 'data'), 'scale': (7, 'f8', 'data'), 'model1': (8, 'f8', 'vol'),
    'model2': (9, 'f8', 'vol'), 'inter_weight': (10, 'f8', 'vol'), 'psum_r': (11, 'f8', 'rot'),
    'prob': (12, 'f8', 'data_rot'), 'quat_norm': (13, 'f8', 'data_mode'),
}

STAGES = ('expand', 'likelihood', 'rescale', 'max', 'norm', 'posterior', 'tomogram', 'reduce',
          'merge')


class IterOpts(C.Structure):
    _fields_ = [('iteration', C.c_int), ('need_scaling', C.c_int), ('update_scale', C.c_int),
                ('known_scale', C.c_int)]


class IterStats(C.Structure):
    _fields_ = [('avg_likelihood', C.c_double), ('mutual_info', C.c_double),
                ('rescale', C.c_double), ('stage_ms', C.c_double * len(STAGES)), ('num_chunks', C.c_int),
                ('view_nonzero', C.c_double)]


class EMCError(RuntimeError):
    """Error reported by the engine (negative return code)."""


_lib = None


def lib():
    """Load libemc_hip.so, raising if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('libemc_hip.so is not built; run `python -m dragonfly_amd.build`')
        L = C.CDLL(LIB_PATH)
        sig = {
            'emc_create': (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
            'emc_destroy': (None, [C.c_void_p]),
            'emc_last_error': (C.c_char_p, []),
            'emc_synchronize': (C.c_int, [C.c_void_p]),
            'emc_num_rot': (C.c_int, [C.c_int]),
            'emc_quat_gen': (C.c_int, [C.c_int, _dp]),
            'emc_set_detector': (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _u8p]),
            'emc_set_modes': (C.c_int, [C.c_void_p, C.c_int]),
            'emc_set_background': (C.c_int, [C.c_void_p, _dp, _dp]),
            'emc_set_quats': (C.c_int, [C.c_void_p, C.c_int, _dp]),
            'emc_set_model': (C.c_int, [C.c_void_p, C.c_long, _dp]),
            'emc_copy_model': (C.c_int, [C.c_void_p, C.c_void_p]),
            'emc_set_frames_sparse': (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _lp, _lp, _ip, _ip, _ip]),
            'emc_set_frames_dense': (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip]),
            'emc_voronoi_subset': (C.c_int, [C.c_int, _dp, C.c_int, _dp, _ip]),
            'emc_nn_forward': (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
            'emc_nn_backward_data': (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
            'emc_nn_backward_weight': (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
            'emc_nn_transpose': (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
            'emc_nn_softmax_xent': (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_void_p, C.c_size_t]),
            'emc_nn_colsum': (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_void_p]),
            'emc_nn_adam': (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t]),
            'emc_nn_polar_hist': (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
            'emc_nn_topk': (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
            'emc_set_precision': (C.c_int, [C.c_void_p, C.c_int]),
            'emc_set_refine': (C.c_int, [C.c_void_p, C.c_int, _ip, _lp, _ip]),
            'emc_set_frame_params': (C.c_int, [C.c_void_p, _dp, _dp, _u8p, _ip, _dp, C.c_double, C.c_int]),
            'emc_set_beta': (C.c_int, [C.c_void_p, _dp]),
            'emc_set_scale': (C.c_int, [C.c_void_p, _dp]),
            'emc_maximize': (C.c_int, [C.c_void_p, C.POINTER(IterOpts), C.POINTER(IterStats)]),
            'emc_update_model': (C.c_int, [C.c_void_p, C.c_double, _dp]),
            'emc_normalize_scale': (C.c_int, [C.c_void_p, _dp]),
            'emc_slice_gen': (C.c_int, [C.c_void_p, _dp, _dp]),
            'emc_slice_gen_static': (C.c_int, [C.c_void_p, _dp, _dp]),
            'emc_subset_mass': (C.c_int, [C.c_void_p, _lp, _ip, _dp]),
            'emc_judge_outside': (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, _dp]),
            'emc_judge_widen': (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _ip, _lp]),
            'emc_judge_widen_cells': (C.c_int, [C.c_void_p, _ip, C.c_long]),
            'emc_cell_topk': (C.c_int, [C.c_void_p, C.c_int, _ip, C.c_int, _ip, C.POINTER(C.c_float)]),
            'emc_slice_merge': (C.c_int, [C.c_void_p, _dp, _dp]),
            'emc_get_field': (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
            'emc_get_sparse_prob': (C.c_int, [C.c_void_p, _ip, C.c_long, _ip, _dp, _lp]),
            'emc_clear_merge': (C.c_int, [C.c_void_p]),
            'emc_cell_posterior': (C.c_int, [C.c_void_p, C.c_int, _ip, C.POINTER(C.c_float)]),
            'emc_set_option': (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
            'emc_get_option': (C.c_int, [C.c_void_p, C.c_char_p, _ip]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise EMCError('engine error %d: %s' % (rc, lib().emc_last_error().decode()))


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def num_rot(num_div):
    n = lib().emc_num_rot(int(num_div))
    if n < 0:
        raise ValueError('num_div %d out of range' % num_div)
    return n


def voronoi_subset(qcoarse, qfine):
    """Index of the nearest coarse orientation of every fine orientation (quaternion.c:844-862)."""
    qc = np.ascontiguousarray(qcoarse, dtype='f8')
    qf = np.ascontiguousarray(qfine, dtype='f8')
    out = np.zeros(len(qf), dtype='i4')
    _check(lib().emc_voronoi_subset(len(qc), _ptr(qc), len(qf), _ptr(qf), _ptr(out, _ip)))
    return out


def quat_gen(num_div):
    """quat_gen (dragonfly/src/quaternion.c:811-842): (num_rot, 5) array, bit-exact."""
    q = np.zeros((num_rot(num_div), 5))
    n = lib().emc_quat_gen(int(num_div), _ptr(q))
    if n != q.shape[0]:
        raise EMCError('emc_quat_gen failed (%d)' % n)
    return q


class Engine:
    """One device-resident EMC problem (an ``emc_engine`` handle)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().emc_create(int(device), C.byref(h)))
        self._h = h
        self.num_pix = self.num_rot = self.num_data = self.size = 0
        self.num_rot_user, self.num_modes = 0, 1

    def close(self):
        if getattr(self, '_h', None):
            lib().emc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_detector(self, qvals, corr, mask):
        qvals = np.ascontiguousarray(qvals, dtype='f8').reshape(-1, 3)
        corr = np.ascontiguousarray(corr, dtype='f8')
        mask = np.ascontiguousarray(mask, dtype='u1')
        if not (qvals.shape[0] == corr.shape[0] == mask.shape[0]):
            raise ValueError('detector arrays disagree on num_pix')
        _check(lib().emc_set_detector(self._h, corr.shape[0], _ptr(qvals), _ptr(corr), _ptr(mask, _u8p)))
        self.num_pix = corr.shape[0]

    def set_modes(self, num_modes):
        """Number of model modes (params num_modes). The engine's orientations become the
        combined r = rotation * num_modes + mode; the model must be set again afterwards."""
        _check(lib().emc_set_modes(self._h, int(num_modes)))
        if int(num_modes) != self.num_modes:
            self.size = 0
        self.num_modes = int(num_modes)
        self.num_rot = self.num_rot_user * self.num_modes

    def set_quats(self, quats):
        quats = np.ascontiguousarray(quats, dtype='f8')
        if quats.ndim != 2 or quats.shape[1] != 5:
            raise ValueError('quats must be (num_rot, 5)')
        _check(lib().emc_set_quats(self._h, quats.shape[0], _ptr(quats)))
        self.num_rot_user = quats.shape[0]
        self.num_rot = self.num_rot_user * self.num_modes

    def set_model(self, model1):
        """model1: (size, size, size) for one mode, (num_modes, size, size, size) for any."""
        model1 = np.ascontiguousarray(model1, dtype='f8')
        if model1.ndim == 3 and self.num_modes == 1:
            model1 = model1[None]
        if model1.ndim != 4 or model1.shape[0] != self.num_modes or len(set(model1.shape[1:])) != 1:
            raise ValueError('model1 must be %d cube(s)' % self.num_modes)
        _check(lib().emc_set_model(self._h, model1.shape[1], _ptr(model1)))
        self.size = model1.shape[1]

    def set_frames(self, ones, multi, ones_accum, multi_accum, place_ones, place_multi, count_multi):
        a = [np.ascontiguousarray(x, dtype=t) for x, t in (
            (ones, 'i4'), (multi, 'i4'), (ones_accum, 'i8'), (multi_accum, 'i8'),
            (place_ones, 'i4'), (place_multi, 'i4'), (count_multi, 'i4'))]
        for i in (4, 5, 6):
            if a[i].size == 0:
                a[i] = np.zeros(1, 'i4')
        _check(lib().emc_set_frames_sparse(
            self._h, a[0].shape[0], _ptr(a[0], _ip), _ptr(a[1], _ip), _ptr(a[2], _lp),
            _ptr(a[3], _lp), _ptr(a[4], _ip), _ptr(a[5], _ip), _ptr(a[6], _ip)))
        self.num_data = a[0].shape[0]

    def set_frames_dense(self, frames):
        frames = np.ascontiguousarray(frames, dtype='i4')
        if frames.ndim != 2:
            raise ValueError('dense frames must be (num_data, num_pix)')
        _check(lib().emc_set_frames_dense(self._h, frames.shape[0], frames.shape[1], _ptr(frames, _ip)))
        self.num_data = frames.shape[0]


    def set_background(self, background, bgscale):
        """Detector background (num_pix,) and per-frame background scales (num_data,); None
        switches it off. Used by the iteration when need_scaling is set."""
        if background is None:
            _check(lib().emc_set_background(self._h, None, None))
            return
        bg = np.ascontiguousarray(background, dtype='f8')
        bs = np.ascontiguousarray(bgscale, dtype='f8')
        if bg.shape != (self.num_pix,) or bs.shape != (self.num_data,):
            raise ValueError('background must be (num_pix,) and bgscale (num_data,)')
        _check(lib().emc_set_background(self._h, _ptr(bg), _ptr(bs)))

    def set_refine(self, num_coarse, quat_mapping, rel_off, rel_quat):
        """Per-frame orientation subsets (refinement); quat_mapping None switches it off."""
        if quat_mapping is None:
            _check(lib().emc_set_refine(self._h, 0, None, None, None))
            return
        qm = np.ascontiguousarray(quat_mapping, dtype='i4')
        ro = np.ascontiguousarray(rel_off, dtype='i8')
        rq = np.ascontiguousarray(rel_quat, dtype='i4')
        if rq.size == 0:
            rq = np.zeros(1, 'i4')
        _check(lib().emc_set_refine(self._h, int(num_coarse), _ptr(qm, _ip), _ptr(ro, _lp), _ptr(rq, _ip)))

    def set_precision(self, mode):
        """0: plain FP64; 1: FP16 log W on matrix cores (dense frames); 2: FP16 hi + lo;
        3: FP32 log W and posterior gathers (any frames); 4, 5: as 1, 2 with BF16 log W and
        counts in the likelihood product (counts <= 256)."""
        _check(lib().emc_set_precision(self._h, int(mode)))

    def set_frame_params(self, scale, beta, blacklist, fcounts, sum_fact, mean_count, num_blacklist=0):
        a = [np.ascontiguousarray(x, dtype=t) for x, t in (
            (scale, 'f8'), (beta, 'f8'), (blacklist, 'u1'), (fcounts, 'i4'), (sum_fact, 'f8'))]
        if any(x.shape[0] != self.num_data for x in a):
            raise ValueError('per-frame arrays must have num_data entries')
        _check(lib().emc_set_frame_params(self._h, _ptr(a[0]), _ptr(a[1]), _ptr(a[2], _u8p),
                                          _ptr(a[3], _ip), _ptr(a[4]), float(mean_count),
                                          int(num_blacklist)))

    def set_beta(self, beta):
        beta = np.ascontiguousarray(beta, dtype='f8')
        if beta.shape[0] != self.num_data:
            raise ValueError('beta must have num_data entries')
        _check(lib().emc_set_beta(self._h, _ptr(beta)))

    def set_scale(self, scale):
        scale = np.ascontiguousarray(scale, dtype='f8')
        if scale.shape[0] != self.num_data:
            raise ValueError('scale must have num_data entries')
        _check(lib().emc_set_scale(self._h, _ptr(scale)))

    def maximize(self, iteration=1, need_scaling=0, update_scale=1, known_scale=0):
        opts = IterOpts(int(iteration), int(need_scaling), int(update_scale), int(known_scale))
        st = IterStats()
        _check(lib().emc_maximize(self._h, C.byref(opts), C.byref(st)))
        return dict(avg_likelihood=st.avg_likelihood, mutual_info=st.mutual_info,
                    rescale=st.rescale, stage_ms=dict(zip(STAGES, list(st.stage_ms))),
                    num_chunks=st.num_chunks, view_nonzero=st.view_nonzero)

    def update_model(self, alpha=0.):
        rms = C.c_double()
        _check(lib().emc_update_model(self._h, float(alpha), C.byref(rms)))
        return rms.value

    def normalize_scale(self):
        m = C.c_double()
        _check(lib().emc_normalize_scale(self._h, C.byref(m)))
        return m.value

    def clear_merge(self):
        _check(lib().emc_clear_merge(self._h))

    def slice_gen(self, quat):
        quat = np.ascontiguousarray(quat, dtype='f8')
        out = np.empty(self.num_pix)
        _check(lib().emc_slice_gen(self._h, _ptr(quat), _ptr(out)))
        return out

    def subset_mass(self, rel_off, rel_quat):
        """Per frame, the posterior mass of the last (plain) iteration on the orientations
        rel_quat[rel_off[d]:rel_off[d + 1]]."""
        rel_off = np.ascontiguousarray(rel_off, dtype='i8')
        rel_quat = np.ascontiguousarray(rel_quat, dtype='i4')
        if rel_off.shape != (self.num_data + 1,):
            raise ValueError('rel_off must have num_data + 1 entries')
        out = np.empty(self.num_data)
        _check(lib().emc_subset_mass(self._h, _ptr(rel_off, _lp), _ptr(rel_quat, _ip) if rel_quat.size else None,
                                     _ptr(out)))
        return out

    def cell_topk(self, num_coarse, quat_mapping, k):
        """Each frame's k most likely coarse cells after an iteration (descending mass) and the
        mass they hold: ((num_data, k) int32, (num_data,) float32)."""
        qm = np.ascontiguousarray(quat_mapping, dtype='i4')
        if qm.shape != (self.num_rot,):
            raise ValueError('quat_mapping must have num_rot entries')
        idx = np.empty((self.num_data, int(k)), dtype='i4')
        mass = np.empty(self.num_data, dtype='f4')
        _check(lib().emc_cell_topk(self._h, int(num_coarse), _ptr(qm, _ip), int(k), _ptr(idx, _ip),
                                   mass.ctypes.data_as(C.POINTER(C.c_float))))
        return idx, mass

    def judge_outside(self, coarse, prob_floor=0.):
        """After an estep_only iteration on refinement subsets (the proposals) and a plain
        estep_only iteration of `coarse` (one orientation per coarse cell, same frames): per
        frame, the estimated posterior mass of the full fine search outside its cells (with
        prob_floor > 0: in cells whose estimated posterior per orientation exceeds it)."""
        out = np.empty(self.num_data)
        _check(lib().emc_judge_outside(self._h, coarse._h, float(prob_floor), _ptr(out)))
        return out

    def copy_model(self, src):
        """This engine's model1 := src's (same device and mode count), copied on the device."""
        _check(lib().emc_copy_model(self._h, src._h))

    def judge_widen(self, coarse, outside_max, max_add, target=None):
        """judge_outside (prob_floor 0) and, for the frames above outside_max, the fewest coarse
        cells to add for the estimate to fall to target (default outside_max / 10: the estimate
        of a widened set reads cells at the coarse sampling only): (outside, num_add, cells) with
        num_add (num_data,) i4 = 0 (kept as proposed), k (k cells), -1 (the full search) and
        cells the added cells frame after frame."""
        out = np.empty(self.num_data)
        num = np.empty(self.num_data, dtype='i4')
        tot = C.c_long(0)
        target = outside_max / 10. if target is None else target
        _check(lib().emc_judge_widen(self._h, coarse._h, float(outside_max), float(target), int(max_add), _ptr(out),
                                     _ptr(num, _ip), C.byref(tot)))
        cells = np.empty(max(tot.value, 1), dtype='i4')
        _check(lib().emc_judge_widen_cells(self._h, _ptr(cells, _ip), cells.size))
        return out, num, cells[:tot.value]

    def slice_merge(self, quat, view):
        quat = np.ascontiguousarray(quat, dtype='f8')
        view = np.ascontiguousarray(view, dtype='f8')
        if view.shape[0] != self.num_pix:
            raise ValueError('view must have num_pix entries')
        _check(lib().emc_slice_merge(self._h, _ptr(quat), _ptr(view)))

    def synchronize(self):
        _check(lib().emc_synchronize(self._h))

    def get(self, name):
        fid, dt, kind = FIELDS[name]
        vol = (self.size,) * 3 if self.num_modes == 1 else (self.num_modes,) + (self.size,) * 3
        shape = {'rot': (self.num_rot,), 'data': (self.num_data,), 'vol': vol,
                 'data_rot': (self.num_data, self.num_rot),
                 'data_mode': (self.num_data, self.num_modes)}[kind]
        out = np.empty(shape, dtype=dt)
        _check(lib().emc_get_field(self._h, fid, out.ctypes.data_as(C.c_void_p), out.size))
        return out

    def sparse_prob(self):
        """Sparse posterior of the last iteration (max_prob.c:295-326): (num_prob, place, prob)
        with num_prob (num_data,) i4 counts of P > 1e-6 and place (i4) / prob (f8) the
        entries frame after frame, orientations ascending."""
        num = np.zeros(self.num_data, dtype='i4')
        tot = C.c_long(0)
        _check(lib().emc_get_sparse_prob(self._h, _ptr(num, _ip), 0, None, None, C.byref(tot)))
        place = np.empty(max(tot.value, 1), dtype='i4')
        prob = np.empty(max(tot.value, 1), dtype='f8')
        _check(lib().emc_get_sparse_prob(self._h, _ptr(num, _ip), place.size, _ptr(place, _ip), _ptr(prob),
                                         C.byref(tot)))
        return num, place[:tot.value], prob[:tot.value]

    def set_option(self, name, value):
        """Kernel-variant option (emc_set_option): lik_tiles, lik_unroll, lik_xcd, tomo_tiles,
        tomo_tile_mode, tomo_parts, tomo_width32, tomo_unroll, tomo_seq, max_chunk_cols,
        merge_mode, subset, slot_pack, lik_split, sparse_rows, full_part_frames, estep_only,
        mstep_only, graph, fused, sig_full_rows."""
        _check(lib().emc_set_option(self._h, name.encode(), int(value)))

    def get_option(self, name):
        v = C.c_int(0)
        _check(lib().emc_get_option(self._h, name.encode(), C.byref(v)))
        return v.value

    def cell_posterior(self, num_coarse, quat_mapping):
        """(num_data, num_coarse) float32: each frame's posterior mass in each coarse cell
        (quat_mapping[r] = cell of orientation r) after the last iteration."""
        qm = np.ascontiguousarray(quat_mapping, dtype='i4')
        if qm.shape != (self.num_rot,):
            raise ValueError('quat_mapping must have num_rot entries')
        out = np.empty((self.num_data, int(num_coarse)), dtype='f4')
        _check(lib().emc_cell_posterior(self._h, int(num_coarse), _ptr(qm, _ip),
                                        out.ctypes.data_as(C.POINTER(C.c_float))))
        return out
