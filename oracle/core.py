"""ctypes binding of the oracle's C restatement (liboracle.so) -- test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_u8p = C.POINTER(C.c_uint8)


class Problem(C.Structure):
    _fields_ = [
        ('num_pix', C.c_int), ('qvals', _dp), ('corr', _dp), ('mask', _u8p),
        ('num_rot', C.c_int), ('quats', _dp),
        ('size', C.c_long), ('model1', _dp), ('model2', _dp), ('inter_weight', _dp),
        ('num_data', C.c_int), ('ftype', C.c_int), ('ones', _ip), ('multi', _ip),
        ('ones_accum', _lp), ('multi_accum', _lp), ('place_ones', _ip), ('place_multi', _ip),
        ('count_multi', _ip), ('int_frames', _ip),
        ('beta', _dp), ('scale', _dp), ('blacklist', _u8p), ('fcounts', _ip), ('sum_fact', _dp),
        ('mean_count', C.c_double), ('num_blacklist', C.c_int),
        ('quat_mapping', _ip), ('rel_off', _lp), ('rel_quat', _ip),
        ('need_scaling', C.c_int), ('update_scale', C.c_int), ('iteration', C.c_int),
        ('known_scale', C.c_int), ('num_threads', C.c_int), ('num_modes', C.c_int),
        ('with_bg', C.c_int), ('background', _dp), ('powder', _dp), ('bgscale', _dp),
    ]


class Result(C.Structure):
    _fields_ = [
        ('u', _dp), ('max_exp', _dp), ('p_norm', _dp), ('rmax', _ip), ('likelihood', _dp),
        ('info', _dp), ('psum_d', _dp), ('prob_dense', _dp),
        ('rescale', C.c_double), ('mutual_info', C.c_double), ('avg_likelihood', C.c_double),
        ('seconds', C.c_double * 7), ('quat_norm', _dp),
    ]


_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE, 'oracle'], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_num_rot.argtypes = [C.c_int]
        L.orc_quat_gen.argtypes = [C.c_int, _dp]
        L.orc_make_rot_quat.argtypes = [_dp, _dp]
        L.orc_slice_gen3d.argtypes = [_dp, _dp, C.c_long, C.c_int, _dp, _dp, _dp]
        L.orc_slice_merge3d.argtypes = [_dp, _dp, _dp, _dp, C.c_long, C.c_int, _dp, _dp, _u8p]
        L.orc_symmetrize_friedel.argtypes = [_dp, _dp, C.c_int]
        L.orc_maximize.argtypes = [C.POINTER(Problem), C.POINTER(Result)]
        L.orc_maximize.restype = C.c_double
        L.orc_update_model.argtypes = [_dp, _dp, _dp, C.c_long, C.c_double]
        L.orc_update_model.restype = C.c_double
        L.orc_update_model_modes.argtypes = [_dp, _dp, _dp, C.c_long, C.c_double, C.c_int]
        L.orc_update_model_modes.restype = C.c_double
        L.orc_voronoi_subset.argtypes = [C.c_int, _dp, C.c_int, _dp, _ip]
        L.orc_tomogram_bg_view.argtypes = [C.POINTER(Problem), _dp, C.c_double, C.c_double, _dp]
        L.orc_calc_powder.argtypes = [C.c_int, _u8p, C.c_int, _ip, _ip, _lp, _lp, _ip, _ip, _ip, _dp]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def quat_gen(num_div):
    n = lib().orc_num_rot(num_div)
    q = np.zeros((n, 5))
    got = lib().orc_quat_gen(num_div, _p(q))
    if got != n:
        raise RuntimeError('orc_quat_gen failed: %d' % got)
    return q


def make_rot_quat(q):
    q = np.ascontiguousarray(q, dtype='f8')
    rot = np.zeros(9)
    lib().orc_make_rot_quat(_p(q), _p(rot))
    return rot.reshape(3, 3)


def slice_gen3d(quat, model, qvals, corr):
    quat = np.ascontiguousarray(quat, dtype='f8')
    model = np.ascontiguousarray(model, dtype='f8')
    qvals = np.ascontiguousarray(qvals, dtype='f8')
    corr = np.ascontiguousarray(corr, dtype='f8')
    size = model.shape[0]
    out = np.empty(len(corr))
    lib().orc_slice_gen3d(_p(quat), _p(model), size, len(corr), _p(qvals), _p(corr), _p(out))
    return out


def slice_merge3d(quat, view, model2, weight, qvals, corr, mask):
    """Accumulate into model2/weight in place (both C-contiguous float64 cubes)."""
    quat = np.ascontiguousarray(quat, dtype='f8')
    view = np.ascontiguousarray(view, dtype='f8')
    qvals = np.ascontiguousarray(qvals, dtype='f8')
    corr = np.ascontiguousarray(corr, dtype='f8')
    mask = np.ascontiguousarray(mask, dtype='u1')
    lib().orc_slice_merge3d(_p(quat), _p(view), _p(model2), _p(weight), model2.shape[0],
                            len(corr), _p(qvals), _p(corr), _p(mask, _u8p))


def symmetrize_friedel(array, weights):
    lib().orc_symmetrize_friedel(_p(array), _p(weights), array.shape[0])


def update_model(model1, model2, inter_weight, alpha=0.):
    if model1.ndim == 4:
        return lib().orc_update_model_modes(_p(model1), _p(model2), _p(inter_weight), model1.shape[1],
                                            float(alpha), model1.shape[0])
    return lib().orc_update_model(_p(model1), _p(model2), _p(inter_weight), model1.shape[0],
                                  float(alpha))


def voronoi_subset(qcoarse, qfine):
    """Nearest coarse orientation of every fine orientation (quaternion.c:844-862)."""
    qc = np.ascontiguousarray(qcoarse, dtype='f8')
    qf = np.ascontiguousarray(qfine, dtype='f8')
    out = np.zeros(len(qf), dtype='i4')
    lib().orc_voronoi_subset(len(qc), _p(qc), len(qf), _p(qf), _p(out, _ip))
    return out


def calc_powder(mask, frames):
    """calc_powder (iterate.c:119-179) for sparse frames: mean photons per mask < 1 pixel."""
    mask = np.ascontiguousarray(mask, dtype='u1')
    a = [np.ascontiguousarray(frames[k], dtype=t) for k, t in (
        ('ones', 'i4'), ('multi', 'i4'), ('ones_accum', 'i8'), ('multi_accum', 'i8'),
        ('place_ones', 'i4'), ('place_multi', 'i4'), ('count_multi', 'i4'))]
    out = np.zeros(len(mask))
    lib().orc_calc_powder(len(mask), _p(mask, _u8p), len(a[0]), _p(a[0], _ip), _p(a[1], _ip), _p(a[2], _lp),
                          _p(a[3], _lp), _p(a[4], _ip), _p(a[5], _ip), _p(a[6], _ip), _p(out))
    return out


def maximize(det, quats, model1, frames, fparams, flags, num_threads=1, dense_prob=False,
             refine=None, background=None, bg_view=None):
    """Run one oracle iteration.

    det      : dict qvals (N,3) f8, corr (N,) f8, mask (N,) u1
    quats    : (R,5) f8
    model1   : (S,S,S) f8, or (M,S,S,S) for M modes: orientations are then the combined
               r = rotation * M + mode and u, prob, rmax count R * M of them; the result holds
               quat_norm (D, M)
    frames   : dict with ftype 'sparse' keys ones, multi, ones_accum, multi_accum, place_ones,
               place_multi, count_multi, or 'dense' key int_frames (D, N) i4
    fparams  : dict beta, scale, blacklist, fcounts, sum_fact, mean_count
    flags    : dict need_scaling, update_scale, iteration, known_scale
    refine   : None, or dict quat_mapping (R,) i4, rel_off (D+1,) i8, rel_quat i4 (frame d is
               evaluated only where quat_mapping[r] is in rel_quat[rel_off[d]:rel_off[d+1]])
    background : None, or dict background (N,) f8 and bgscale (D,) f8 (a detector with a
               background; powder from calc_powder); with need_scaling the likelihood, the
               tomogram and the scale update take it into account
    bg_view  : None, or dict pr (D,), psum_r, scalemax: instead of an iteration, return the
               view update_tomogram_bg finds for one orientation with these inputs
    Returns a dict of outputs; fparams['scale'] is updated in place when scaling is on.
    """
    qvals = np.ascontiguousarray(det['qvals'], dtype='f8')
    corr = np.ascontiguousarray(det['corr'], dtype='f8')
    mask = np.ascontiguousarray(det['mask'], dtype='u1')
    quats = np.ascontiguousarray(quats, dtype='f8')
    model1 = np.ascontiguousarray(model1, dtype='f8')
    M = model1.shape[0] if model1.ndim == 4 else 1
    S = model1.shape[-1]
    model2 = np.zeros_like(model1)
    inter_weight = np.zeros_like(model1)
    dense = 'int_frames' in frames
    D = frames['int_frames'].shape[0] if dense else len(frames['ones'])
    R = quats.shape[0] * M
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a

    prob = Problem()
    prob.num_pix = len(corr)
    prob.qvals, prob.corr, prob.mask = _p(qvals), _p(corr), _p(mask, _u8p)
    prob.num_rot, prob.quats, prob.num_modes = quats.shape[0], _p(quats), M
    prob.size, prob.model1, prob.model2, prob.inter_weight = S, _p(model1), _p(model2), _p(inter_weight)
    prob.num_data = D
    if dense:
        prob.ftype = 1
        prob.int_frames = _p(arr(frames['int_frames'], 'i4'), _ip)
    else:
        prob.ftype = 0
        prob.ones = _p(arr(frames['ones'], 'i4'), _ip)
        prob.multi = _p(arr(frames['multi'], 'i4'), _ip)
        prob.ones_accum = _p(arr(frames['ones_accum'], 'i8'), _lp)
        prob.multi_accum = _p(arr(frames['multi_accum'], 'i8'), _lp)
        prob.place_ones = _p(arr(frames['place_ones'], 'i4'), _ip)
        prob.place_multi = _p(arr(frames['place_multi'], 'i4'), _ip)
        prob.count_multi = _p(arr(frames['count_multi'], 'i4'), _ip)
    scale = fparams['scale']
    if scale.dtype != np.float64 or not scale.flags.c_contiguous:
        raise TypeError('scale must be contiguous float64 (updated in place)')
    blacklist = arr(fparams['blacklist'], 'u1')
    prob.beta = _p(arr(fparams['beta'], 'f8'))
    prob.scale = _p(scale)
    prob.blacklist = _p(blacklist, _u8p)
    prob.fcounts = _p(arr(fparams['fcounts'], 'i4'), _ip)
    prob.sum_fact = _p(arr(fparams['sum_fact'], 'f8'))
    prob.mean_count = float(fparams['mean_count'])
    prob.num_blacklist = int(fparams.get('num_blacklist', 0))
    prob.need_scaling = int(flags.get('need_scaling', 0))
    prob.update_scale = int(flags.get('update_scale', 1))
    prob.iteration = int(flags.get('iteration', 1))
    prob.known_scale = int(flags.get('known_scale', 0))
    prob.num_threads = int(num_threads)
    if refine is not None:
        prob.quat_mapping = _p(arr(refine['quat_mapping'], 'i4'), _ip)
        prob.rel_off = _p(arr(refine['rel_off'], 'i8'), _lp)
        prob.rel_quat = _p(arr(refine['rel_quat'], 'i4'), _ip)

    if background is not None:
        prob.with_bg = 1
        prob.background = _p(arr(background['background'], 'f8'))
        prob.bgscale = _p(arr(background['bgscale'], 'f8'))
        prob.powder = _p(arr(calc_powder(mask, frames), 'f8'))
    if bg_view is not None:
        view = np.zeros(len(corr))
        lib().orc_tomogram_bg_view(C.byref(prob), _p(arr(bg_view['pr'], 'f8')), float(bg_view['psum_r']),
                                   float(bg_view['scalemax']), _p(view))
        return view

    out = dict(u=np.zeros(R), max_exp=np.zeros(D), p_norm=np.zeros(D), rmax=np.zeros(D, 'i4'),
               likelihood=np.zeros(D), info=np.zeros(D), psum_d=np.zeros(D))
    res = Result()
    res.u, res.max_exp, res.p_norm = _p(out['u']), _p(out['max_exp']), _p(out['p_norm'])
    res.rmax = _p(out['rmax'], _ip)
    res.likelihood, res.info, res.psum_d = _p(out['likelihood']), _p(out['info']), _p(out['psum_d'])
    if dense_prob:
        out['prob'] = np.zeros((D, R))
        res.prob_dense = _p(out['prob'])
    if M > 1:
        out['quat_norm'] = np.zeros((D, M))
        res.quat_norm = _p(out['quat_norm'])
    avg = lib().orc_maximize(C.byref(prob), C.byref(res))
    out.update(avg_likelihood=avg, rescale=res.rescale, mutual_info=res.mutual_info,
               model2=model2, inter_weight=inter_weight, seconds=list(res.seconds))
    return out
