"""ctypes binding of the reference's own quaternion.c + model.c -- test infrastructure only.

``_ref/libdfref.so`` is compiled (by ``oracle/Makefile`` target ``ref``) from the reference
sources where they lie, unchanged. It is used to pin the oracle's orientation sampling and
geometry bit for bit and to write the committed golden fixtures. The struct layouts below
restate the reference headers dragonfly/src/quaternion.h:11-15, detector.h:6-16 and
model.h:15-20 so the compiled functions can be called.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_ref', 'libdfref.so')

_dp = C.POINTER(C.c_double)


class CQuat(C.Structure):
    _fields_ = [('num_div', C.c_int), ('num_rot', C.c_int), ('num_rot_p', C.c_int),
                ('quats', _dp), ('reduced', C.c_int), ('icosahedral_flag', C.c_int),
                ('octahedral_flag', C.c_int)]


class CDet(C.Structure):
    _fields_ = [('fname', C.c_char_p), ('num_pix', C.c_int), ('qvals', _dp), ('corr', _dp),
                ('raw_mask', C.POINTER(C.c_uint8)), ('with_bg', C.c_uint8),
                ('background', _dp), ('powder', _dp), ('detd', C.c_double),
                ('ewald_rad', C.c_double)]


class CModel(C.Structure):
    _fields_ = [('mtype', C.c_int), ('size', C.c_long), ('center', C.c_long), ('vol', C.c_long),
                ('num_modes', C.c_int), ('model1', _dp), ('model2', _dp),
                ('inter_weight', _dp)]


_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.quat_gen.argtypes = [C.c_int, C.POINTER(CQuat)]
        L.make_rot_quat.argtypes = [_dp, _dp]
        L.slice_gen3d.argtypes = [_dp, C.c_int, _dp, C.POINTER(CDet), C.POINTER(CModel)]
        L.slice_merge3d.argtypes = [_dp, C.c_int, _dp, _dp, _dp, C.c_long, C.POINTER(CDet)]
        L.symmetrize_friedel.argtypes = [_dp, _dp, C.c_int]
        L.voronoi_subset.argtypes = [C.POINTER(CQuat), C.POINTER(CQuat), C.POINTER(C.c_int)]
        _lib = L
    return _lib


def quat_gen(num_div):
    q = CQuat()
    q.num_div = num_div
    n = lib().quat_gen(num_div, C.byref(q))
    return np.ctypeslib.as_array(q.quats, shape=(n * 5,)).copy().reshape(n, 5)


def voronoi_subset(qcoarse, qfine):
    """voronoi_subset (quaternion.c:844-862) on the given quaternion arrays."""
    qc = np.ascontiguousarray(qcoarse, dtype='f8')
    qf = np.ascontiguousarray(qfine, dtype='f8')
    a, b = CQuat(), CQuat()
    a.num_rot, a.quats = len(qc), qc.ctypes.data_as(_dp)
    b.num_rot, b.quats = len(qf), qf.ctypes.data_as(_dp)
    out = np.zeros(len(qf), dtype='i4')
    lib().voronoi_subset(C.byref(a), C.byref(b), out.ctypes.data_as(C.POINTER(C.c_int)))
    return out


def make_rot_quat(quat):
    quat = np.ascontiguousarray(quat, dtype='f8')
    rot = np.zeros(9)
    lib().make_rot_quat(quat.ctypes.data_as(_dp), rot.ctypes.data_as(_dp))
    return rot.reshape(3, 3)


def _det(qvals, corr, mask, keep):
    qvals = np.ascontiguousarray(qvals, dtype='f8')
    corr = np.ascontiguousarray(corr, dtype='f8')
    mask = np.ascontiguousarray(mask, dtype='u1')
    keep += [qvals, corr, mask]
    d = CDet()
    d.num_pix = len(corr)
    d.qvals = qvals.ctypes.data_as(_dp)
    d.corr = corr.ctypes.data_as(_dp)
    d.raw_mask = mask.ctypes.data_as(C.POINTER(C.c_uint8))
    return d


def slice_gen3d(quat, model, qvals, corr, mask=None):
    keep = []
    if mask is None:
        mask = np.zeros(len(corr), 'u1')
    det = _det(qvals, corr, mask, keep)
    model = np.ascontiguousarray(model, dtype='f8')
    m = CModel()
    m.size = model.shape[0]
    m.center = m.size // 2
    m.vol = m.size ** 3
    m.num_modes = 1
    m.model1 = model.ctypes.data_as(_dp)
    quat = np.ascontiguousarray(quat, dtype='f8')
    out = np.empty(len(corr))
    lib().slice_gen3d(quat.ctypes.data_as(_dp), 0, out.ctypes.data_as(_dp), C.byref(det), C.byref(m))
    return out


def slice_merge3d(quat, view, model2, weight, qvals, corr, mask):
    keep = []
    det = _det(qvals, corr, mask, keep)
    quat = np.ascontiguousarray(quat, dtype='f8')
    view = np.ascontiguousarray(view, dtype='f8')
    lib().slice_merge3d(quat.ctypes.data_as(_dp), 0, view.ctypes.data_as(_dp),
                        model2.ctypes.data_as(_dp), weight.ctypes.data_as(_dp),
                        model2.shape[0], C.byref(det))


def symmetrize_friedel(array, weights):
    lib().symmetrize_friedel(array.ctypes.data_as(_dp), weights.ctypes.data_as(_dp), array.shape[0])
