"""emc_slice_gen_static (the entry point that allocates no device memory) against
emc_slice_gen through the C ABI: bit-identical views on two calls in a row, the second on a
smaller detector and model, and a clear error for a detector above the staging maximum."""
import ctypes as C
import re

import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _static_max_pix():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, 'include', 'emc_hip.h')).read()
    return int(re.search(r'#define EMC_STATIC_MAX_PIX (\d+)', text).group(1))


def _static(eng, quat):
    from dragonfly_amd import _lib
    quat = np.ascontiguousarray(quat, dtype='f8')
    out = np.full(eng.num_pix, np.nan)
    rc = _lib.lib().emc_slice_gen_static(eng._h, quat.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return rc, out


def _problem(size, nx, ny):
    from dragonfly_amd import _lib, synth
    det = synth.flat_detector(nx, ny, qmax=(size - 3) / 2. - 0.05, stoprad=2., gap=0)
    rng = np.random.default_rng(size)
    model = rng.random((size,) * 3) + 0.1
    return det, model, _lib.quat_gen(2)


def test_static_slice_gen_matches_slice_gen_then_smaller():
    from dragonfly_amd._lib import Engine
    from oracle import core
    for size, nx, ny, rows in ((65, 160, 150, (0, 7, 301)), (21, 22, 20, (3, 419))):
        print('case: emc_slice_gen_static, %d^3 model, %dx%d detector' % (size, nx, ny), flush=True)
        det, model, quats = _problem(size, nx, ny)
        eng = Engine(0)
        eng.set_detector(det['qvals'], det['corr'], det['mask'])
        eng.set_model(model)
        for r in rows:
            rc, g = _static(eng, quats[r])
            assert rc == 0, r
            ref = eng.slice_gen(quats[r])
            assert np.array_equal(g.view('u8'), ref.view('u8')), 'static slice differs at r=%d' % r
            o = core.slice_gen3d(quats[r], model, det['qvals'], det['corr'])
            assert np.array_equal(g.view('u8'), o.view('u8')), 'static slice differs from the oracle at r=%d' % r
        eng.close()


def test_static_slice_gen_rejects_oversize_detector():
    from dragonfly_amd import _lib
    from dragonfly_amd._lib import Engine
    n = _static_max_pix() + 1
    print('case: emc_slice_gen_static with %d pixels' % n, flush=True)
    rng = np.random.default_rng(1)
    qvals = rng.uniform(-3., 3., (n, 3))
    eng = Engine(0)
    eng.set_detector(qvals, np.ones(n), np.zeros(n, 'u1'))
    eng.set_model(np.ones((9, 9, 9)))
    rc, _ = _static(eng, np.array([1., 0., 0., 0., 1.]))
    assert rc == -1
    assert 'maximum' in _lib.lib().emc_last_error().decode()
    # the allocating entry point still serves this detector
    assert np.isfinite(eng.slice_gen(np.array([1., 0., 0., 0., 1.]))).all()
    eng.close()
