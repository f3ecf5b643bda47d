"""cv_divergence_noalloc (Divergence on device-resident fields, no device
allocation) against cv_divergence on the same fields: the same maximum and
first index, bit for bit, on two calls in a row, the second on a smaller
grid; and its argument checks."""
import ctypes as C

import numpy as np
import pytest

from curvib_old_amd import _lib, cases
from curvib_old_amd.solver import from_case

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _to_device(h, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0      # host to device
    return p


def test_noalloc_divergence_matches_cv_divergence():
    L = _lib.load()
    h = _hip()
    for cells, seed in (((48, 40, 32), 3), ((24, 16, 20), 4)):
        case = cases.ChannelCase(cells=cells, amp=0.1, sphere=(0.5, 0.5, 0.5, 0.4), seed=seed, dt=2e-4, re=300.0)
        u, (xyz, nv, uc) = from_case(case, mglevels=2)
        bufs = []
        try:
            rng = np.random.default_rng(seed)
            uc2 = uc + rng.uniform(-1e-3, 1e-3, uc.shape) * (nv < 0.1)[..., None]
            u.put("ucont", uc2)
            v_ref, i_ref = u.Divergence()
            mx, my, mz = case.storage
            bufs = [_to_device(h, x) for x in (u.get("ucont"), u.get("aj"), u.get("nvert"))]
            out = _to_device(h, np.full(2, -7.0))
            bufs.append(out)
            rc = L.cv_divergence_noalloc(0, bufs[0], bufs[1], bufs[2], mx, my, mz, None, out)
            assert rc == _lib.CV_OK, _lib.last_error()
            assert h.hipDeviceSynchronize() == 0
            res = np.zeros(2)
            assert h.hipMemcpy(res.ctypes.data_as(C.c_void_p), out, 16, 2) == 0      # device to host
            print(f"noalloc divergence {cells}: {res[0]:.6e} at {int(res[1])}, cv_divergence {v_ref:.6e} at {i_ref}",
                  flush=True)
            assert res[0] == v_ref and int(res[1]) == i_ref, (cells, res, v_ref, i_ref)
        finally:
            for p in bufs:
                h.hipFree(p)
            u.close()


def test_noalloc_divergence_refuses_bad_arguments():
    L = _lib.load()
    fake = C.c_void_p(16)     # never dereferenced: the checks come first
    assert L.cv_divergence_noalloc(0, None, fake, fake, 8, 8, 8, None, fake) == _lib.CV_ERR_ARG
    assert L.cv_divergence_noalloc(0, fake, fake, fake, 2, 8, 8, None, fake) == _lib.CV_ERR_ARG
    assert L.cv_divergence_noalloc(0, fake, fake, fake, 2048, 1024, 1024, None, fake) == _lib.CV_ERR_ARG
    assert L.cv_divergence_noalloc(-1, fake, fake, fake, 8, 8, 8, None, fake) == _lib.CV_ERR_ARG
