"""The allocation-free solve (ubg_mg_solve_device_noalloc, include/ubg_c.h)
against the allocating entry point it mirrors, ubg_mg_solve_host on a new MG
of the same flag: bit-identical pressure on two calls in a row, the second on
a smaller grid (the shared static scratch keeps the first call's contents),
and sizes beyond the static array rejected before any device work.

The caller-owned device arrays are allocated here, by the test acting as
the caller, with the HIP runtime libubg.so itself links (libamdhip64); the
solve runs in libubg.so.
"""
import ctypes as C

import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import _lib as L
from ubootgl_amd import workloads

pytestmark = pytest.mark.gpu


_hip = None


def hip():
    global _hip
    if _hip is None:
        L.lib()  # loads the runtime libubg.so links
        _hip = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]),
                           ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipDeviceSynchronize", [])):
            getattr(_hip, name).argtypes = args
            getattr(_hip, name).restype = C.c_int
    return _hip


class DeviceArray:
    """A caller-owned device copy of a numpy array."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), self.a.nbytes) == 0
        assert hip().hipMemcpy(self.ptr, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        hip().hipFree(self.ptr)


def _noalloc(p, f, flag, h, zg, ncycles, params=None):
    dp = DeviceArray(np.asarray(p, np.float32))
    df = DeviceArray(np.asarray(f, np.float32))
    dfl = DeviceArray(np.asarray(flag, np.uint8))
    H, W = p.shape
    prm = C.byref(params._p) if params is not None else None
    rc = L.lib().ubg_mg_solve_device_noalloc(dp.ptr, df.ptr, dfl.ptr, W, H, float(h), int(zg),
                                             int(ncycles), prm, None)
    L.check(rc)
    assert hip().hipDeviceSynchronize() == 0
    out = dp.get()
    for d in (dp, df, dfl):
        d.free()
    return out


def _existing(p, f, flag, h, zg, ncycles):
    mg = U.MG(flag)
    out = np.array(p, np.float32, copy=True)
    for _ in range(ncycles):
        mg.solve(out, f, flag, h, zeroGradientBC=bool(zg))
    mg.close()
    return out


def _problem(W, H, seed):
    d = workloads.small_domain(W, H, seed=seed, ndisks=4)
    g = np.random.default_rng(seed)
    p = (0.1 * g.standard_normal((H, W))).astype(np.float32) * d.flag
    f = g.standard_normal((H, W)).astype(np.float32) * d.flag
    h = float(np.float32(1.0 / W))
    return p, f, d.flag, h


def test_noalloc_matches_solve_host_twice_second_smaller(gpu):
    for (W, H, zg, nc, seed) in ((1024, 600, 1, 2, 31), (257, 129, 0, 1, 32)):
        p, f, flag, h = _problem(W, H, seed)
        got = _noalloc(p, f, flag, h, zg, nc)
        ref = _existing(p, f, flag, h, zg, nc)
        np.testing.assert_array_equal(got, ref, err_msg=f"{W}x{H} zg={zg} cycles={nc}")
        assert np.abs(got - p).max() > 0  # the solve did something


def test_noalloc_rejects_sizes_beyond_scratch(gpu):
    W, H = 4096, 2048
    rc = L.lib().ubg_mg_solve_device_noalloc(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), W, H,
                                             1.0, 0, 1, None, None)
    assert rc == L.E_BUDGET
    assert b"static array" in L.lib().ubg_last_error()
