"""Served refinement solves over libubg.so (ubg_server_*).

A request is `timesteps` consecutive solves on one registered geometry, each
to `tol` relative to ||F f|| of its rhs and warm-started from the previous
one; the C++ worker thread admits requests by HBM budget and advances all
admitted requests together (continuous batching).  No counterpart in the
reference.
"""
import ctypes as C
import threading

import numpy as np

from . import _lib as L
from ._lib import check, fptr, lib

STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class Server:
    def __init__(self, hbm_budget=200 * 10**9, max_active=1024, tol=1e-6, vcycles_per_iter=4,
                 max_iter=200, params=None, steps_per_turn=1):
        prm = L.ServerParams()
        lib().ubg_server_params_default(C.byref(prm))
        prm.hbm_budget = int(hbm_budget)
        prm.max_active = int(max_active)
        prm.tol = float(tol)
        prm.vcycles_per_iter = int(vcycles_per_iter)
        prm.max_iter = int(max_iter)
        prm.steps_per_turn = int(steps_per_turn)
        h = C.c_void_p()
        check(lib().ubg_server_create(C.byref(prm), C.byref(params._p) if params else None,
                                      C.byref(h)))
        self._h = h
        self._geoms = {}
        self._keep = {}  # request id -> arrays the server reads / writes
        self._lock = threading.Lock()

    def add_geometry(self, flag, h):
        flag = np.ascontiguousarray(flag, np.float32)
        H, W = flag.shape
        gid = C.c_int()
        check(lib().ubg_server_add_geometry(self._h, W, H, fptr(flag), float(h), C.byref(gid)))
        self._geoms[gid.value] = (W, H)
        return gid.value

    def submit(self, gid, rhs_list, fnorms=None, want_x=False):
        if gid not in self._geoms:
            raise L.UbgError(L.E_INVALID, "bad geometry id")
        W, H = self._geoms[gid]
        arrs = [np.ascontiguousarray(r, np.float32) for r in rhs_list]
        for a in arrs:
            assert a.shape == (H, W)
        ptrs = (L._fp * len(arrs))(*[fptr(a) for a in arrs])
        fn = None
        if fnorms is not None:
            fn = np.ascontiguousarray(fnorms, np.float64)
        x = np.zeros((H, W), np.float64) if want_x else None
        rid = C.c_uint64()
        dp = C.POINTER(C.c_double)
        check(lib().ubg_server_submit(self._h, gid, len(arrs), ptrs,
                                      fn.ctypes.data_as(dp) if fn is not None else None,
                                      x.ctypes.data_as(dp) if x is not None else None,
                                      C.byref(rid)))
        with self._lock:
            self._keep[rid.value] = (arrs, ptrs, fn, x)
        return rid.value

    def wait(self, rid, timeout=-1.0):
        info = L.RequestInfo()
        rc = lib().ubg_server_wait(self._h, rid, float(timeout), C.byref(info))
        out = dict(id=info.id, state=STATES[info.state], timesteps_done=info.timesteps_done,
                   iterations=info.iterations, converged=bool(info.converged),
                   rel_residual=info.rel_residual, t_submit=info.t_submit,
                   t_admit=info.t_admit, t_done=info.t_done)
        if info.state >= 2:
            with self._lock:
                kept = self._keep.pop(rid, None)
            if kept is not None and kept[3] is not None:
                out["x"] = kept[3]
        if rc != L.E_STATE:
            check(rc)
        return out

    def drain(self):
        check(lib().ubg_server_drain(self._h))

    def stats(self):
        st = L.ServerStats()
        check(lib().ubg_server_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in L.ServerStats._fields_}

    def clock(self):
        return lib().ubg_server_clock(self._h)

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().ubg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
