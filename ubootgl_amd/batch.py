"""Batched solves over libubg.so (ubg_batch_*, ubg_sims_step).

Refinement solves: n independent Poisson problems (mgtest form: fluid outer
ring held as Dirichlet data, residual operator of calculateResidualField,
pressure_solver.cpp:91-116) of uneven sizes, each solved by FP64 iterative
refinement around FP32 V-cycles exactly as MG.ir_solve does for one domain.

Projections: the simulation's own pressure solve (MG::solve with the
zero-gradient BC, pressure_solver.cpp:250-257) on n domains, and whole
Simulation::step calls of n simulations with their solves batched.

Every multigrid stage of all domains is one launch.  No counterpart in the
reference, which steps one Simulation at a time.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, fptr, lib

TILE_SHAPES = [(32, 8), (64, 4), (64, 8), (64, 16), (128, 4), (128, 8), (256, 4)]


def batch_bytes(sizes):
    """Device bytes a batch of (W, H) domains allocates (host-side rule)."""
    ws = (C.c_int * len(sizes))(*[int(w) for w, _ in sizes])
    hs = (C.c_int * len(sizes))(*[int(h) for _, h in sizes])
    out = C.c_size_t()
    check(lib().ubg_batch_bytes(len(sizes), ws, hs, C.byref(out)))
    return out.value


class Batch:
    def __init__(self, sizes, params=None):
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        n = len(self.sizes)
        ws = (C.c_int * n)(*[w for w, _ in self.sizes])
        hs = (C.c_int * n)(*[h for _, h in self.sizes])
        h = C.c_void_p()
        check(lib().ubg_batch_create(n, ws, hs, C.byref(params._p) if params else None,
                                     C.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.sizes)

    def set_domain(self, d, flag, rhs, h):
        flag = np.ascontiguousarray(flag, np.float32)
        rhs = np.ascontiguousarray(rhs, np.float32)
        W, H = self.sizes[d]
        assert flag.shape == (H, W) and rhs.shape == (H, W)
        check(lib().ubg_batch_set_domain(self._h, d, fptr(flag), fptr(rhs), float(h)))

    def set_problem(self, d, P):
        self.set_domain(d, P.flag, P.f, P.h)

    def set_rhs(self, d, rhs):
        check(lib().ubg_batch_set_rhs(self._h, d, fptr(np.ascontiguousarray(rhs, np.float32))))

    def set_guess(self, d, x=None):
        if x is None:
            check(lib().ubg_batch_set_guess(self._h, d, None))
            return
        x = np.ascontiguousarray(x, np.float64)
        check(lib().ubg_batch_set_guess(self._h, d, x.ctypes.data_as(C.POINTER(C.c_double))))

    def solve(self, tol=1e-6, vcycles_per_iter=4, max_iter=200, raise_on_stall=True):
        """Solve every domain; returns the number of two-step graph replays."""
        prm = L.IRParams(tol, vcycles_per_iter, max_iter)
        reps = C.c_int()
        rc = lib().ubg_batch_solve(self._h, C.byref(prm), C.byref(reps))
        if rc != L.E_NOT_CONVERGED or raise_on_stall:
            check(rc)
        return reps.value

    def result(self, d):
        res = L.IRResult()
        check(lib().ubg_batch_result(self._h, d, C.byref(res)))
        n = min(res.iterations, L.IR_HISTORY)
        return dict(iterations=res.iterations, vcycles=res.vcycles,
                    converged=bool(res.converged), r0_norm=res.r0_norm,
                    rel_residual=res.rel_residual, history=list(res.history[:n]))

    def solution(self, d):
        W, H = self.sizes[d]
        out = np.zeros((H, W), np.float64)
        check(lib().ubg_batch_get_solution(self._h, d, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def autotune(self, db_path=None):
        t = L.BatchTuning()
        check(lib().ubg_batch_autotune(self._h, db_path.encode() if db_path else None,
                                       C.byref(t)))
        return dict(tile_w=t.tile_w, tile_h=t.tile_h, blocks=t.blocks, pass_rows=t.pass_rows,
                    sweep_ms=t.sweep_ms)

    def set_tuning(self, tile_w, tile_h, blocks=0, pass_rows=0):
        """tile_w x tile_h: smoother tile (used when a smoothing count exceeds
        3); pass_rows: rows of the fused level-pass tiles (8, 16 or 32; 0 =
        chosen per level from its tile count)."""
        t = L.BatchTuning(tile_w, tile_h, blocks, pass_rows, 0.0)
        check(lib().ubg_batch_set_tuning(self._h, C.byref(t)))

    def tuning(self):
        t = L.BatchTuning()
        check(lib().ubg_batch_get_tuning(self._h, C.byref(t)))
        return dict(tile_w=t.tile_w, tile_h=t.tile_h, blocks=t.blocks, pass_rows=t.pass_rows)

    # ---- projections (the simulation's pressure solve) ----
    def set_projection_domain(self, d, flag, rhs, h):
        """Domain d for projections: any 0/1 flag (walls included)."""
        flag = np.ascontiguousarray(flag, np.float32)
        rhs = np.ascontiguousarray(rhs, np.float32)
        W, H = self.sizes[d]
        assert flag.shape == (H, W) and rhs.shape == (H, W)
        check(lib().ubg_batch_set_projection_domain(self._h, d, fptr(flag), fptr(rhs), float(h)))

    def set_pressure(self, d, p=None):
        if p is None:
            check(lib().ubg_batch_set_pressure(self._h, d, None))
            return
        p = np.ascontiguousarray(p, np.float32)
        W, H = self.sizes[d]
        assert p.shape == (H, W)
        check(lib().ubg_batch_set_pressure(self._h, d, fptr(p)))

    def pressure(self, d):
        W, H = self.sizes[d]
        out = np.zeros((H, W), np.float32)
        check(lib().ubg_batch_get_pressure(self._h, d, fptr(out)))
        return out

    def vcycles(self, ncycles, zero_gradient_bc=True):
        """ncycles V-cycles (MG::solve) of every domain, batched."""
        check(lib().ubg_batch_vcycles(self._h, int(ncycles), 1 if zero_gradient_bc else 0))

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().ubg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SimBatch(Batch):
    """The pressure solves of several Simulations, stepped together:
    step(dts) is one Simulation.step of each, bit-identical to stepping them
    one by one, with one launch per multigrid stage for all of them."""

    def __init__(self, sims):
        self.sims = list(sims)
        n = len(self.sims)
        hs = (C.c_void_p * n)(*[s._h.value for s in self.sims])
        h = C.c_void_p()
        check(lib().ubg_batch_create_for_sims(hs, n, C.byref(h)))
        self._h = h
        self.sizes = [(s.width, s.height) for s in self.sims]

    def step(self, dts):
        n = len(self.sims)
        if np.isscalar(dts):
            dts = [dts] * n
        d = np.ascontiguousarray(dts, np.float32)
        assert d.shape == (n,)
        hs = (C.c_void_p * n)(*[s._h.value for s in self.sims])
        check(lib().ubg_sims_step(self._h, hs, n, fptr(d)))
