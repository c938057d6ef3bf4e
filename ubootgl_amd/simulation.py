"""Python mirror of the reference's Simulation / MG interface over libubg.so.

Names and argument meaning follow simulation.hpp:18-141 and
pressure_solver.hpp:3-76; fields live on the GPU and are exchanged with host
numpy arrays explicitly (`get(name)` / `set(name, array)`), which is the one
behavioural difference from the reference's public std::vector members.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, fptr, lib

_FIELD_IDS = {n: i for i, n in enumerate(L.FIELD_NAMES)}


def _fid(name):
    if isinstance(name, int):
        return name
    return _FIELD_IDS[name]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def level_sizes(W, H):
    """MG level sizes: halve while both sizes exceed 3 (pressure_solver.hpp:16-31)."""
    ws = (C.c_int * 64)()
    hs = (C.c_int * 64)()
    n = lib().ubg_mg_level_sizes(W, H, ws, hs, 64)
    return [(ws[i], hs[i]) for i in range(n)]


class Params:
    """Solver parameters; defaults are the reference's (3/3/5 sweeps, 2 V-cycles, alpha 1)."""

    def __init__(self, **kw):
        self._p = L.default_params()
        for k, v in kw.items():
            if not hasattr(self._p, k):
                raise AttributeError(k)
            setattr(self._p, k, v)

    def __getattr__(self, k):
        return getattr(self._p, k)


class MG:
    """pressure_solver.hpp:13-76.  MG(width, height) or MG(flag) (which also
    runs updateFields); solve() takes host grids like the reference."""

    def __init__(self, width_or_flag, height=None, _handle=None, _owner=None):
        self._owner = _owner
        self.sizes = None
        if _handle is not None:
            self._h = _handle
            self._borrowed = True
        else:
            self._borrowed = False
            if height is None:
                flag = _f32(width_or_flag)
                H, W = flag.shape
            else:
                W, H, flag = int(width_or_flag), int(height), None
            h = C.c_void_p()
            check(lib().ubg_mg_create(W, H, C.byref(h)))
            self._h = h
            self.set_sizes(W, H)
            if flag is not None:
                self.updateFields(flag)
        n = C.c_int()
        check(lib().ubg_mg_num_levels(self._h, C.byref(n)))
        self.levels = n.value

    def _level_shape(self, level):
        if self.sizes is None:
            raise RuntimeError("sizes unknown")
        W, H = self.sizes[level]
        return (H, W)

    def set_sizes(self, W, H):
        self.sizes = level_sizes(W, H)

    def set_params(self, params):
        check(lib().ubg_mg_set_params(self._h, C.byref(params._p)))

    def updateFields(self, flag):
        flag = _f32(flag)
        if self.sizes is None:
            self.set_sizes(flag.shape[1], flag.shape[0])
        check(lib().ubg_mg_update_fields(self._h, fptr(flag)))

    def solve(self, p, f, flag, h, zeroGradientBC=False):
        """MG::solve on host grids; p (float32, C-contiguous) is updated in place."""
        assert p.dtype == np.float32 and p.flags.c_contiguous
        if self.sizes is None:
            self.set_sizes(p.shape[1], p.shape[0])
        fl = None if flag is None else _f32(flag)
        check(lib().ubg_mg_solve_host(self._h, fptr(p), fptr(_f32(f)), fptr(fl), h,
                                      int(bool(zeroGradientBC))))
        return p

    # resident-field interface -------------------------------------------------
    def set_level(self, level, which, arr):
        check(lib().ubg_mg_set_level_field(self._h, level, which, fptr(_f32(arr))))

    def get_level(self, level, which):
        out = np.zeros(self._level_shape(level), np.float32)
        check(lib().ubg_mg_get_level_field(self._h, level, which, fptr(out)))
        return out

    def get_flag(self, level):
        out = np.zeros(self._level_shape(level), np.float32)
        check(lib().ubg_mg_get_level_flag(self._h, level, fptr(out)))
        return out

    def level_op(self, op, level, h, alpha=1.0):
        check(lib().ubg_mg_level_op(self._h, op, level, h, alpha))

    def solve_resident(self, h, zeroGradientBC=False, ncycles=1):
        check(lib().ubg_mg_solve(self._h, h, int(bool(zeroGradientBC)), ncycles))

    def gs(self, h, alpha=1.0, nsweeps=1):
        """gs (pressure_solver.cpp:26-33) on the resident level-0 fields."""
        check(lib().ubg_mg_gs(self._h, h, alpha, nsweeps))

    def rbgs(self, h, alpha=1.0, nsweeps=1):
        check(lib().ubg_mg_rbgs(self._h, h, alpha, nsweeps))

    def residual(self, h, want_field=True):
        """calculateResidualField on level 0: (r, ||r||_2)."""
        r = np.zeros(self._level_shape(0), np.float32) if want_field else None
        n = C.c_double()
        check(lib().ubg_mg_residual(self._h, h, fptr(r), C.byref(n)))
        return r, n.value

    def sync(self):
        check(lib().ubg_mg_sync(self._h))

    # FP64 iterative refinement around FP32 V-cycles (extension) --------------
    def ir_set_rhs(self, f):
        """float32 right-hand side of the refinement solve (H x W)."""
        check(lib().ubg_mg_ir_set_rhs(self._h, fptr(_f32(f))))

    def ir_set_guess(self, x=None):
        """float64 initial guess incl. the Dirichlet ring; None = zeros."""
        if x is None:
            check(lib().ubg_mg_ir_set_guess(self._h, None))
            return
        x = np.ascontiguousarray(x, np.float64)
        check(lib().ubg_mg_ir_set_guess(self._h, x.ctypes.data_as(C.POINTER(C.c_double))))

    def ir_solve(self, h, tol=1e-10, vcycles_per_iter=4, max_iter=200, raise_on_stall=True,
                 ref_norm=0.0):
        """Refine to ||r|| <= tol ||r0|| (or tol * ref_norm when ref_norm > 0);
        returns a dict of the ubg_ir_result."""
        prm = L.IRParams(tol, vcycles_per_iter, max_iter, ref_norm)
        res = L.IRResult()
        rc = lib().ubg_mg_ir_solve(self._h, h, C.byref(prm), C.byref(res))
        if rc != L.E_NOT_CONVERGED or raise_on_stall:
            check(rc)
        n = min(res.iterations, L.IR_HISTORY)
        return dict(iterations=res.iterations, vcycles=res.vcycles,
                    converged=bool(res.converged), r0_norm=res.r0_norm,
                    rel_residual=res.rel_residual, history=list(res.history[:n]))

    def ir_solution(self):
        out = np.zeros(self._level_shape(0), np.float64)
        check(lib().ubg_mg_ir_get_solution(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and not self._borrowed:
            lib().ubg_mg_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """simulation.hpp:18-141: Simulation(flag, pwidth, mu), or
    Simulation.blank(pwidth, mu, width, height) for the reference's
    Simulation(pwidth, mu, width, height) (simulation.hpp:20-30)."""

    def __init__(self, flag, pwidth, mu, params=None, _blank_size=None):
        h = C.c_void_p()
        if _blank_size is not None:
            self.width, self.height = _blank_size
            check(lib().ubg_sim_create_blank(self.width, self.height, pwidth, mu, C.byref(h)))
        else:
            flag = _f32(flag)
            self.height, self.width = flag.shape
            check(lib().ubg_sim_create(fptr(flag), self.width, self.height, pwidth, mu,
                                       C.byref(h)))
        self._h = h
        self.pwidth, self.mu = float(pwidth), float(mu)
        hh = C.c_float()
        check(lib().ubg_sim_dims(self._h, None, None, C.byref(hh), None, None))
        self.h = hh.value
        if params is not None:
            self.set_params(params)
        mgh = C.c_void_p()
        check(lib().ubg_sim_get_mg(self._h, C.byref(mgh)))
        self.mg = MG(None, _handle=mgh, _owner=self)
        self.mg.set_sizes(self.width, self.height)
        self.dt = 0.0

    @classmethod
    def blank(cls, pwidth, mu, width, height, params=None):
        """Simulation(pwidth, mu, width, height): all fields zero, the flag
        too; the MG's coarse flags stay 1 until the flag is set."""
        return cls(None, pwidth, mu, params=params, _blank_size=(int(width), int(height)))

    def set_params(self, params):
        check(lib().ubg_sim_set_params(self._h, C.byref(params._p)))

    def shape(self, name):
        w, h = C.c_int(), C.c_int()
        check(lib().ubg_sim_field_shape(self._h, _fid(name), C.byref(w), C.byref(h)))
        return (h.value, w.value)

    def get(self, name):
        out = np.zeros(self.shape(name), np.float32)
        check(lib().ubg_sim_get_field(self._h, _fid(name), fptr(out)))
        return out

    def set(self, name, arr):
        arr = _f32(arr)
        if arr.shape != self.shape(name):
            raise ValueError(f"{name}: shape {arr.shape} != {self.shape(name)}")
        check(lib().ubg_sim_set_field(self._h, _fid(name), fptr(arr)))

    # reference methods --------------------------------------------------------
    def step(self, timestep):
        self.dt = timestep
        check(lib().ubg_sim_step(self._h, timestep))

    def run(self, dt, nsteps):
        check(lib().ubg_sim_run(self._h, dt, nsteps))

    def stage(self, st, dt=None):
        check(lib().ubg_sim_stage(self._h, st, self.dt if dt is None else dt))

    def applyAccumulatedVelocity(self):
        self.stage(L.STAGE_ACCUM)

    def diffuse(self):
        self.stage(L.STAGE_DIFFUSE)

    def advect(self):
        self.stage(L.STAGE_ADVECT)

    def setVBCs(self):
        self.stage(L.STAGE_SET_VBCS)

    def project(self):
        self.stage(L.STAGE_PROJECT)

    def setPBC(self):
        self.stage(L.STAGE_SET_PBC)

    def saveCurrentVelocityFields(self):
        self.stage(L.STAGE_SAVE_CURRENT)

    def setBCs(self, west, east, north, south):
        """The bcWest / bcEast / bcNorth / bcSouth members (simulation.hpp:
        127-128): INFLOW 0, OUTFLOW 1, OUTFLOW_ZERO_PRESSURE 2, NOSLIP 3."""
        check(lib().ubg_sim_set_bcs(self._h, int(west), int(east), int(north), int(south)))

    def setGrids(self, x, y, val):
        check(lib().ubg_sim_set_grids(self._h, int(x), int(y), float(val)))

    def add_sink(self, x, y, z):
        check(lib().ubg_sim_add_sink(self._h, x, y, z))

    def sinks(self):
        n = C.c_int()
        check(lib().ubg_sim_get_sinks(self._h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 3), np.float32)
        if n.value:
            check(lib().ubg_sim_get_sinks(self._h, fptr(out), n.value, C.byref(n)))
        return out

    def advectFloatingItems(self, items, game_dt, simple=False):
        """Simulation::advectFloatingItems / advectFloatingItemsSimple
        (advect_floating_items.cpp:16-274): items (numpy array of
        ITEM_DTYPE, physical coordinates) are updated in place; their
        reaction goes into vx_accum / vy_accum."""
        assert items.dtype == L.ITEM_DTYPE and items.flags.c_contiguous
        check(lib().ubg_sim_advect_items(self._h, items.ctypes.data_as(C.c_void_p), len(items),
                                         float(game_dt), int(bool(simple))))
        return items

    def shiftMap(self, new_flag):
        """The simulation part of UbootGlApp::shiftMap (ubootgl_app.cpp:248-296):
        fields shift one cell left, the new flag is applied with setGrids,
        inlet velocity, current fields saved, MG reclassified."""
        check(lib().ubg_sim_shift_map(self._h, fptr(_f32(new_flag))))

    def collocated(self):
        """(vxy, mag): the (2H-1) x (2W-1) x 2 collocated velocity image and its
        magnitude from vx_current / vy_current (VelocityTextures::
        updateFromStaggered, velocity_textures.cpp:63-93)."""
        W, H = self.width, self.height
        vxy = np.zeros((2 * H - 1, 2 * W - 1, 2), np.float32)
        mag = np.zeros((2 * H - 1, 2 * W - 1), np.float32)
        check(lib().ubg_sim_collocate(self._h, vxy.ctypes.data_as(C.POINTER(C.c_float)),
                                      fptr(mag), 0))
        return vxy, mag

    def _sample_flag(self, mode, pc):
        pc = _f32(pc).reshape(-1, 2)
        n = pc.shape[0]
        out = np.zeros((n, 2) if mode == L.FLAG_NORMAL else (n,), np.float32)
        check(lib().ubg_sim_sample_flag(self._h, mode, fptr(pc), fptr(out), n))
        return out

    def psampleFlagNearest(self, pc):
        """simulation.cpp:422-425 at physical positions pc (n, 2); cells
        outside the grid give 0."""
        return self._sample_flag(L.FLAG_NEAREST, pc)

    def psampleFlagLinear(self, pc):
        """simulation.cpp:398-411 at physical positions pc (n, 2)."""
        return self._sample_flag(L.FLAG_LINEAR, pc)

    def psampleFlagNormal(self, pc):
        """simulation.cpp:413-420 at physical positions pc (n, 2) -> (n, 2)."""
        return self._sample_flag(L.FLAG_NORMAL, pc)

    def getDT(self):
        """simulation.cpp:226-238."""
        out = C.c_float()
        check(lib().ubg_sim_get_dt(self._h, C.byref(out)))
        return out.value

    def centerP(self):
        """simulation.cpp:210-224."""
        check(lib().ubg_sim_center_p(self._h))

    def bilinearVel(self, c):
        c = _f32(c).reshape(-1, 2)
        out = np.zeros_like(c)
        check(lib().ubg_sim_bilinear_vel(self._h, fptr(c), fptr(out), c.shape[0]))
        return out

    def set_tracers(self, xy, age=None):
        xy = _f32(xy).reshape(-1, 2)
        a = None if age is None else _f32(age)
        check(lib().ubg_sim_set_tracers(self._h, fptr(xy), fptr(a), xy.shape[0]))
        self.ntracers = xy.shape[0]

    def get_tracers(self):
        n = self.ntracers
        xy = np.zeros((n, 2), np.float32)
        age = np.zeros(n, np.float32)
        check(lib().ubg_sim_get_tracers(self._h, fptr(xy), fptr(age), n))
        return xy, age

    def advect_tracers(self, dt, seed):
        check(lib().ubg_sim_advect_tracers(self._h, dt, seed))

    def sync(self):
        check(lib().ubg_sim_sync(self._h))

    def timer_enable(self, timer=L.TIMER_SMOOTH_L0, on=True):
        check(lib().ubg_sim_timer_enable(self._h, timer, int(on)))

    def timer_read(self, timer=L.TIMER_SMOOTH_L0):
        """(total ms, launches) of the kernel class `timer` since enabling."""
        ms, n = C.c_double(), C.c_long()
        check(lib().ubg_sim_timer_read(self._h, timer, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None) is not None:
            self.mg._h = None
            lib().ubg_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def terrain_flags(rgba, scale):
    """Terrain flags from a decoded RGBA8 image (H x W x 4, row 0 at the top):
    (flag_full (H x W), flag_sim (H/scale x W/scale)) -- terrain.cpp:3-45."""
    rgba = np.ascontiguousarray(rgba, np.uint8)
    H, W = rgba.shape[:2]
    full = np.zeros((H, W), np.float32)
    sim = np.zeros((H // scale, W // scale), np.float32)
    check(lib().ubg_terrain_flags(rgba.ctypes.data_as(C.POINTER(C.c_uint8)), W, H, int(scale),
                                  fptr(full), fptr(sim)))
    return full, sim


def flag_components(flag):
    """Fluid components of a 0/1 flag under 4-neighbour connectivity, on the
    GPU: (labels int32 H x W, n), labels 0..n-1 in raster order of each
    component's first cell, -1 on solid cells (ubg_flag_components)."""
    fl = _f32(flag)
    H, W = fl.shape
    lab = np.zeros((H, W), np.int32)
    n = C.c_int()
    check(lib().ubg_flag_components(fptr(fl), W, H, lab.ctypes.data_as(C.POINTER(C.c_int)),
                                    C.byref(n)))
    return lab, n.value


def remove_component_means(flag, f):
    """f with its mean over each fluid component removed, on the GPU
    (ubg_remove_component_means); returns (f', n)."""
    fl = _f32(flag)
    out = np.array(f, np.float32, copy=True, order="C")
    H, W = fl.shape
    assert out.shape == (H, W)
    n = C.c_int()
    check(lib().ubg_remove_component_means(fptr(fl), W, H, fptr(out), C.byref(n)))
    return out, n.value
