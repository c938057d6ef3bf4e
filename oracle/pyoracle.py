"""ctypes wrapper of the CPU oracle (oracle/ubg_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the ubootgl_amd package.  It restates
the reference's CPU solver (simulation.cpp, pressure_solver.cpp,
interpolators.hpp) so that the HIP path can be checked against it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libubg_oracle.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

# field ids (same numbering as the product's UBG_FIELD_*)
VX, VX_BACK, VY, VY_BACK, VX_ACCUM, VY_ACCUM, VX_CURRENT, VY_CURRENT, P, F, FLAG, R = range(12)
FIELD_NAMES = ["vx", "vx_back", "vy", "vy_back", "vx_accum", "vy_accum",
               "vx_current", "vy_current", "p", "f", "flag", "r"]
# stage ids (same numbering as the product's UBG_STAGE_*)
(ST_ACCUM, ST_DIFFUSE, ST_ADVECT, ST_SET_VBCS, ST_PROJECT, ST_SET_PBC,
 ST_SAVE_CURRENT, ST_DIVERGENCE, ST_CORRECT_VELOCITY, ST_MG_SOLVE) = range(10)

# the timing build of the same source (reference flags -Ofast -march=native,
# eight-wide advect), loaded only by bench.py's cpu_baseline leg
NATIVE_LIB_PATH = os.path.join(_HERE, "_build", "libubg_oracle_native.so")

_lib = None
_lib_path = LIB_PATH


def use_timing_build(path=NATIVE_LIB_PATH):
    """Switch this process to the timing build (bench.py's cpu_baseline only:
    its -Ofast numerics make it no checker)."""
    global _lib, _lib_path
    _lib, _lib_path = None, path


def lib():
    global _lib
    if _lib is None:
        _lib = load(_lib_path)
    return _lib


def load(path):
    """A ctypes handle of one build of the oracle with every signature set."""
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
    L = C.CDLL(path)
    sig = {
        "oracle_set_threads": (C.c_int, [C.c_int]),
        "oracle_mg_level_sizes": (C.c_int, [C.c_int, C.c_int, _i32p, _i32p, C.c_int]),
        "oracle_coarsen_flag": (None, [_f32p, C.c_int, _f32p, C.c_int, C.c_int]),
        "oracle_rbgs_f32": (None, [_f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float]),
        "oracle_rbgs_f64": (None, [_f64p, _f64p, _f64p, C.c_int, C.c_int, C.c_double, C.c_double]),
        "oracle_residual_f32": (C.c_double, [_f32p, _f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float]),
        "oracle_residual_f64": (C.c_double, [_f64p, _f64p, _f64p, _f64p, C.c_int, C.c_int, C.c_double]),
        "oracle_restrict_f32": (None, [_f32p, C.c_int, _f32p, C.c_int, C.c_int]),
        "oracle_prolongate_f32": (None, [_f32p, C.c_int, C.c_int, _f32p, _f32p, C.c_int, _f32p]),
        "oracle_correct_f32": (None, [_f32p, _f32p, C.c_int, C.c_int]),
        "oracle_zero_gradient_f32": (None, [_f32p, C.c_int, C.c_int]),
        "oracle_bicubic": (C.c_float, [_f32p, C.c_int, C.c_int, C.c_float, C.c_float]),
        "oracle_bilinear": (C.c_float, [_f32p, C.c_int, C.c_int, C.c_float, C.c_float]),
        "oracle_cubic_hermite": (C.c_float, [C.c_float] * 5),
        "oracle_wang_hash": (C.c_uint32, [C.c_uint32]),
        "oracle_check_count_division": (C.c_long, [C.c_uint32]),
        "oracle_lcg": (C.c_uint32, [C.c_uint32]),
        "oracle_mg_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int]),
        "oracle_mg_destroy": (None, [C.c_void_p, C.c_int]),
        "oracle_mg_set_params": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float]),
        "oracle_mg_update_fields": (None, [C.c_void_p, _f32p]),
        "oracle_mg64_update_fields": (None, [C.c_void_p, _f64p]),
        "oracle_mg_get_flag": (None, [C.c_void_p, C.c_int, _f32p]),
        "oracle_mg_solve": (None, [C.c_void_p, _f32p, _f32p, _f32p, C.c_float, C.c_int]),
        "oracle_mg64_solve": (None, [C.c_void_p, _f64p, _f64p, _f64p, C.c_double, C.c_int]),
        "oracle_sim_create": (C.c_void_p, [_f32p, C.c_int, C.c_int, C.c_float, C.c_float]),
        "oracle_sim_destroy": (None, [C.c_void_p]),
        "oracle_sim_set_params": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float]),
        "oracle_sim_set_field": (C.c_int, [C.c_void_p, C.c_int, _f32p]),
        "oracle_sim_get_field": (C.c_int, [C.c_void_p, C.c_int, _f32p]),
        "oracle_sim_h": (C.c_float, [C.c_void_p]),
        "oracle_sim_add_sink": (None, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
        "oracle_sim_num_sinks": (C.c_int, [C.c_void_p]),
        "oracle_sim_step": (None, [C.c_void_p, C.c_float]),
        "oracle_sim_set_bcs": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
        "oracle_sim_stage": (C.c_int, [C.c_void_p, C.c_int, C.c_float]),
        "oracle_sim_bilinear_vel": (None, [C.c_void_p, _f32p, _f32p, C.c_int]),
        "oracle_flag_linear": (C.c_float, [_f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]),
        "oracle_flag_nearest": (C.c_float, [_f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]),
        "oracle_flag_normal": (None, [_f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                      C.c_float, _f32p]),
        "oracle_gs_f32": (None, [_f32p, _f32p, _f32p, C.c_int, C.c_int, C.c_float, C.c_float]),
        "oracle_sim_center_p": (None, [C.c_void_p]),
        "oracle_sim_get_dt": (C.c_float, [C.c_void_p]),
        "oracle_sim_advect_tracers": (None, [C.c_void_p, _f32p, _f32p, C.c_int, C.c_float, C.c_uint32]),
        "oracle_sim_advect_items": (None, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_int]),
        "oracle_cephes_sin": (C.c_float, [C.c_float]),
        "oracle_cephes_cos": (C.c_float, [C.c_float]),
        "oracle_components": (C.c_int, [_f32p, C.c_int, C.c_int, _i32p]),
        "oracle_remove_component_means": (C.c_int, [_f32p, C.c_int, C.c_int, _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def set_threads(n=0):
    return lib().oracle_set_threads(int(n))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def level_sizes(W, H):
    ws = np.zeros(64, np.int32)
    hs = np.zeros(64, np.int32)
    n = lib().oracle_mg_level_sizes(W, H, ws, hs, 64)
    return [(int(ws[i]), int(hs[i])) for i in range(n)]


def rbgs(p, f, flag, h, alpha=1.0):
    """pressure_solver.cpp:49-89 (two-pass semantics); p updated in place."""
    H, W = p.shape
    if p.dtype == np.float64:
        lib().oracle_rbgs_f64(p, f64(f), f64(flag), W, H, h, alpha)
    else:
        lib().oracle_rbgs_f32(p, f32(f), f32(flag), W, H, h, alpha)
    return p


def residual(p, f, flag, h):
    """pressure_solver.cpp:91-116; returns (r, ||r||_2)."""
    H, W = p.shape
    if p.dtype == np.float64:
        r = np.zeros_like(p)
        n = lib().oracle_residual_f64(p, f64(f), f64(flag), r, W, H, h)
    else:
        r = np.zeros((H, W), np.float32)
        n = lib().oracle_residual_f32(f32(p), f32(f), f32(flag), r, W, H, h)
    return r, n


def gs(p, f, flag, h, alpha=1.0):
    """pressure_solver.cpp:26-33 (lexicographic sweep); float32 p in place."""
    H, W = p.shape
    lib().oracle_gs_f32(p, f32(f), f32(flag), W, H, h, alpha)
    return p


def flag_sample(flag, pwidth, h, mode, pts):
    """psampleFlagNearest (mode 0), psampleFlagLinear (1), psampleFlagNormal
    (2) (simulation.cpp:398-425) at physical positions pts (n, 2); returns
    (n,) or (n, 2) float32."""
    flag = f32(flag)
    H, W = flag.shape
    pts = np.asarray(pts, np.float32).reshape(-1, 2)
    if mode == 2:
        out = np.zeros((len(pts), 2), np.float32)
        tmp = np.zeros(2, np.float32)
        for i, (x, y) in enumerate(pts):
            lib().oracle_flag_normal(flag, W, H, pwidth, h, float(x), float(y), tmp)
            out[i] = tmp
        return out
    fn = lib().oracle_flag_nearest if mode == 0 else lib().oracle_flag_linear
    return np.array([fn(flag, W, H, pwidth, float(x), float(y)) for x, y in pts], np.float32)


def restrict(r, Wc, Hc):
    H, W = r.shape
    rc = np.zeros((Hc, Wc), np.float32)
    lib().oracle_restrict_f32(f32(r), W, rc, Wc, Hc)
    return rc


def prolongate(W, H, ec, flagc, flag):
    Hc, Wc = ec.shape
    e = np.zeros((H, W), np.float32)
    lib().oracle_prolongate_f32(e, W, H, f32(ec), f32(flagc), Wc, f32(flag))
    return e


def correct(p, e):
    H, W = p.shape
    lib().oracle_correct_f32(p, f32(e), W, H)
    return p


def zero_gradient(p):
    H, W = p.shape
    lib().oracle_zero_gradient_f32(p, W, H)
    return p


def coarsen_flag(flag, Wc, Hc):
    H, W = flag.shape
    fc = np.zeros((Hc, Wc), np.float32)
    lib().oracle_coarsen_flag(f32(flag), W, fc, Wc, Hc)
    return fc


def bicubic(g, cx, cy):
    gh, gw = g.shape
    return lib().oracle_bicubic(f32(g), gw, gh, cx, cy)


def bilinear(g, cx, cy):
    gh, gw = g.shape
    return lib().oracle_bilinear(f32(g), gw, gh, cx, cy)


class OracleMG:
    """pressure_solver.hpp:13-76 (MG), fp32 or fp64 storage."""

    def __init__(self, W, H, fp64=False, nu_pre=3, nu_post=3, nu_coarse=5, alpha=1.0):
        self.W, self.H, self.fp64 = W, H, bool(fp64)
        self.h_ = lib().oracle_mg_create(W, H, int(self.fp64))
        lib().oracle_mg_set_params(self.h_, int(self.fp64), nu_pre, nu_post, nu_coarse, alpha)
        self.sizes = level_sizes(W, H)

    def update_fields(self, flag):
        if self.fp64:
            lib().oracle_mg64_update_fields(self.h_, f64(flag))
        else:
            lib().oracle_mg_update_fields(self.h_, f32(flag))

    def flag_level(self, level):
        assert not self.fp64
        W, H = self.sizes[level]
        out = np.zeros((H, W), np.float32)
        lib().oracle_mg_get_flag(self.h_, level, out)
        return out

    def solve(self, p, f, flag, h, zero_gradient_bc=False):
        if self.fp64:
            lib().oracle_mg64_solve(self.h_, p, f64(f), f64(flag), h, int(zero_gradient_bc))
        else:
            lib().oracle_mg_solve(self.h_, p, f32(f), f32(flag), h, int(zero_gradient_bc))
        return p

    def __del__(self):
        try:
            lib().oracle_mg_destroy(self.h_, int(self.fp64))
        except Exception:
            pass


class OracleSim:
    """simulation.hpp:18-141 (Simulation) with the reference constructor."""

    def __init__(self, flag, pwidth, mu, nu_pre=3, nu_post=3, nu_coarse=5, vcycles=2, alpha=1.0):
        flag = f32(flag)
        self.H, self.W = flag.shape
        self.h_ = lib().oracle_sim_create(flag, self.W, self.H, pwidth, mu)
        lib().oracle_sim_set_params(self.h_, nu_pre, nu_post, nu_coarse, vcycles, alpha)
        self.pwidth, self.mu = pwidth, mu

    @property
    def h(self):
        return lib().oracle_sim_h(self.h_)

    def shape(self, fid):
        W, H = self.W, self.H
        if fid in (VX, VX_BACK, VX_ACCUM, VX_CURRENT):
            return (H, W - 1)
        if fid in (VY, VY_BACK, VY_ACCUM, VY_CURRENT):
            return (H - 1, W)
        return (H, W)

    def get(self, fid):
        out = np.zeros(self.shape(fid), np.float32)
        assert lib().oracle_sim_get_field(self.h_, fid, out) == 0
        return out

    def set(self, fid, arr):
        arr = f32(arr)
        assert arr.shape == self.shape(fid)
        assert lib().oracle_sim_set_field(self.h_, fid, arr) == 0

    def step(self, dt):
        lib().oracle_sim_step(self.h_, dt)

    def center_p(self):
        lib().oracle_sim_center_p(self.h_)

    def get_dt(self):
        return lib().oracle_sim_get_dt(self.h_)

    def stage(self, st, dt):
        assert lib().oracle_sim_stage(self.h_, st, dt) == 0

    def set_bcs(self, west, east, north, south):
        lib().oracle_sim_set_bcs(self.h_, west, east, north, south)

    def add_sink(self, x, y, z):
        lib().oracle_sim_add_sink(self.h_, x, y, z)

    def num_sinks(self):
        return lib().oracle_sim_num_sinks(self.h_)

    def bilinear_vel(self, c):
        c = f32(c).reshape(-1, 2)
        out = np.zeros_like(c)
        lib().oracle_sim_bilinear_vel(self.h_, c, out, c.shape[0])
        return out

    def advect_items(self, items, game_dt, simple=False):
        """advect_floating_items.cpp:16-274 on this simulation; items is a
        numpy array with the ubg_item layout (updated in place)."""
        lib().oracle_sim_advect_items(self.h_, items.ctypes.data, len(items), float(game_dt),
                                      int(bool(simple)))
        return items

    def advect_tracers(self, xy, age, dt, seed):
        """xy (n,2) float32 and age (n,) float32 are updated in place."""
        assert xy.dtype == np.float32 and xy.flags.c_contiguous
        assert age.dtype == np.float32 and age.flags.c_contiguous
        lib().oracle_sim_advect_tracers(self.h_, xy.reshape(-1), age, age.shape[0], dt, seed)

    def __del__(self):
        try:
            lib().oracle_sim_destroy(self.h_)
        except Exception:
            pass


def solve_poisson_fp64(flag, f, h, tol=1e-13, nu_pre=3, nu_post=3, nu_coarse=5, max_cycles=2000,
                       x0=None):
    """The FP64 CPU multigrid reference for the refinement solves: float64
    V-cycles (pressure_solver.cpp:201-248 in double) on x from x0 (default 0,
    the Dirichlet ring included) until ||r||_2 <= tol ||r_0||_2.  Returns
    (x, history of relative residuals)."""
    H, W = flag.shape
    mg = OracleMG(W, H, fp64=True, nu_pre=nu_pre, nu_post=nu_post, nu_coarse=nu_coarse)
    mg.update_fields(flag)
    x = np.zeros((H, W), np.float64) if x0 is None else np.array(x0, np.float64, order="C")
    ff, fl = f64(f), f64(flag)
    _, r0 = residual(x, ff, fl, float(h))
    hist = []
    if r0 == 0.0:
        return x, hist
    for _ in range(max_cycles):
        mg.solve(x, ff, fl, float(h))
        _, rn = residual(x, ff, fl, float(h))
        hist.append(rn / r0)
        if hist[-1] <= tol:
            break
    return x, hist


def collocate(vx, vy):
    """Collocated velocity image of interp_shader.cs:14-33 (dispatched by
    VelocityTextures::updateFromStaggered, velocity_textures.cpp:63-93) with
    the default sampler (linear filtering between texel centres, GL_REPEAT),
    in float32 in the kernel's operation order: pixel (gx, gy) samples vx at
    texel point ((gx-1)/2, gy/2) and vy at (gx/2, (gy-1)/2).  Returns
    (vxy (2H-1, 2W-1, 2), mag (2H-1, 2W-1))."""
    vx = np.asarray(vx, np.float32)
    vy = np.asarray(vy, np.float32)
    H = vx.shape[0]
    W = vx.shape[1] + 1
    gx = np.arange(2 * W - 1)
    gy = np.arange(2 * H - 1)
    half, one = np.float32(0.5), np.float32(1.0)

    def linear(g, twx, twy):
        gh, gw = g.shape
        i0, j0 = np.floor_divide(twx, 2), np.floor_divide(twy, 2)
        wx = np.where(twx % 2 == 1, half, np.float32(0))[None, :]
        wy = np.where(twy % 2 == 1, half, np.float32(0))[:, None]
        xa, xb = np.mod(i0, gw), np.mod(i0 + 1, gw)
        ya, yb = np.mod(j0, gh), np.mod(j0 + 1, gh)
        t00, t10 = g[np.ix_(ya, xa)], g[np.ix_(ya, xb)]
        t01, t11 = g[np.ix_(yb, xa)], g[np.ix_(yb, xb)]
        r0 = (one - wx) * t00 + wx * t10
        r1 = (one - wx) * t01 + wx * t11
        return (one - wy) * r0 + wy * r1

    a = linear(vx, gx - 1, gy)
    b = linear(vy, gx, gy - 1)
    mag = np.sqrt(a * a + b * b)
    return np.stack([a, b], -1).astype(np.float32), mag.astype(np.float32)


def shift_map(fields, new_flag):
    """The simulation part of UbootGlApp::shiftMap (ubootgl_app.cpp:248-296)
    on a dict of float32 fields (vx, vx_back, vy, vy_back, p): the left shifts
    (:253-269), setGrids(c, new_flag(c)) for every cell (:271-275,
    simulation.hpp:82-98), the inlet velocity (:278-287), the current copies
    (:292).  Returns the new dict incl. flag, vx_current, vy_current."""
    f = {k: np.array(v, np.float32, copy=True) for k, v in fields.items()}
    flag = np.asarray(new_flag, np.float32)
    H, W = flag.shape
    # for x in 2 .. w-1: f(x-1) = f(x), b(x-1) = f(x)  (f(x) not yet overwritten)
    for a, b in (("vx", "vx_back"), ("vy", "vy_back")):
        w = f[a].shape[1]
        moved = f[a][:, 2:w].copy()
        f[a][:, 1:w - 1] = moved
        f[b][:, 1:w - 1] = moved
    f["p"][:, 0:W - 1] = f["p"][:, 1:W].copy()
    solid = flag == 0
    f["vx"][solid[:, :W - 1]] = 0.0
    f["vx"][solid[:, 1:]] = 0.0
    f["vy"][solid[:H - 1, :]] = 0.0
    f["vy"][solid[1:, :]] = 0.0
    f["p"][solid] = 0.0
    area = np.float32(1.0)
    for y in range(H - 1):
        area = np.float32(area + flag[y, 0])
    vin = np.float32(np.float32(0.07) * np.float32(H)) / area
    col = (vin * flag[:, 0]).astype(np.float32)
    f["vx"][:, 0] = col
    f["vx_back"][:, 0] = col
    f["flag"] = flag.copy()
    f["vx_current"] = f["vx"].copy()
    f["vy_current"] = f["vy"].copy()
    return f


def terrain_flags(rgba, scale):
    """Terrain's flags (terrain.cpp:3-45): full = red / 255 with rows flipped;
    sim(x, y) = (block sum in row-major order / scale / scale) > 0.99, float32."""
    rgba = np.asarray(rgba, np.uint8)
    H, W = rgba.shape[:2]
    full = (rgba[::-1, :, 0].astype(np.float32) / np.float32(255.0)).astype(np.float32)
    sh, sw = H // scale, W // scale
    sim = np.zeros((sh, sw), np.float32)
    acc = np.zeros((sh, sw), np.float32)
    blocks = full[:sh * scale, :sw * scale].reshape(sh, scale, sw, scale)
    for sy in range(scale):
        for sx in range(scale):
            acc = (acc + blocks[:, sy, :, sx]).astype(np.float32)
    mean = (acc / np.float32(scale) / np.float32(scale)).astype(np.float32)
    sim[mean > np.float32(0.99)] = 1.0
    return full, sim


def cephes_sin(x):
    return lib().oracle_cephes_sin(float(x))


def cephes_cos(x):
    return lib().oracle_cephes_cos(float(x))


def components(flag):
    """Fluid components under 4-neighbour connectivity: (labels int32, n),
    labels 0..n-1 in raster order of each component's first cell, -1 on
    solid cells."""
    fl = f32(flag)
    H, W = fl.shape
    lab = np.zeros((H, W), np.int32)
    n = lib().oracle_components(fl, W, H, lab)
    return lab, n


def remove_component_means(flag, f):
    """f with its mean over each fluid component removed (64-bit fixed-point
    sums, float64 means); returns (f', n)."""
    fl = f32(flag)
    out = f32(f).copy()
    H, W = fl.shape
    n = lib().oracle_remove_component_means(fl, W, H, out)
    return out, n
