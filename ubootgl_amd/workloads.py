"""Seeded synthetic workloads (flag fields, velocity fields, tracers).

Nothing in the reference fixes these inputs; they are defined here, generated
on the host in float64 with numpy's PCG64 and rounded to float32 once, so the
GPU path and the CPU oracle always see identical bytes.

Frames: cell (i, j) has its centre at grid coordinate (i, j); vx(i, j) sits at
(i + 1/2, j) and vy(i, j) at (i, j + 1/2), as in the reference's staggered grid
(simulation.hpp:22-23, advect_floating_items.cpp:11-14).
"""
import math

import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def walls(W, H):
    """1 = fluid, 0 = solid; walls on all four sides."""
    f = np.ones((H, W), np.float64)
    f[0, :] = f[-1, :] = 0.0
    f[:, 0] = f[:, -1] = 0.0
    return f


def stamp_disks(flag, cx, cy, r):
    """Cell (i, j) becomes solid iff (i - cx)^2 + (j - cy)^2 <= r^2 (float64)."""
    H, W = flag.shape
    for x0, y0, rr in zip(cx, cy, r):
        i0, i1 = max(0, int(math.floor(x0 - rr))), min(W - 1, int(math.ceil(x0 + rr)))
        j0, j1 = max(0, int(math.floor(y0 - rr))), min(H - 1, int(math.ceil(y0 + rr)))
        if i1 < i0 or j1 < j0:
            continue
        ii = np.arange(i0, i1 + 1, dtype=np.float64)
        jj = np.arange(j0, j1 + 1, dtype=np.float64)
        d2 = (ii[None, :] - x0) ** 2 + (jj[:, None] - y0) ** 2
        sub = flag[j0:j1 + 1, i0:i1 + 1]
        sub[d2 <= rr * rr] = 0.0
    return flag


def solid_fraction(flag):
    return 1.0 - float(flag[1:-1, 1:-1].mean())


def mode_field(shape, X, Y, modes):
    """sum over modes of A sin(2 pi (kx X + ky Y) + phase), X/Y in domain units."""
    out = np.zeros(shape, np.float64)
    for kx, ky, ph, amp in modes:
        out += amp * np.sin(2.0 * np.pi * (kx * X + ky * Y) + ph)
    return out


def draw_modes(g, n, kmax=6, amp=1.0):
    return [(int(g.integers(1, kmax + 1)), int(g.integers(1, kmax + 1)),
             float(g.uniform(0.0, 2.0 * np.pi)), amp) for _ in range(n)]


def velocity(flag, g, nmodes=8, kmax=6):
    """Sinusoid-mode staggered velocity (float64), face-masked by the flag."""
    H, W = flag.shape
    mx = draw_modes(g, nmodes, kmax)
    my = draw_modes(g, nmodes, kmax)
    # vx(i, j) at (i + 1/2, j); vy(i, j) at (i, j + 1/2)
    Xx = (np.arange(W - 1, dtype=np.float64)[None, :] + 0.5) / W
    Yx = np.arange(H, dtype=np.float64)[:, None] / H
    Xy = np.arange(W, dtype=np.float64)[None, :] / W
    Yy = (np.arange(H - 1, dtype=np.float64)[:, None] + 0.5) / H
    vx = mode_field((H, W - 1), Xx, Yx, mx) * flag[:, :-1] * flag[:, 1:]
    vy = mode_field((H - 1, W), Xy, Yy, my) * flag[:-1, :] * flag[1:, :]
    return vx, vy


class Domain:
    """One synthetic fluid domain: flag, staggered velocity, h, pwidth, dt."""

    def __init__(self, flag, vx, vy, mu=0.001, cfl=0.5):
        self.flag = np.ascontiguousarray(flag, np.float32)
        self.H, self.W = self.flag.shape
        self.vx = np.ascontiguousarray(vx, np.float32)
        self.vy = np.ascontiguousarray(vy, np.float32)
        # h = 1/W, hence pwidth = (W-1)/W since h = pwidth/(W-1) (simulation.hpp:63)
        self.pwidth = np.float32((self.W - 1) / self.W)
        self.h = np.float32(self.pwidth / np.float32(self.W - 1))
        self.mu = np.float32(mu)
        vmax = max(float(np.abs(self.vx).max()), float(np.abs(self.vy).max()), 1e-12)
        self.dt = np.float32(cfl * float(self.h) / vmax)

    def init_simulation_fields(self):
        """vx/vy as a freshly constructed Simulation holds them after the
        workload's field is written: interior from the field, then the
        constructor's vx(0, y) = 1, vy(x, 0) = 0 (simulation.hpp:55-60).  Both
        buffers of each DoubleBuffered2DGrid get the same values."""
        vx = self.vx.copy()
        vy = self.vy.copy()
        vx[:, 0] = 1.0
        vy[0, :] = 0.0
        return vx, vy


def config0(seed=0):
    """512x256, walls + 12 disks of radius U[4,24], 8 modes (kx,ky in 1..6)."""
    W, H = 512, 256
    g = rng(seed)
    flag = walls(W, H)
    n = 12
    cx = g.uniform(0, W, n)
    cy = g.uniform(0, H, n)
    r = g.uniform(4, 24, n)
    stamp_disks(flag, cx, cy, r)
    vx, vy = velocity(flag, g, 8, 6)
    return Domain(flag, vx, vy)


def disks_to_fraction(W, H, g, ndisks, rmin, rmax, target, tol=0.0005, base=None):
    """`ndisks` disks with log-uniform base radii in [rmin, rmax]; the scale s
    in r = clamp(s*rho, rmin, rmax) is bisected until the rasterised solid
    fraction (interior cells) is within tol of target.  `base` makes the
    starting flag field (default: walls)."""
    base = base or walls
    cx = g.uniform(0, W, ndisks)
    cy = g.uniform(0, H, ndisks)
    rho = np.exp(g.uniform(math.log(rmin), math.log(rmax), ndisks))
    lo, hi = 0.0, 1.0
    best = None
    for _ in range(40):
        s = 0.5 * (lo + hi)
        r = np.clip(s * rho, rmin, rmax)
        fl = stamp_disks(base(W, H), cx, cy, r)
        fr = solid_fraction(fl)
        best = fl
        if abs(fr - target) <= tol:
            break
        if fr < target:
            lo = s
        else:
            hi = s
    return best


def config1(seed=1):
    """4096x2048, 400 disks (radius 8..128, scaled to 5% solid), 8 modes."""
    W, H = 4096, 2048
    g = rng(seed)
    flag = disks_to_fraction(W, H, g, 400, 8.0, 128.0, 0.05)
    vx, vy = velocity(flag, g, 8, 6)
    return Domain(flag, vx, vy)


def tracers_in_fluid(d, n, seed):
    """n points uniform in the fluid cells, physical coordinates of the flag
    sampler (cell i spans [i, i+1) * pwidth / W, simulation.cpp:398-411)."""
    g = rng(seed)
    fl = d.flag[1:-1, 1:-1] > 0.5
    jj, ii = np.nonzero(fl)
    pick = g.integers(0, ii.size, n)
    scale = float(d.pwidth) / d.W
    x = (ii[pick] + 1 + g.uniform(0, 1, n)) * scale
    y = (jj[pick] + 1 + g.uniform(0, 1, n)) * scale
    return np.ascontiguousarray(np.stack([x, y], 1), np.float32)


def small_domain(W, H, seed, ndisks=3, nmodes=4):
    """A small random domain for parity tests."""
    g = rng(seed)
    flag = walls(W, H)
    if ndisks:
        rmax = max(2.0, min(W, H) / 6.0)
        stamp_disks(flag, g.uniform(0, W, ndisks), g.uniform(0, H, ndisks),
                    g.uniform(1.5, rmax, ndisks))
    vx, vy = velocity(flag, g, nmodes, 4)
    return Domain(flag, vx, vy)


# ---------------------------------------------------------------------------
# Standalone projection solves (mixed-precision, batched and served solves)
# ---------------------------------------------------------------------------

def open_ring(W, H):
    """All fluid: the outer ring holds Dirichlet values, as in mgtest
    (mgtest/mgtest.cpp:14-29)."""
    return np.ones((H, W), np.float64)


def fill_enclosed(flag):
    """Make solid every fluid cell with no 4-connected fluid path to the
    outer ring: such a pocket would be a pure-Neumann (singular) problem."""
    from scipy import ndimage
    lab, _ = ndimage.label(flag > 0.5)  # 4-connectivity, the stencil's
    ring = np.unique(np.concatenate([lab[0], lab[-1], lab[:, 0], lab[:, -1]]))
    keep = np.isin(lab, ring[ring > 0])
    return np.where(keep, 1.0, 0.0)


def divergence_rhs(flag, vx, vy, h):
    """f = -(1/h)(vx(x) - vx(x-1) + vy(y) - vy(y-1)) on interior fluid cells
    (the divergence of simulation.cpp:166-171), 0 on the ring and solids;
    float64."""
    H, W = flag.shape
    f = np.zeros((H, W), np.float64)
    f[1:-1, 1:-1] = -(vx[1:-1, 1:] - vx[1:-1, :-1] + vy[1:, 1:-1] - vy[:-1, 1:-1]) / h
    return f * flag


class PoissonProblem:
    """One standalone projection solve: flag (1 = fluid; ring fluid and held
    at x = 0), float32 rhs f, grid spacing h = 1/W."""

    def __init__(self, flag, f, h):
        self.flag = np.ascontiguousarray(flag, np.float32)
        self.H, self.W = self.flag.shape
        self.f = np.ascontiguousarray(f, np.float32)
        self.h = np.float32(h)

    @property
    def cells(self):
        return self.W * self.H


def poisson_problem(W, H, seed, solid=0.03, ndisks=None, rmin=None, rmax=None, nmodes=16,
                    kmax=6):
    """Seeded projection solve on a W x H grid: disks (log-uniform radii in
    [rmin, rmax], scaled to the `solid` fraction), enclosed pockets filled,
    f = -div of a face-masked `nmodes`-mode velocity field."""
    g = rng(seed)
    if ndisks is None:
        ndisks = max(16, W * H // 20000)
    rmin = rmin if rmin is not None else max(1.5, min(W, H) / 256)
    rmax = rmax if rmax is not None else max(2 * rmin, min(W, H) / 8)
    if solid > 0:
        flag = disks_to_fraction(W, H, g, ndisks, rmin, rmax, solid,
                                 tol=max(0.0005, 0.05 * solid), base=open_ring)
    else:
        flag = open_ring(W, H)
    flag[0, :] = flag[-1, :] = 1.0
    flag[:, 0] = flag[:, -1] = 1.0
    flag = fill_enclosed(flag)
    h = np.float32(1.0 / W)
    vx, vy = velocity(flag, g, nmodes, kmax)
    f = divergence_rhs(flag, vx, vy, float(h))
    return PoissonProblem(flag, f, h)


def config3(seed=3):
    """8192x4096 mixed-precision solve: 800 disks (radius 8..128) scaled to
    3% solid, rhs = -div of a 16-mode field."""
    return poisson_problem(8192, 4096, seed, solid=0.03, ndisks=800, rmin=8.0, rmax=128.0,
                           nmodes=16)


def batch_sizes(n, seed, wmin=128, wmax=2048):
    """Domain sizes of the batched workload: W = 32 round(exp(U[ln wmin,
    ln wmax]) / 32), aspect a ~ U[1, 4], H = max(32, 32 round(W / a / 32))."""
    g = rng(seed)
    out = []
    for _ in range(n):
        W = int(32 * round(math.exp(g.uniform(math.log(wmin), math.log(wmax))) / 32))
        a = g.uniform(1.0, 4.0)
        H = max(32, int(32 * round(W / a / 32)))
        out.append((W, H))
    return out


def config2(n=256, seed=2):
    """256 independent solve domains of uneven size (batch_sizes), each with
    U[0, 8%] disk obstacles and its own 8-mode velocity field."""
    g = rng(seed)
    sizes = batch_sizes(n, seed)
    probs = []
    for i, (W, H) in enumerate(sizes):
        solid = float(g.uniform(0.0, 0.08))
        probs.append(poisson_problem(W, H, seed=10_000 * seed + i, solid=solid, nmodes=8))
    return probs


def serving_sizes(n, seed, lo=64, wmax=4096, hmax=2048):
    """W log-uniform in [lo, wmax], H log-uniform in [lo, hmax], independent,
    rounded to multiples of 32."""
    g = rng(seed)
    out = []
    for _ in range(n):
        W = int(32 * round(math.exp(g.uniform(math.log(lo), math.log(wmax))) / 32))
        H = int(32 * round(math.exp(g.uniform(math.log(lo), math.log(hmax))) / 32))
        out.append((max(lo, W), max(lo, H)))
    return out


class ServingTemplate:
    """A served domain: flag, h and the rhs of `timesteps` consecutive
    timesteps (the velocity's mode phases advance by `dphase` per step)."""

    def __init__(self, flag, rhs, h):
        self.flag = np.ascontiguousarray(flag, np.float32)
        self.H, self.W = self.flag.shape
        self.rhs = [np.ascontiguousarray(r, np.float32) for r in rhs]
        self.h = np.float32(h)
        inner = self.flag[1:-1, 1:-1] > 0.5
        self.fnorm = [float(np.sqrt((r[1:-1, 1:-1][inner].astype(np.float64) ** 2).sum()))
                      for r in self.rhs]

    @property
    def cells(self):
        return self.W * self.H


def serving_template(W, H, seed, timesteps=3, dphase=0.25, nmodes=8):
    g = rng(seed)
    solid = float(g.uniform(0.0, 0.08))
    P = poisson_problem(W, H, seed=seed, solid=solid, nmodes=nmodes)
    flag = P.flag.astype(np.float64)
    g2 = rng(seed + 1)
    mx = draw_modes(g2, nmodes, 6)
    my = draw_modes(g2, nmodes, 6)
    Xx = (np.arange(W - 1, dtype=np.float64)[None, :] + 0.5) / W
    Yx = np.arange(H, dtype=np.float64)[:, None] / H
    Xy = np.arange(W, dtype=np.float64)[None, :] / W
    Yy = (np.arange(H - 1, dtype=np.float64)[:, None] + 0.5) / H
    rhs = []
    for t in range(timesteps):
        sx = [(kx, ky, ph + t * dphase, a) for kx, ky, ph, a in mx]
        sy = [(kx, ky, ph + t * dphase, a) for kx, ky, ph, a in my]
        vx = mode_field((H, W - 1), Xx, Yx, sx) * flag[:, :-1] * flag[:, 1:]
        vy = mode_field((H - 1, W), Xy, Yy, sy) * flag[:-1, :] * flag[1:, :]
        rhs.append(divergence_rhs(flag, vx, vy, float(P.h)))
    return ServingTemplate(P.flag, rhs, P.h)


def serving_templates(n, seed=7, timesteps=3):
    """n served domains with sizes from serving_sizes (64x64 .. 4096x2048)."""
    return [serving_template(W, H, seed * 1000 + i, timesteps)
            for i, (W, H) in enumerate(serving_sizes(n, seed))]


def arrivals(rate, duration, seed=7):
    """Poisson arrival times (s) at `rate` per second over [0, duration)."""
    g = rng(seed)
    t, out = 0.0, []
    while True:
        t += g.exponential(1.0 / rate)
        if t >= duration:
            return np.array(out)
        out.append(t)
