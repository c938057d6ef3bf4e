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


def disks_to_fraction(W, H, g, ndisks, rmin, rmax, target, tol=0.0005):
    """`ndisks` disks with log-uniform base radii in [rmin, rmax]; the scale s
    in r = clamp(s*rho, rmin, rmax) is bisected until the rasterised solid
    fraction (interior cells) is within tol of target."""
    cx = g.uniform(0, W, ndisks)
    cy = g.uniform(0, H, ndisks)
    rho = np.exp(g.uniform(math.log(rmin), math.log(rmax), ndisks))
    lo, hi = 0.0, 1.0
    best = None
    for _ in range(40):
        s = 0.5 * (lo + hi)
        r = np.clip(s * rho, rmin, rmax)
        fl = stamp_disks(walls(W, H), cx, cy, r)
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
