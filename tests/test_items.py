"""Floating items coupled to the fluid (advect_floating_items.cpp:16-274).

Items are integrated one per thread; their state must equal the oracle's bit
for bit (both use the Cephes sin/cos and IEEE-rounded basic operations in the
reference's order).  The reaction scattered into vx_accum / vy_accum is summed
with float atomics, so those fields are compared within a relative tolerance
of 1e-5 (summation order only).

The oracle is pinned by: the Cephes sin/cos against math.sin/cos (<= 2 ulp);
the out-of-domain rule (an item outside the domain only moves and turns:
closed form); the collision response (an item driven into a solid wall
bumps and has its normal velocity reflected); the sign of the scatter (an
item moving through still fluid pushes the fluid ahead of it forward).
"""
import math

import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import _lib as L
from ubootgl_amd import workloads
from oracle import pyoracle as O

ACC_TOL = 1e-5


PWIDTH = 0.8  # the game's simulation width (ubootgl_app.cpp:30)


def make_items(n, seed, W, H, simple=False):
    """Items at the game's physical scales: ships / torpedoes of 0.004-0.009 x
    0.0008-0.0022 with mass 0.6-1.3 (ubootgl_app.cpp:94-96, torpedo.cpp:11-16);
    debris 0.0003 s^2 with mass 0.1 s (0.15..0.45), s in 1..3
    (explosion.cpp:36-52)."""
    g = np.random.default_rng(seed)
    it = np.zeros(n, L.ITEM_DTYPE)
    cell = PWIDTH / W
    if simple:
        s2 = g.uniform(1, 3, n) ** 2
        it["size"] = (0.0003 * s2)[:, None]
        it["mass"] = 0.1 * s2 * g.uniform(0.15, 0.45, n)
    else:
        it["size"][:, 0] = g.uniform(0.004, 0.009, n)
        it["size"][:, 1] = g.uniform(0.0008, 0.0022, n)
        it["mass"] = g.uniform(0.6, 1.3, n)
    it["pos"][:, 0] = g.uniform(4, W - 5, n) * cell
    it["pos"][:, 1] = g.uniform(4, H - 5, n) * cell
    it["rotation"] = g.uniform(0, 2 * math.pi, n)
    it["vel"] = g.normal(0, 0.05, (n, 2))
    it["ang_vel"] = g.normal(0, 0.5, n)
    return it


def test_cephes_trig():
    xs = np.linspace(-12.0, 12.0, 20001).astype(np.float32)
    for f, ref in ((O.cephes_sin, math.sin), (O.cephes_cos, math.cos)):
        for x in xs[::7]:
            got = np.float32(f(float(x)))
            want = np.float32(ref(float(x)))
            ulp = np.spacing(np.float32(max(abs(float(want)), 1e-6)))
            assert abs(float(got) - float(want)) <= 2 * float(ulp), (x, got, want)
    assert O.cephes_sin(0.0) == 0.0 and O.cephes_cos(0.0) == 1.0


def test_items_oracle_out_of_domain_closed_form():
    d = workloads.small_domain(64, 32, seed=1)
    sim = O.OracleSim(d.flag, float(d.pwidth), float(d.mu))
    it = np.zeros(1, L.ITEM_DTYPE)
    it["pos"] = (-1.0, -1.0)  # outside: the sub-steps only move and turn it
    it["vel"] = (0.5, 0.25)
    it["ang_vel"] = 1.5
    it["mass"] = 1.0
    it["size"] = (0.1, 0.1)
    it["force"] = (3.0, 4.0)
    dt = 0.01
    steps = int(min(15.0, max(1.0, float(np.float32(0.5) * np.float32(dt) / np.float32(sim.h)) * 2.5)))
    sim.advect_items(it, dt)
    sub = np.float32(dt) / np.float32(steps)
    x = np.float32(-1.0)
    rot = 0.0
    for _ in range(steps):
        x = np.float32(x + np.float32(sub * np.float32(0.5)))
        rot = float(np.float32(math.fmod(float(np.float32(rot) + np.float32(sub * np.float32(1.5))) + 2 * math.pi,
                                         2 * math.pi)))
    assert it["pos"][0, 0] == x
    assert it["rotation"][0] == np.float32(rot)
    assert tuple(it["vel"][0]) == (0.5, 0.25) and it["bump_count"][0] == 0
    assert tuple(it["force"][0]) == (0.0, 0.0) and it["ang_force"][0] == 0.0


def test_items_oracle_collision_and_scatter_sign():
    W, H = 96, 48
    flag = workloads.walls(W, H).astype(np.float32)
    flag[:12, :] = 0  # a floor
    sim = O.OracleSim(flag, PWIDTH, 0.001)
    sim.set(O.VX, np.zeros((H, W - 1), np.float32))
    sim.set(O.VY, np.zeros((H - 1, W), np.float32))
    cell = PWIDTH / W
    it = np.zeros(2, L.ITEM_DTYPE)
    it["size"] = (0.009, 0.0022)
    it["mass"] = 1.3
    it["pos"][0] = (40 * cell, 12.4 * cell)   # falling onto the floor (top at 12 cells)
    it["vel"][0] = (0.0, -0.3)
    it["pos"][1] = (30 * cell, 30 * cell)     # moving right through still fluid
    it["vel"][1] = (0.2, 0.0)
    sim.advect_items(it, 1.0 / 60)
    assert it["bump_count"][0] > 0 and it["vel"][0, 1] > 0  # reflected upward
    acc = sim.get(O.VX_ACCUM)
    assert acc[25:36, 25:40].sum() > 0  # fluid ahead pushed forward


@pytest.mark.gpu
@pytest.mark.parametrize("simple", [False, True])
def test_items_gpu_matches_oracle(gpu, simple):
    d = workloads.small_domain(160, 80, seed=7, ndisks=4)
    sim = U.Simulation(d.flag, PWIDTH, d.mu)
    ref = O.OracleSim(d.flag, PWIDTH, float(d.mu))
    vx, vy = d.init_simulation_fields()
    vx, vy = vx * np.float32(0.05), vy * np.float32(0.05)  # the game's flow speeds
    for name, fid, arr in (("vx", O.VX, vx), ("vx_back", O.VX_BACK, vx), ("vy", O.VY, vy),
                           ("vy_back", O.VY_BACK, vy)):
        sim.set(name, arr)
        ref.set(fid, arr)
    dt = float(d.dt)
    for _ in range(2):
        sim.step(dt)
        ref.step(dt)
    items = make_items(96, 3, 160, 80, simple)
    items[-1]["pos"] = (np.nan, np.nan)  # a non-finite item is left where it is
    items_ref = items.copy()
    for frame in range(3):
        sim.advectFloatingItems(items, 1.0 / 60, simple=simple)
        ref.advect_items(items_ref, 1.0 / 60, simple=simple)
        for k in ("pos", "vel", "rotation", "ang_vel", "bump_count", "force", "ang_force"):
            assert np.array_equal(items[k], items_ref[k], equal_nan=True), (frame, k)
        assert np.isnan(items["pos"][-1]).all() and np.isfinite(items["pos"][:-1]).all()
        for name, fid in (("vx_accum", O.VX_ACCUM), ("vy_accum", O.VY_ACCUM)):
            a, b = sim.get(name), ref.get(fid)
            scale = max(float(np.abs(b).max()), 1e-30)
            assert float(np.abs(a - b).max()) <= ACC_TOL * scale, (frame, name)
    # the accumulated reaction enters the next step (applyAccumulatedVelocity)
    sim.step(dt)
    ref.step(dt)
    a, b = sim.get("vx"), ref.get(O.VX)
    assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b)
