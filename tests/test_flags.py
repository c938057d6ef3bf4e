"""Flag producers around the path: Terrain's flags from an image
(terrain.cpp:3-45) and the simulation part of a map shift
(ubootgl_app.cpp:248-296).

The oracles (oracle/pyoracle.terrain_flags, .shift_map) are pinned by closed
forms: a white image gives all fluid, a block with one dark pixel out of 100
has mean exactly 0.99 (not fluid), rows are flipped; a shift moves a ramp one
cell left, keeps the untouched edge columns, zeroes the faces of solid cells
and sets the inlet to 0.07 H / (1 + fluid cells of column 0).  The GPU
results must equal the oracle's bit for bit, and a step after the shift must
match the oracle's step.
"""
import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import workloads
from oracle import pyoracle as O


def test_terrain_oracle_closed_forms():
    img = np.full((40, 60, 4), 255, np.uint8)
    full, sim = O.terrain_flags(img, 10)
    assert full.shape == (40, 60) and sim.shape == (4, 6)
    assert np.all(full == 1.0) and np.all(sim == 1.0)
    img[0, 0, 0] = 0  # top-left pixel of the image = last row of the grid
    full, sim = O.terrain_flags(img, 10)
    assert full[39, 0] == 0.0
    assert sim[3, 0] == 0.0 and sim.sum() == 23  # mean 0.99 is not > 0.99
    img[0, 0, 0] = 128
    _, sim = O.terrain_flags(img, 10)
    assert sim[3, 0] == 1.0


def test_shift_map_oracle_closed_forms():
    H, W = 6, 9
    x = np.arange(W, dtype=np.float32)
    f = {"vx": np.tile(x[:W - 1] + 1, (H, 1)), "vx_back": np.zeros((H, W - 1), np.float32),
         "vy": np.tile(x + 1, (H - 1, 1)), "vy_back": np.zeros((H - 1, W), np.float32),
         "p": np.tile(x + 1, (H, 1))}
    flag = np.ones((H, W), np.float32)
    g = O.shift_map(f, flag)
    assert np.all(g["vx"][:, 1:W - 2] == x[2:W - 1] + 1)      # moved left
    assert np.all(g["vx"][:, W - 2] == W - 1)                 # last column kept
    assert np.all(g["vx_back"][:, 1:W - 2] == g["vx"][:, 1:W - 2])
    assert np.all(g["p"][:, :W - 1] == x[1:] + 1) and np.all(g["p"][:, W - 1] == W)
    assert np.allclose(g["vx"][:, 0], 0.07 * H / (1 + (H - 1)))
    flag[2, 4] = 0
    g = O.shift_map(f, flag)
    assert g["vx"][2, 4] == 0 and g["vx"][2, 3] == 0 and g["vy"][2, 4] == 0
    assert g["vy"][1, 4] == 0 and g["p"][2, 4] == 0
    assert g["vx"][2, 5] != 0


@pytest.mark.gpu
def test_terrain_gpu_matches_oracle(gpu):
    g = np.random.default_rng(4)
    img = g.integers(0, 256, (123, 250, 4)).astype(np.uint8)
    img[:60, :100, 0] = 255  # fluid blocks, with a few dark pixels
    img[5, 7, 0] = 250
    for scale in (1, 3, 10):
        full, sim = U.terrain_flags(img, scale)
        rf, rs = O.terrain_flags(img, scale)
        assert np.array_equal(full, rf) and np.array_equal(sim, rs)


@pytest.mark.gpu
def test_shift_map_gpu_matches_oracle(gpu):
    d = workloads.small_domain(100, 52, seed=12)
    sim = U.Simulation(d.flag, d.pwidth, d.mu)
    vx, vy = d.init_simulation_fields()
    for name, arr in (("vx", vx), ("vx_back", vx), ("vy", vy), ("vy_back", vy)):
        sim.set(name, arr)
    dt = float(d.dt)
    for _ in range(2):
        sim.step(dt)
    before = {k: sim.get(k) for k in ("vx", "vx_back", "vy", "vy_back", "p")}
    new_flag = np.roll(d.flag, -1, axis=1)
    new_flag[:, -2] = d.flag[:, -2]  # the terrain's newly generated column
    sim.shiftMap(new_flag)
    ref = O.shift_map(before, new_flag)
    for k in ("vx", "vx_back", "vy", "vy_back", "p", "flag", "vx_current", "vy_current"):
        assert np.array_equal(sim.get(k), ref[k]), k
    # the MG classification follows the new flag, and the next step matches
    osim = O.OracleSim(new_flag, float(d.pwidth), float(d.mu))
    for fid, k in ((O.VX, "vx"), (O.VX_BACK, "vx_back"), (O.VY, "vy"), (O.VY_BACK, "vy_back"),
                   (O.P, "p")):
        osim.set(fid, ref[k])
    sim.step(dt)
    osim.step(dt)
    assert np.array_equal(sim.get("vx"), osim.get(O.VX))
    assert np.array_equal(sim.get("p"), osim.get(O.P))
