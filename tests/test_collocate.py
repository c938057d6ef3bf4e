"""Collocated velocity export (VelocityTextures::updateFromStaggered,
velocity_textures.cpp:63-93, interp_shader.cs:14-33).

The oracle (oracle/pyoracle.collocate) is pinned by closed forms: constant
fields stay constant, a field linear in the texel index is interpolated
exactly at the collocated points, and the left / bottom edges blend with the
opposite edge (GL_REPEAT of the default sampler).  The GPU image must equal
the oracle's bit for bit (the magnitude within one float32 ulp).  Hardware
texture filtering of the reference's GL path is parity unpinned.
"""
import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import workloads
from oracle import pyoracle as O


def test_collocate_constant_and_linear():
    W, H = 13, 9
    vxy, mag = O.collocate(np.full((H, W - 1), 2.0), np.full((H - 1, W), -3.0))
    assert np.all(vxy[..., 0] == 2.0) and np.all(vxy[..., 1] == -3.0)
    assert np.allclose(mag, np.hypot(2.0, 3.0))
    # vx(i, j) = i + 100 j: pixel (gx, gy) samples texel point ((gx-1)/2, gy/2)
    i = np.arange(W - 1, dtype=np.float32)[None, :]
    j = np.arange(H, dtype=np.float32)[:, None]
    vx = i + 100 * j
    vy = np.zeros((H - 1, W), np.float32)
    vxy, _ = O.collocate(vx, vy)
    gx = np.arange(2 * W - 1)
    gy = np.arange(0, 2 * H - 1, 2)  # even rows sit on texel rows
    for g in gx[1:-1]:
        expect = (g - 1) / 2 + 100 * gy / 2
        assert np.allclose(vxy[gy, g, 0], expect)
    # gx = 0: half way between texel -1 (wraps to W-2) and texel 0; the last
    # column, texel point W - 1.5, between texel W-2 and W-1 (wraps to 0)
    assert np.allclose(vxy[gy, 0, 0], (W - 2 + 0) / 2 + 100 * gy / 2)
    assert np.allclose(vxy[gy, -1, 0], (W - 2 + 0) / 2 + 100 * gy / 2)
    # odd rows: midway between rows, except the last is never odd
    assert np.allclose(vxy[1, 1, 0], 0.0 + 100 * 0.5)


@pytest.mark.gpu
def test_collocate_gpu_matches_oracle(gpu):
    d = workloads.small_domain(96, 40, seed=5)
    sim = U.Simulation(d.flag, d.pwidth, d.mu)
    g = np.random.default_rng(1)
    vxc = g.standard_normal((40, 95)).astype(np.float32)
    vyc = g.standard_normal((39, 96)).astype(np.float32)
    sim.set("vx_current", vxc)
    sim.set("vy_current", vyc)
    vxy, mag = sim.collocated()
    rvxy, rmag = O.collocate(vxc, vyc)
    assert np.array_equal(vxy, rvxy)
    ulp = np.spacing(np.maximum(np.abs(rmag), np.float32(1e-30)))
    assert np.all(np.abs(mag - rmag) <= ulp)
    # after real steps, the image of the saved current fields
    vx, vy = d.init_simulation_fields()
    for name, arr in (("vx", vx), ("vx_back", vx), ("vy", vy), ("vy_back", vy)):
        sim.set(name, arr)
    for _ in range(2):
        sim.step(float(d.dt))
    vxy, mag = sim.collocated()
    rvxy, rmag = O.collocate(sim.get("vx_current"), sim.get("vy_current"))
    assert np.array_equal(vxy, rvxy)
    assert np.all(np.abs(mag - rmag) <= np.spacing(np.maximum(np.abs(rmag), np.float32(1e-30))))
