"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must reproduce them exactly (CPU), and the HIP
path must match them to the parity tolerance (GPU)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEP_FIXTURES = ["step_64x32.npz", "step_100x52_nu22.npz"]


def load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_step(z):
    nu_pre, nu_post, nu_coarse, vcycles = (int(v) for v in z["params"])
    s = O.OracleSim(z["flag"], float(z["pwidth"]), float(z["mu"]), nu_pre=nu_pre,
                    nu_post=nu_post, nu_coarse=nu_coarse, vcycles=vcycles)
    for fid, a in ((O.VX, z["vx0"]), (O.VX_BACK, z["vx0"]), (O.VY, z["vy0"]),
                   (O.VY_BACK, z["vy0"])):
        s.set(fid, a)
    for _ in range(int(z["nsteps"])):
        s.step(float(z["dt"]))
    return s


@pytest.mark.parametrize("name", STEP_FIXTURES)
def test_oracle_reproduces_step_fixture(name):
    z = load(name)
    s = oracle_step(z)
    for fid, key in ((O.VX, "vx"), (O.VY, "vy"), (O.P, "p"), (O.VX_BACK, "vx_back")):
        np.testing.assert_array_equal(s.get(fid), z[key])


def test_oracle_reproduces_mg_fixture():
    z = load("mg_67x41.npz")
    H, W = z["flag"].shape
    mg = O.OracleMG(W, H)
    mg.update_fields(z["flag"])
    p = z["p0"].copy()
    for _ in range(int(z["ncycles"])):
        mg.solve(p, z["f"], z["flag"], float(z["h"]), True)
    np.testing.assert_array_equal(p, z["p"])
    np.testing.assert_array_equal(mg.flag_level(1), z["flagc1"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", STEP_FIXTURES)
def test_gpu_matches_step_fixture(gpu, name):
    import ubootgl_amd as U
    z = load(name)
    nu_pre, nu_post, nu_coarse, vcycles = (int(v) for v in z["params"])
    sim = U.Simulation(z["flag"], float(z["pwidth"]), float(z["mu"]),
                       params=U.Params(nu_pre=nu_pre, nu_post=nu_post, nu_coarse=nu_coarse,
                                       vcycles_per_step=vcycles))
    for name_, key in (("vx", "vx0"), ("vx_back", "vx0"), ("vy", "vy0"), ("vy_back", "vy0")):
        sim.set(name_, z[key])
    for _ in range(int(z["nsteps"])):
        sim.step(float(z["dt"]))
    for key in ("vx", "vy", "p", "vx_back", "vy_back"):
        assert rel(sim.get(key), z[key]) < 1e-5, key
        np.testing.assert_array_equal(sim.get(key), z[key])


@pytest.mark.gpu
def test_gpu_matches_mg_fixture(gpu):
    import ubootgl_amd as U
    z = load("mg_67x41.npz")
    mg = U.MG(z["flag"])
    np.testing.assert_array_equal(mg.get_flag(1), z["flagc1"])
    np.testing.assert_array_equal(mg.get_flag(2), z["flagc2"])
    p = z["p0"].copy()
    for _ in range(int(z["ncycles"])):
        mg.solve(p, z["f"], z["flag"], float(z["h"]), True)
    np.testing.assert_array_equal(p, z["p"])
