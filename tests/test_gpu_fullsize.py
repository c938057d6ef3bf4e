"""Parity at the benchmark's full size: 4096x2048, 5% disks, 2/2 smoothing,
2 V-cycles per step.  At this size the smoother runs thousands of waves in
several rounds, so ordering hazards between waves (a wave reading another's
new values) would show up here and not on small grids."""
import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import workloads
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def domain():
    return workloads.config1()


def test_fullsize_classification_bit_exact(gpu, domain):
    d = domain
    mg = U.MG(d.flag)
    om = O.OracleMG(d.W, d.H)
    om.update_fields(d.flag)
    for l in range(mg.levels):
        np.testing.assert_array_equal(mg.get_flag(l), om.flag_level(l))


def test_fullsize_steps_match_oracle(gpu, domain):
    """The benchmark workload as stated: 100 timesteps of the 4096x2048
    domain with 1 M tracers; fields and tracers compared after 10 and after
    100 steps."""
    d = domain
    P = U.Params(nu_pre=2, nu_post=2, nu_coarse=5, vcycles_per_step=2)
    vx, vy = d.init_simulation_fields()
    sim = U.Simulation(d.flag, d.pwidth, d.mu, params=P)
    ref = O.OracleSim(d.flag, float(d.pwidth), float(d.mu), nu_pre=2, nu_post=2, nu_coarse=5,
                      vcycles=2)
    for name, fid, a in (("vx", O.VX, vx), ("vx_back", O.VX_BACK, vx), ("vy", O.VY, vy),
                         ("vy_back", O.VY_BACK, vy)):
        sim.set(name, a)
        ref.set(fid, a)
    xy = workloads.tracers_in_fluid(d, 1_000_000, seed=2)
    age = np.zeros(xy.shape[0], np.float32)
    sim.set_tracers(xy, age)
    rxy, rage = xy.copy(), age.copy()
    dt = float(d.dt)
    done = 0
    for upto in (10, 100):
        sim.run(dt, upto - done)  # graph-replayed steps + tracer advection (seeds 1..n)
        for k in range(done, upto):
            ref.step(dt)
            ref.advect_tracers(rxy, rage, dt, k + 1)
        done = upto
        for name, fid in (("vx", O.VX), ("vy", O.VY), ("p", O.P), ("vx_back", O.VX_BACK),
                          ("vy_back", O.VY_BACK)):
            a, b = sim.get(name), ref.get(fid)
            assert rel(a, b) < 1e-5, f"step {upto} {name}: rel L2 {rel(a, b):.3e}"
            np.testing.assert_array_equal(a, b, err_msg=f"step {upto}: {name} not bit-identical")
        gxy, gage = sim.get_tracers()
        np.testing.assert_array_equal(gxy, rxy, err_msg=f"step {upto}: tracer positions")
        np.testing.assert_array_equal(gage, rage, err_msg=f"step {upto}: tracer ages")
        print(f"step {upto}: fields and 1 M tracers bit-identical", flush=True)


def test_fullsize_smoother_sweeps_match(gpu, domain):
    """Eight level-0 sweeps with zero-gradient BC, the kernel the roofline
    figure measures, against the two-pass oracle sweep."""
    d = domain
    g = np.random.default_rng(5)
    p = g.standard_normal((d.H, d.W)).astype(np.float32)
    f = g.standard_normal((d.H, d.W)).astype(np.float32)
    mg = U.MG(d.flag)
    mg.set_level(0, 0, p)
    mg.set_level(0, 1, f)
    h = float(d.h)
    for _ in range(4):
        mg.level_op(0, 0, h)
        mg.level_op(1, 0, h)
    ref = p.copy()
    for _ in range(4):
        O.rbgs(ref, f, d.flag, h, 1.0)
        O.zero_gradient(ref)
    np.testing.assert_array_equal(mg.get_level(0, 0), ref)


@pytest.mark.parametrize("W,H,zg,nu", [(2049, 2049, True, 3), (2050, 2049, False, 3),
                                        (2049, 2049, True, 1), (2050, 2049, True, 2)])
def test_fused_level0_odd_shapes(gpu, W, H, zg, nu):
    """Levels of at least 4 M cells run the row-streamed sweeps and the fused
    residual/restriction and prolongation passes.  An odd width takes their
    unaligned load path (mask bytes loaded one by one), an odd height ends
    on a row block that holds only the last row, which zero-gradient BCs set
    from the row above (nu = 1: the prolongation pass's output is final, no
    sweep after it rewrites that row; nu = 2: three V-cycles in one call);
    three V-cycles are bit-identical to the oracle's."""
    assert W * H >= 1 << 22
    g = np.random.default_rng(W + H)
    flag = (g.random((H, W)) > 0.1).astype(np.float32)
    flag[0, :] = flag[-1, :] = 0
    flag[:, 0] = flag[:, -1] = 0
    f = g.standard_normal((H, W)).astype(np.float32)
    h = np.float32(1.0 / W)
    mg = U.MG(flag)
    mg.set_params(U.Params(nu_pre=nu, nu_post=nu, nu_coarse=5))
    om = O.OracleMG(W, H, nu_pre=nu, nu_post=nu)
    om.update_fields(flag)
    pg = np.zeros((H, W), np.float32)
    pr = pg.copy()
    if nu == 2:
        mg.set_level(0, 0, pg)
        mg.set_level(0, 1, f)
        mg.solve_resident(float(h), zg, ncycles=3)
        pg = mg.get_level(0, 0)
        for _ in range(3):
            om.solve(pr, f, flag, float(h), zg)
    else:
        for _ in range(3):
            mg.solve(pg, f, flag, float(h), zg)
            om.solve(pr, f, flag, float(h), zg)
    assert rel(pg, pr) < 1e-5
    np.testing.assert_array_equal(pg, pr)
