"""GPU parity: libubg.so against the CPU oracle on identical seeded inputs.

The parity contract is: flags and classification bit-exact, velocity and
pressure within 1e-5 relative L2.  Both sides are built without FMA
contraction and perform the same IEEE operations in the same order, so the
fields are in fact bit-identical; these tests assert that (exact equality),
which implies the 1e-5 bound.
"""
import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import _lib as L
from ubootgl_amd import workloads
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def rand_flag(W, H, seed, frac=0.2):
    g = np.random.default_rng(seed)
    f = (g.random((H, W)) > frac).astype(np.float32)
    f[0, :] = f[-1, :] = 0
    f[:, 0] = f[:, -1] = 0
    return f


SIZES = [(8, 8), (16, 9), (33, 17), (64, 32), (100, 60), (257, 129)]


@pytest.mark.parametrize("W,H", SIZES)
def test_classification_bit_exact(gpu, W, H):
    flag = rand_flag(W, H, 1, 0.35)
    mg = U.MG(flag)
    om = O.OracleMG(W, H)
    om.update_fields(flag)
    assert mg.levels == len(om.sizes) == len(U.level_sizes(W, H))
    for l in range(mg.levels):
        np.testing.assert_array_equal(mg.get_flag(l), om.flag_level(l))


@pytest.mark.parametrize("W,H", [(16, 9), (64, 32), (257, 129)])
def test_rbgs_sweep(gpu, W, H):
    g = np.random.default_rng(2)
    flag = rand_flag(W, H, 3)
    p = g.standard_normal((H, W)).astype(np.float32)
    f = g.standard_normal((H, W)).astype(np.float32)
    h = np.float32(1.0 / W)
    mg = U.MG(flag)
    mg.set_level(0, L.LEVEL_P, p)
    mg.set_level(0, L.LEVEL_F, f)
    mg.rbgs(float(h), 1.0, 3)
    got = mg.get_level(0, L.LEVEL_P)
    ref = p.copy()
    for _ in range(3):
        O.rbgs(ref, f, flag, float(h), 1.0)
    np.testing.assert_array_equal(got, ref)


def test_rbgs_alpha(gpu):
    W, H = 48, 40
    g = np.random.default_rng(4)
    flag = rand_flag(W, H, 5)
    p = g.standard_normal((H, W)).astype(np.float32)
    f = g.standard_normal((H, W)).astype(np.float32)
    mg = U.MG(flag)
    mg.set_level(0, L.LEVEL_P, p)
    mg.set_level(0, L.LEVEL_F, f)
    mg.level_op(L.MGOP_RBGS, 0, 0.02, 0.7)
    ref = O.rbgs(p.copy(), f, flag, 0.02, 0.7)
    np.testing.assert_array_equal(mg.get_level(0, L.LEVEL_P), ref)


@pytest.mark.parametrize("W,H", [(33, 17), (128, 64)])
def test_residual_restrict_prolong(gpu, W, H):
    g = np.random.default_rng(6)
    flag = rand_flag(W, H, 7)
    p = g.standard_normal((H, W)).astype(np.float32)
    f = g.standard_normal((H, W)).astype(np.float32)
    h = 1.0 / W
    mg = U.MG(flag)
    om = O.OracleMG(W, H)
    om.update_fields(flag)
    mg.set_level(0, L.LEVEL_P, p)
    mg.set_level(0, L.LEVEL_F, f)
    # residual field
    r_gpu, n_gpu = mg.residual(h)
    r_ref, n_ref = O.residual(p, f, flag, h)
    np.testing.assert_array_equal(r_gpu, r_ref)
    assert abs(n_gpu - n_ref) <= 1e-12 * n_ref
    # residual + restrict
    mg.level_op(L.MGOP_RESTRICT_RESIDUAL, 0, h)
    Wc, Hc = mg.sizes[1]
    rc_ref = O.restrict(r_ref, Wc, Hc)
    np.testing.assert_array_equal(mg.get_level(1, L.LEVEL_F), rc_ref)
    assert not mg.get_level(1, L.LEVEL_P).any()
    # prolongate + correct with a random coarse correction
    ec = g.standard_normal((Hc, Wc)).astype(np.float32)
    ec[0, :] = ec[-1, :] = 0
    ec[:, 0] = ec[:, -1] = 0
    mg.set_level(1, L.LEVEL_P, ec)
    mg.level_op(L.MGOP_PROLONG_CORRECT, 0, h)
    e = O.prolongate(W, H, ec, om.flag_level(1), flag)
    ref = O.correct(p.copy(), e)
    np.testing.assert_array_equal(mg.get_level(0, L.LEVEL_P), ref)


def test_zero_gradient(gpu):
    W, H = 40, 24
    g = np.random.default_rng(8)
    p = g.standard_normal((H, W)).astype(np.float32)
    mg = U.MG(W, H)
    mg.set_level(0, L.LEVEL_P, p)
    mg.level_op(L.MGOP_ZERO_GRADIENT, 0, 1.0)
    np.testing.assert_array_equal(mg.get_level(0, L.LEVEL_P), O.zero_gradient(p.copy()))


@pytest.mark.parametrize("W,H,zg", [(8, 8, False), (64, 32, True), (100, 37, True),
                                    (256, 256, False), (512, 256, True)])
def test_vcycle(gpu, W, H, zg):
    g = np.random.default_rng(9)
    flag = rand_flag(W, H, 10, 0.1)
    f = g.standard_normal((H, W)).astype(np.float32)
    p = np.zeros((H, W), np.float32)
    h = np.float32(1.0 / W)
    mg = U.MG(flag)
    om = O.OracleMG(W, H)
    om.update_fields(flag)
    pg, pr = p.copy(), p.copy()
    for _ in range(3):
        mg.solve(pg, f, flag, float(h), zg)
        om.solve(pr, f, flag, float(h), zg)
    assert rel(pg, pr) < TOL
    np.testing.assert_array_equal(pg, pr)


def test_mgtest_kat_gpu(gpu):
    """mgtest/mgtest.cpp:10-53 on the GPU: Dirichlet harmonic problem."""
    N = 257
    h = np.float32(1.0 / (N - 1))
    u = np.zeros((N, N), np.float32)
    x = np.arange(N)
    u[N - 1, 1:N - 1] = (np.sinh(np.pi) * np.sin(x[1:N - 1] / (N - 1.0) * np.pi)).astype(np.float32)
    rhs = np.zeros_like(u)
    flag = np.ones_like(u)
    mg = U.MG(N, N)
    om = O.OracleMG(N, N)
    ug, ur = u.copy(), u.copy()
    for _ in range(5):
        mg.solve(ug, rhs, flag, float(h))
        om.solve(ur, rhs, flag, float(h))
    np.testing.assert_array_equal(ug, ur)


def make_pair(d, params=None, **oparams):
    vx, vy = d.init_simulation_fields()
    sim = U.Simulation(d.flag, d.pwidth, d.mu, params=params)
    ref = O.OracleSim(d.flag, float(d.pwidth), float(d.mu), **oparams)
    for name, fid, arr in (("vx", O.VX, vx), ("vx_back", O.VX_BACK, vx),
                           ("vy", O.VY, vy), ("vy_back", O.VY_BACK, vy)):
        sim.set(name, arr)
        ref.set(fid, arr)
    return sim, ref


FIELDS = [("vx", O.VX), ("vx_back", O.VX_BACK), ("vy", O.VY), ("vy_back", O.VY_BACK),
          ("p", O.P), ("f", O.F)]


# f is the divergence of the advected velocity: a difference of nearly equal
# face values, so fp32 rounding differences in v are amplified there.  The
# parity contract is on velocity and pressure (1e-5); f is held to 1e-4.
F_TOL = 1e-4


def compare(sim, ref, fields=FIELDS, tol=TOL):
    for name, fid in fields:
        a, b = sim.get(name), ref.get(fid)
        r = rel(a, b)
        t = F_TOL if name == "f" else tol
        assert r < t, f"{name}: rel L2 {r:.3e}"
        np.testing.assert_array_equal(a, b, err_msg=f"{name} not bit-identical")


@pytest.mark.parametrize("stage", [L.STAGE_DIFFUSE, L.STAGE_ADVECT, L.STAGE_SET_VBCS,
                                   L.STAGE_DIVERGENCE, L.STAGE_MG_SOLVE, L.STAGE_PROJECT])
def test_stage(gpu, stage):
    d = workloads.small_domain(96, 48, seed=11, ndisks=3)
    sim, ref = make_pair(d)
    # put the state somewhere non-trivial first: one full step on both
    sim.step(float(d.dt))
    ref.step(float(d.dt))
    sim.stage(stage, float(d.dt))
    ref.stage(stage, float(d.dt))
    compare(sim, ref)


def test_flag_roundtrip_and_setgrids(gpu):
    d = workloads.small_domain(40, 24, seed=12)
    sim = U.Simulation(d.flag, d.pwidth, d.mu)
    np.testing.assert_array_equal(sim.get("flag"), d.flag)
    sim.setGrids(5, 6, 0.0)
    fl = d.flag.copy()
    fl[6, 5] = 0
    np.testing.assert_array_equal(sim.get("flag"), fl)
    with pytest.raises(U.UbgError):
        sim.setGrids(1, 1, 0.5)


@pytest.mark.parametrize("W,H,nsteps", [(64, 32, 3), (96, 48, 5), (130, 70, 4)])
def test_step_small(gpu, W, H, nsteps):
    d = workloads.small_domain(W, H, seed=W + H, ndisks=3)
    sim, ref = make_pair(d)
    for _ in range(nsteps):
        sim.step(float(d.dt))
        ref.step(float(d.dt))
    compare(sim, ref, FIELDS + [("vx_current", O.VX_CURRENT), ("vy_current", O.VY_CURRENT)])


# (west, east, north, south): every kind on every side at least once
BC_SETS = [(0, 1, 3, 2), (1, 0, 1, 3), (3, 3, 0, 0), (2, 2, 2, 1)]


@pytest.mark.parametrize("bcs", BC_SETS)
def test_step_bc_kinds(gpu, bcs):
    """Steps under every boundary-condition kind: the fused passes write the
    boundary cells exactly as setVBCs / setPBC do."""
    d = workloads.small_domain(72, 40, seed=17, ndisks=3)
    sim, ref = make_pair(d)
    sim.setBCs(*bcs)
    ref.set_bcs(*bcs)
    for _ in range(3):
        sim.step(float(d.dt))
        ref.step(float(d.dt))
    compare(sim, ref, FIELDS + [("vx_current", O.VX_CURRENT), ("vy_current", O.VY_CURRENT)])


@pytest.mark.parametrize("W,H,bcs", [(8, 8, (0, 1, 3, 2)), (9, 11, (1, 0, 1, 3)),
                                     (13, 8, (3, 3, 0, 0)), (17, 9, (2, 2, 2, 1)),
                                     (33, 20, (0, 2, 3, 3))])
def test_step_tiny_and_odd(gpu, W, H, bcs):
    """The smallest and odd domains: boundary duties of neighbouring edges
    and corners meet, tiles are partial, the hierarchy is 1-3 levels."""
    d = workloads.small_domain(W, H, seed=W * 7 + H, ndisks=1 if min(W, H) > 12 else 0)
    sim, ref = make_pair(d)
    sim.setBCs(*bcs)
    ref.set_bcs(*bcs)
    for _ in range(3):
        sim.step(float(d.dt))
        ref.step(float(d.dt))
    compare(sim, ref, FIELDS + [("vx_current", O.VX_CURRENT), ("vy_current", O.VY_CURRENT)])


def test_step_after_host_writes(gpu):
    """Accumulated velocity written by the host, then a step; a velocity
    written by the host, then the diffuse stage alone: vy's boundary is
    refreshed first, as the reference's setVBCs after the first vx
    iteration does."""
    d = workloads.small_domain(64, 40, seed=18, ndisks=2)
    sim, ref = make_pair(d)
    sim.setBCs(1, 1, 1, 1)
    ref.set_bcs(1, 1, 1, 1)
    g = np.random.default_rng(5)
    ax = g.normal(0, 0.2, (40, 63)).astype(np.float32)
    ay = g.normal(0, 0.2, (39, 64)).astype(np.float32)
    sim.set("vx_accum", ax)
    sim.set("vy_accum", ay)
    ref.set(O.VX_ACCUM, ax)
    ref.set(O.VY_ACCUM, ay)
    for _ in range(2):
        sim.step(float(d.dt))
        ref.step(float(d.dt))
    compare(sim, ref)
    vy = g.normal(0, 0.5, (39, 64)).astype(np.float32)
    sim.set("vy", vy)
    ref.set(O.VY, vy)
    sim.stage(L.STAGE_DIFFUSE, float(d.dt))
    ref.stage(L.STAGE_DIFFUSE, float(d.dt))
    compare(sim, ref)
    sim.step(float(d.dt))
    ref.step(float(d.dt))
    compare(sim, ref)


def test_step_config0(gpu):
    """512x256 config: one step, 10 V-cycles, 3/3/5 smoothing."""
    d = workloads.config0()
    P = U.Params(vcycles_per_step=10)
    sim, ref = make_pair(d, params=P, vcycles=10)
    sim.step(float(d.dt))
    ref.step(float(d.dt))
    compare(sim, ref)


def test_step_eager_equals_graph(gpu):
    d = workloads.small_domain(80, 40, seed=13)
    a, _ = make_pair(d, params=U.Params(use_graph=1))
    b, _ = make_pair(d, params=U.Params(use_graph=0))
    for _ in range(4):
        a.step(float(d.dt))
        b.step(float(d.dt))
    for name in ("vx", "vy", "p", "vx_back", "vy_back"):
        np.testing.assert_array_equal(a.get(name), b.get(name))


def test_sinks(gpu):
    d = workloads.small_domain(64, 48, seed=14)
    sim, ref = make_pair(d)
    for s in (sim,):
        s.add_sink(0.3, 0.2, 5.0)
    ref.add_sink(0.3, 0.2, 5.0)
    for _ in range(3):
        sim.step(float(d.dt))
        ref.step(float(d.dt))
    compare(sim, ref)
    assert sim.sinks().shape[0] == ref.num_sinks()


def test_bilinear_vel(gpu):
    d = workloads.small_domain(64, 40, seed=15)
    sim, ref = make_pair(d)
    sim.step(float(d.dt))
    ref.step(float(d.dt))
    g = np.random.default_rng(16)
    c = np.stack([g.uniform(-2, 66, 500), g.uniform(-2, 42, 500)], 1).astype(np.float32)
    np.testing.assert_array_equal(sim.bilinearVel(c), ref.bilinear_vel(c))


def test_tracers(gpu):
    d = workloads.small_domain(64, 40, seed=17, ndisks=4)
    sim, ref = make_pair(d)
    xy = workloads.tracers_in_fluid(d, 2000, seed=18)
    age = np.linspace(0, 6.5, xy.shape[0]).astype(np.float32)
    sim.set_tracers(xy, age)
    rxy, rage = xy.copy(), age.copy()
    for k in range(5):
        sim.step(float(d.dt))
        ref.step(float(d.dt))
        sim.advect_tracers(float(d.dt), 100 + k)
        ref.advect_tracers(rxy, rage, float(d.dt), 100 + k)
    gxy, gage = sim.get_tracers()
    np.testing.assert_array_equal(gxy, rxy)
    np.testing.assert_array_equal(gage, rage)


@pytest.mark.parametrize("use_graph", [1, 0])
def test_run_tracers_equal_step_by_step(gpu, use_graph):
    """Simulation.run (steps with the tracer advection after each, seeds
    continuing across calls) equals step() + advect_tracers() with seeds 1,
    2, ... -- fields and tracers, graph-replayed and eager."""
    d = workloads.small_domain(96, 48, seed=19, ndisks=5)
    prm = U.Params(use_graph=use_graph)
    a, _ = make_pair(d, params=prm)
    b, _ = make_pair(d, params=prm)
    xy = workloads.tracers_in_fluid(d, 3000, seed=20)
    age = np.linspace(0, 6.5, xy.shape[0]).astype(np.float32)
    a.set_tracers(xy, age)
    b.set_tracers(xy, age)
    dt = float(d.dt)
    a.run(dt, 6)
    a.run(dt, 4)  # a second run continues the seeds
    for k in range(10):
        b.step(dt)
        b.advect_tracers(dt, 1 + k)
    for name in ("vx", "vy", "p", "vx_back", "vy_back", "vx_current", "vy_current"):
        np.testing.assert_array_equal(a.get(name), b.get(name), err_msg=name)
    axy, aage = a.get_tracers()
    bxy, bage = b.get_tracers()
    np.testing.assert_array_equal(axy, bxy)
    np.testing.assert_array_equal(aage, bage)


def test_error_codes(gpu):
    with pytest.raises(U.UbgError) as e:
        U.MG(7, 100)
    assert e.value.code == L.E_DOMAIN_TOO_SMALL
    bad = np.ones((16, 16), np.float32) * 0.5
    with pytest.raises(U.UbgError) as e:
        U.Simulation(bad, 0.9, 0.001)
    assert e.value.code == L.E_FLAG_VALUE


def test_advect_flat_loads_bit_exact(gpu, monkeypatch):
    """advect's 64-bit flat-load path (used for velocity grids of 2 GiB or
    more) forced on a small grid: bit-identical to the oracle, as the
    buffer-load path is."""
    monkeypatch.setenv("UBG_ADVECT_FLAT_LOADS", "1")
    d = workloads.small_domain(96, 48, seed=51, ndisks=3)
    sim, ref = make_pair(d)
    for _ in range(2):
        sim.step(float(d.dt))
        ref.step(float(d.dt))
    sim.stage(L.STAGE_ADVECT, float(d.dt))
    ref.stage(L.STAGE_ADVECT, float(d.dt))
    compare(sim, ref)


@pytest.mark.parametrize("rows,nu", [(r, 3) for r in range(8, 37, 2)] + [(38, 2), (40, 2)])
def test_tile_pass_row_heights(gpu, monkeypatch, rows, nu):
    """Every tile row height the LDS level passes can pick (even, 8..40; the
    64 KiB of LDS a tile may use allow 36 rows at nu = 3 and 40 at nu = 2),
    forced on an odd-sized hierarchy whose levels all run as tiles: three
    V-cycles bit-identical to the oracle's."""
    monkeypatch.setenv("UBG_TILE_ROWS", str(rows))
    Wd, Hd = 1001, 603
    g = np.random.default_rng(52)
    flag = rand_flag(Wd, Hd, 53, 0.1)
    f = g.standard_normal((Hd, Wd)).astype(np.float32)
    h = float(np.float32(1.0 / Wd))
    mg = U.MG(flag)
    mg.set_params(U.Params(nu_pre=nu, nu_post=nu, nu_coarse=5))
    om = O.OracleMG(Wd, Hd, nu_pre=nu, nu_post=nu)
    om.update_fields(flag)
    pg = np.zeros((Hd, Wd), np.float32)
    pr = pg.copy()
    for _ in range(3):
        mg.solve(pg, f, flag, h, False)
        om.solve(pr, f, flag, h, False)
    np.testing.assert_array_equal(pg, pr)
    mg.close()
