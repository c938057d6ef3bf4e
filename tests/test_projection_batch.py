"""Batched projections: the simulation's own pressure solve (MG::solve with
the zero-gradient BC, pressure_solver.cpp:250-257 as Simulation::project
calls it, simulation.cpp:189-190) on many domains, one launch per
multigrid stage (ubg_batch_vcycles), and whole steps of several
simulations with their solves batched (ubg_sims_step).

Parity contract: bit-identical to the single-domain solve (MG.solve) and to
Simulation.step on each simulation alone, which the parity tests hold
bit-identical to the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import _lib as L
from ubootgl_amd import batch as B
from ubootgl_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

# widths and heights from the batch minimum (16) up, odd and even, levels
# with and without row-streamed neighbours
SIZES = [(16, 16), (37, 23), (96, 48), (130, 70), (64, 200), (257, 129), (512, 96)]
FIELDS = ["vx", "vx_back", "vy", "vy_back", "p", "f", "vx_current", "vy_current"]


def domain(w, h, seed):
    d = W.small_domain(w, h, seed=seed, ndisks=2 if min(w, h) > 20 else 0)
    g = np.random.default_rng(seed + 7)
    f = g.standard_normal((h, w)).astype(np.float32)
    p0 = g.standard_normal((h, w)).astype(np.float32)
    return d.flag.astype(np.float32), f, p0, np.float32(1.0 / w)


@pytest.mark.parametrize("zg", [True, False])
@pytest.mark.parametrize("nu", [(2, 2, 5), (3, 3, 5), (1, 2, 4)])
def test_batched_vcycles_equal_single_solves(gpu, zg, nu):
    """Three V-cycles of every domain from a random start, batched, against
    three MG.solve calls on each domain alone (bit for bit)."""
    P = U.Params(nu_pre=nu[0], nu_post=nu[1], nu_coarse=nu[2])
    doms = [domain(w, h, 40 + i) for i, (w, h) in enumerate(SIZES)]
    bt = B.Batch(SIZES, params=P)
    for i, (flag, f, p0, h) in enumerate(doms):
        bt.set_projection_domain(i, flag, f, h)
        bt.set_pressure(i, p0)
    bt.vcycles(3, zg)
    for i, (flag, f, p0, h) in enumerate(doms):
        mg = U.MG(flag)
        mg.set_params(P)
        p = p0.copy()
        for _ in range(3):
            mg.solve(p, f, flag, float(h), zg)
        np.testing.assert_array_equal(bt.pressure(i), p, err_msg=f"domain {i} {SIZES[i]}")
        mg.close()
    bt.close()


def test_batched_vcycles_match_oracle(gpu):
    """One domain of the batch against the oracle's MG directly."""
    sizes = [(96, 48), (37, 23)]
    doms = [domain(w, h, 60 + i) for i, (w, h) in enumerate(sizes)]
    bt = B.Batch(sizes)
    for i, (flag, f, p0, h) in enumerate(doms):
        bt.set_projection_domain(i, flag, f, h)
        bt.set_pressure(i, p0)
    bt.vcycles(2, True)
    bt.vcycles(1, True)  # a second call continues from the result
    for i, (flag, f, p0, h) in enumerate(doms):
        om = O.OracleMG(*sizes[i])
        om.update_fields(flag)
        p = p0.copy()
        for _ in range(3):
            om.solve(p, f, flag, float(h), True)
        np.testing.assert_array_equal(bt.pressure(i), p)
    bt.close()


def make_sims(specs, params):
    """Pairs of identical simulations: (individual, batched)."""
    out = []
    for k, (w, h, bcs, seed) in enumerate(specs):
        d = W.small_domain(w, h, seed=seed, ndisks=2)
        vx, vy = d.init_simulation_fields()
        pair = []
        for _ in range(2):
            s = U.Simulation(d.flag, d.pwidth, d.mu, params=params)
            for name, a in (("vx", vx), ("vx_back", vx), ("vy", vy), ("vy_back", vy)):
                s.set(name, a)
            s.setBCs(*bcs)
            pair.append(s)
        out.append((pair[0], pair[1], float(d.dt)))
    return out


SPECS = [(72, 40, (0, 1, 3, 2), 71), (96, 48, (1, 0, 1, 3), 72), (130, 70, (3, 3, 0, 0), 73),
         (33, 20, (2, 2, 2, 1), 74), (200, 64, (0, 2, 3, 3), 75)]


@pytest.mark.parametrize("params", [None, (2, 2, 5, 2)])
def test_sims_step_equals_individual_steps(gpu, params):
    """Three steps of five simulations of different sizes and boundary
    kinds, their pressure solves batched, against the same simulations
    stepped one by one: every field bit-identical after each step.  A sink
    on one of them exercises the eager part of the step; a flag change
    between steps rebuilds that simulation's domain."""
    P = None if params is None else U.Params(nu_pre=params[0], nu_post=params[1],
                                             nu_coarse=params[2], vcycles_per_step=params[3])
    sims = make_sims(SPECS, P)
    sims[1][0].add_sink(0.3, 0.15, 5.0)
    sims[1][1].add_sink(0.3, 0.15, 5.0)
    sb = B.SimBatch([b for _, b, _ in sims])
    dts = [dt for _, _, dt in sims]
    for step in range(3):
        if step == 2:
            for s in (sims[2][0], sims[2][1]):
                s.setGrids(60, 30, 0.0)  # a new solid cell: the flags change
        for a, _, dt in sims:
            a.step(dt)
        sb.step(dts)
        for k, (a, b, _) in enumerate(sims):
            for name in FIELDS:
                np.testing.assert_array_equal(b.get(name), a.get(name),
                                              err_msg=f"step {step} sim {k} {name}")
    sb.close()


def test_projection_batch_errors(gpu):
    flag = W.walls(64, 48).astype(np.float32)
    f = np.zeros((48, 64), np.float32)
    bt = B.Batch([(64, 48)])
    bt.set_projection_domain(0, flag, f, 1.0 / 64)
    with pytest.raises(L.UbgError) as e:  # a walled domain is not a refinement problem
        bt.solve()
    assert e.value.code == L.E_FLAG_VALUE
    with pytest.raises(L.UbgError) as e:
        bt.vcycles(0)
    assert e.value.code == L.E_INVALID
    bt.close()
    d = W.small_domain(64, 48, seed=3)
    s1 = U.Simulation(d.flag, d.pwidth, d.mu)
    s2 = U.Simulation(d.flag, d.pwidth, d.mu, params=U.Params(nu_pre=2, nu_post=2))
    with pytest.raises(L.UbgError) as e:  # the solver parameters must agree
        B.SimBatch([s1, s2])
    assert e.value.code == L.E_INVALID
    tiny = W.small_domain(12, 12, seed=4, ndisks=0)
    s3 = U.Simulation(tiny.flag, tiny.pwidth, tiny.mu)
    with pytest.raises(L.UbgError) as e:
        B.SimBatch([s1, s3])
    assert e.value.code == L.E_DOMAIN_TOO_SMALL
    # a simulation listed twice would be advanced twice in the same launches
    with pytest.raises(L.UbgError) as e:
        B.SimBatch([s1, s1])
    assert e.value.code == L.E_INVALID
    s4 = U.Simulation(d.flag, d.pwidth, d.mu)
    sb = B.SimBatch([s1, s4])
    before = s1.get("vx")
    hs = (C.c_void_p * 2)(s1._h.value, s1._h.value)
    dts = np.full(2, d.dt, np.float32)
    with pytest.raises(L.UbgError) as e:
        L.check(L.lib().ubg_sims_step(sb._h, hs, 2, L.fptr(dts)))
    assert e.value.code == L.E_INVALID
    np.testing.assert_array_equal(s1.get("vx"), before)  # rejected before any work
    sb.close()


def test_sims_step_interleaved_with_own_work(gpu):
    """Batched steps mixed with each simulation's own work: a step of its
    own between batched steps, host writes of a velocity field and of the
    accumulation field (the next batched step runs applyAccumulatedVelocity
    on the simulation's stream first), reads right after a batched step,
    and a batched step after a destroyed-and-rebuilt batch.  Both sets stay
    bit-identical to the simulations stepped alone."""
    sims = make_sims(SPECS[:3], None)
    sb = B.SimBatch([b for _, b, _ in sims])
    dts = [dt for _, _, dt in sims]
    g = np.random.default_rng(9)

    def both(fn):
        for a, b, dt in sims:
            fn(a, dt)
            fn(b, dt)

    def check(tag):
        for k, (a, b, _) in enumerate(sims):
            for name in FIELDS:
                np.testing.assert_array_equal(b.get(name), a.get(name),
                                              err_msg=f"{tag} sim {k} {name}")

    for a, _, dt in sims:
        a.step(dt)
    sb.step(dts)
    check("first")
    both(lambda s, dt: s.step(dt))  # own steps between batched ones
    check("own step")
    for a, b, _ in sims:
        acc = (g.standard_normal(a.shape("vx_accum")) * 1e-3).astype(np.float32)
        a.set("vx_accum", acc)
        b.set("vx_accum", acc)
    for a, _, dt in sims:
        a.step(dt)
    sb.step(dts)
    check("accumulation")
    sb.close()
    sb2 = B.SimBatch([b for _, b, _ in sims])
    for a, _, dt in sims:
        a.step(dt)
    sb2.step(dts)
    check("new batch")
    sb2.close()
