"""Served solves (ubg_server_*): continuous batching with HBM-budget
admission.

Contract: each request's result (every timestep to 1e-6 relative to ||F f||,
warm-started from the previous timestep) is bit-identical whether it runs
alone or among other requests that arrive while it is on the device; its
residual, recomputed in float64 by the oracle's calculateResidualField, meets
the tolerance; the device memory in use never exceeds the budget.
"""
import time

import numpy as np
import pytest

from ubootgl_amd import _lib as L
from ubootgl_amd import batch as B
from ubootgl_amd import serve as S
from ubootgl_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (96, 160), (256, 128), (480, 96), (128, 384), (320, 320)]


@pytest.fixture(scope="module")
def templates():
    return [W.serving_template(w, h, seed=70 + i) for i, (w, h) in enumerate(SIZES)]


def run_alone(tpl, gpu_budget=2 * 10**9, steps_per_turn=1):
    srv = S.Server(hbm_budget=gpu_budget, tol=1e-6, steps_per_turn=steps_per_turn)
    gid = srv.add_geometry(tpl.flag, tpl.h)
    rid = srv.submit(gid, tpl.rhs, tpl.fnorm, want_x=True)
    info = srv.wait(rid)
    srv.close()
    return info


def test_serve_alone_equals_mix(gpu, templates):
    alone = [run_alone(t) for t in templates]
    srv = S.Server(hbm_budget=2 * 10**9, tol=1e-6)
    gids = [srv.add_geometry(t.flag, t.h) for t in templates]
    rng = np.random.default_rng(3)
    order = list(rng.permutation(len(templates))) * 3  # each template three times
    rids = []
    for k, i in enumerate(order):
        rids.append((i, srv.submit(gids[i], templates[i].rhs, templates[i].fnorm, want_x=True)))
        if k % 4 == 3:
            time.sleep(0.002)  # later requests arrive while earlier ones run
    for i, rid in rids:
        info = srv.wait(rid)
        assert info["state"] == "done" and info["converged"]
        assert info["timesteps_done"] == 3
        assert info["iterations"] == alone[i]["iterations"]
        assert np.array_equal(info["x"], alone[i]["x"]), f"template {i}"
    st = srv.stats()
    assert st["completed"] == len(order) and st["failed"] == 0
    assert st["solves"] == 3 * len(order)
    assert st["peak_active"] > 1  # requests did overlap on the device
    srv.close()


def test_serve_steps_per_turn_invariant(gpu, templates):
    """Several refinement steps per worker turn: a domain that converges
    inside a turn skips the remaining ones on the device, so results and
    iteration counts do not change."""
    for t in (templates[0], templates[3]):
        a = run_alone(t)
        b = run_alone(t, steps_per_turn=3)
        assert b["iterations"] == a["iterations"]
        assert np.array_equal(b["x"], a["x"])


def test_serve_residual_independent_check(gpu, templates):
    """Last timestep's solution against a float64 residual computed by the
    oracle: ||r|| <= tol * ||F f|| (with a little rounding slack)."""
    t = templates[2]
    info = run_alone(t)
    _, rn = O.residual(info["x"], t.rhs[-1], t.flag, float(t.h))
    assert rn <= 1.0001e-6 * t.fnorm[-1]
    assert info["rel_residual"] <= 1e-6


def test_serve_equals_sequential_warm_solves(gpu, templates):
    """A request's timesteps equal the single-domain refinement solver run
    one timestep after another, warm-started, with ref_norm = ||F f|| of each
    rhs (the criterion the server applies), bit for bit."""
    import ubootgl_amd as U
    t = templates[1]
    info = run_alone(t)
    mg = U.MG(t.flag)
    its = 0
    for k in range(len(t.rhs)):
        mg.ir_set_rhs(t.rhs[k])
        if k == 0:
            mg.ir_set_guess(None)
        r = mg.ir_solve(float(t.h), tol=1e-6, vcycles_per_iter=4, ref_norm=t.fnorm[k])
        assert r["converged"] and r["rel_residual"] <= 1e-6
        its += r["iterations"]
    x = mg.ir_solution()
    mg.close()
    assert np.array_equal(x, info["x"])
    # a warm start that already meets the tolerance takes no step
    mg = U.MG(t.flag)
    mg.ir_set_rhs(t.rhs[-1])
    mg.ir_set_guess(x)
    r = mg.ir_solve(float(t.h), tol=1e-6, ref_norm=t.fnorm[-1])
    assert r["converged"] and r["iterations"] == 0
    mg.close()


def test_serve_budget_admission(gpu, templates):
    """A budget that holds the geometries and about two requests: the rest
    queue and are admitted as memory frees up; all finish, the arena's peak
    stays within the budget."""
    t = templates[5]
    per = B.batch_bytes([(t.W, t.H)])
    budget = int(2.6 * per)
    srv = S.Server(hbm_budget=budget, tol=1e-6)
    gid = srv.add_geometry(t.flag, t.h)
    rids = [srv.submit(gid, t.rhs, t.fnorm) for _ in range(6)]
    infos = [srv.wait(r) for r in rids]
    assert all(i["state"] == "done" for i in infos)
    st = srv.stats()
    assert st["peak_bytes"] <= budget
    assert st["peak_active"] <= 2
    # admission is in arrival order
    admits = [i["t_admit"] for i in infos]
    assert admits == sorted(admits)
    srv.close()


def test_serve_errors(gpu, templates):
    srv = S.Server(hbm_budget=64 * 2**20, tol=1e-6, max_active=4)
    walls = W.walls(64, 64).astype(np.float32)
    with pytest.raises(L.UbgError) as e:
        srv.add_geometry(walls, 1 / 64)
    assert e.value.code == L.E_FLAG_VALUE
    big = templates[5]
    gid = srv.add_geometry(templates[0].flag, templates[0].h)
    with pytest.raises(L.UbgError) as e:
        srv.submit(99, templates[0].rhs)
    assert e.value.code == L.E_INVALID
    # fnorm computed by the server when not given
    rid = srv.submit(gid, templates[0].rhs, want_x=True)
    info = srv.wait(rid)
    assert info["state"] == "done" and info["converged"]
    with pytest.raises(L.UbgError) as e:
        srv.wait(12345)
    assert e.value.code == L.E_INVALID
    srv.close()
    srv = S.Server(hbm_budget=2**21, tol=1e-6)
    with pytest.raises(L.UbgError) as e:
        gid = srv.add_geometry(big.flag, big.h)
        srv.submit(gid, big.rhs)
    assert e.value.code == L.E_BUDGET
    srv.close()
