"""The batched and served workloads at their stated sizes, checked off the
GPU in float64.

* Batched solves: 256 independent domains of uneven size (W log-uniform in
  128..2048, multiples of 32, aspect 1..4, 0-8% disks, own 8-mode field),
  one batch (autotuned, graph-replayed) to 1e-6 relative residual.  Every
  domain's residual is recomputed by the oracle's float64
  calculateResidualField (pressure_solver.cpp:91-116) from the returned
  solution; eight sampled domains are compared with the FP64 CPU multigrid's
  solution solved to 1e-13.
* Served solves: a Poisson arrival stream at 2000 requests/s (seed 7) over
  48 domain templates from 64x64 up to 4096x2048, each request three
  warm-started solves to 1e-6 of ||F f||, under a 200 GB HBM budget.  Every
  request's last solution is checked with the oracle's float64 residual, and
  a sample is re-run alone and must be bit-identical to its run in the mix.
"""
import time

import numpy as np
import pytest

from ubootgl_amd import batch as B
from ubootgl_amd import serve as S
from ubootgl_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def batch_problems():
    t = time.time()
    probs = W.config2(256)
    print(f"256 domains, {sum(P.cells for P in probs) / 1e6:.1f} M cells, generated in "
          f"{time.time() - t:.1f}s", flush=True)
    return probs


def test_batched_256_uneven_domains_float64_residuals(gpu, batch_problems):
    probs = batch_problems
    bt = B.Batch([(P.W, P.H) for P in probs])
    for d, P in enumerate(probs):
        bt.set_problem(d, P)
        bt.set_guess(d)
    tune = bt.autotune()
    print(f"autotune: {tune}", flush=True)
    reps = bt.solve(tol=1e-6, vcycles_per_iter=4)
    assert reps >= 1
    worst = 0.0
    for d, P in enumerate(probs):
        res = bt.result(d)
        assert res["converged"], f"domain {d} {P.W}x{P.H}: {res}"
        x = bt.solution(d)
        # ||r_0|| of the zero start = ||F f|| on the interior
        _, r0 = O.residual(np.zeros((P.H, P.W)), P.f, P.flag, float(P.h))
        _, rn = O.residual(x, P.f, P.flag, float(P.h))
        assert rn <= 1.0001e-6 * r0, f"domain {d} {P.W}x{P.H}: {rn / r0:.3e}"
        worst = max(worst, rn / r0)
    print(f"worst float64 relative residual over 256 domains: {worst:.3e}", flush=True)
    # eight domains against the FP64 CPU multigrid solved to 1e-13: a 1e-6
    # residual leaves a relative error of a few 1e-6 (measured on the CPU
    # for 512x256 .. 2048x1024); 1e-4 is the bound asserted
    g = np.random.default_rng(9)
    small = [d for d, P in enumerate(probs) if P.cells <= 400_000]
    for d in g.choice(small, 8, replace=False):
        P = probs[d]
        xr, hist = O.solve_poisson_fp64(P.flag, P.f, float(P.h), tol=1e-13)
        e = rel(bt.solution(d), xr)
        print(f"domain {d} {P.W}x{P.H}: rel error vs FP64 oracle {e:.2e}", flush=True)
        assert e < 1e-4
    bt.close()


@pytest.fixture(scope="module")
def serving_templates():
    t = time.time()
    tpls = W.serving_templates(48, seed=7)
    print(f"48 templates up to {max(t.cells for t in tpls) / 1e6:.1f} M cells, generated in "
          f"{time.time() - t:.1f}s", flush=True)
    return tpls


def _alone(tpl):
    srv = S.Server(hbm_budget=200 * 10**9, tol=1e-6, vcycles_per_iter=4)
    gid = srv.add_geometry(tpl.flag, tpl.h)
    info = srv.wait(srv.submit(gid, tpl.rhs, tpl.fnorm, want_x=True))
    srv.close()
    return info


def test_served_poisson_stream_200GB(gpu, serving_templates):
    tpls = serving_templates
    srv = S.Server(hbm_budget=200 * 10**9, tol=1e-6, vcycles_per_iter=4)
    gids = [srv.add_geometry(t.flag, t.h) for t in tpls]
    arr = W.arrivals(2000.0, 0.04, seed=7)  # the first 40 ms of the stream
    pick = W.rng(8).integers(0, len(tpls), arr.size)
    rids = []
    t0 = time.perf_counter()
    for ta, i in zip(arr, pick):
        while time.perf_counter() - t0 < ta:
            time.sleep(0.0002)
        rids.append((int(i), srv.submit(gids[i], tpls[i].rhs, tpls[i].fnorm, want_x=True)))
    infos = [(i, srv.wait(r)) for i, r in rids]
    st = srv.stats()
    srv.close()
    print(f"{len(infos)} requests, peak active {st['peak_active']}, peak "
          f"{st['peak_bytes'] / 1e9:.2f} GB", flush=True)
    assert st["completed"] == len(infos) and st["failed"] == 0
    assert st["peak_active"] > 1 and st["peak_bytes"] <= 200 * 10**9
    worst = 0.0
    for i, info in infos:
        t = tpls[i]
        assert info["state"] == "done" and info["converged"] and info["timesteps_done"] == 3
        _, rn = O.residual(info["x"], t.rhs[-1], t.flag, float(t.h))
        assert rn <= 1.0001e-6 * t.fnorm[-1], f"template {i}: {rn / t.fnorm[-1]:.3e}"
        worst = max(worst, rn / t.fnorm[-1])
    print(f"worst float64 relative residual over the stream: {worst:.3e}", flush=True)
    # alone vs in the mix: the largest and three other distinct templates
    seen = sorted({i for i, _ in infos}, key=lambda i: -tpls[i].cells)
    sample = seen[:1] + seen[1::max(1, len(seen) // 3)][:3]
    for i in sample:
        mixed = next(info for j, info in infos if j == i)
        alone = _alone(tpls[i])
        assert alone["iterations"] == mixed["iterations"], f"template {i}"
        assert np.array_equal(alone["x"], mixed["x"]), f"template {i}"
