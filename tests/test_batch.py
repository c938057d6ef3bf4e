"""Batched refinement solves (ubg_batch_*).

Parity contract: every domain of a batch ends bit-identical to the
single-domain refinement solve of the same problem (MG.ir_solve, itself
checked against the FP64 CPU multigrid in test_refine.py), with the same
iteration count and residual history -- whatever the batch mix, the tile
shape or the workgroup count.
"""
import os

import numpy as np
import pytest

import ubootgl_amd as U
from ubootgl_amd import _lib as L
from ubootgl_amd import batch as B
from ubootgl_amd import workloads as W
from oracle import pyoracle as O

MIXED_SIZES = [(64, 64), (96, 32), (128, 128), (160, 64), (256, 96), (320, 320),
               (512, 128), (384, 224), (16, 16), (33, 47), (1024, 256), (200, 130)]


def problems(sizes, seed=21):
    g = np.random.default_rng(seed)
    return [W.poisson_problem(w, h, seed=seed * 100 + i, solid=float(g.uniform(0, 0.08)),
                              nmodes=8) for i, (w, h) in enumerate(sizes)]


def single(P, tol, k=4):
    mg = U.MG(P.flag)
    mg.ir_set_rhs(P.f)
    mg.ir_set_guess(None)
    res = mg.ir_solve(float(P.h), tol=tol, vcycles_per_iter=k)
    x = mg.ir_solution()
    mg.close()
    return res, x


# ------------------------------------------------------------------- CPU


def test_batch_sizes_rule():
    s = W.batch_sizes(500, seed=2)
    for w, h in s:
        assert w % 32 == 0 and h % 32 == 0
        assert 128 <= w <= 2048 and h >= 32
        assert h <= w + 16  # aspect >= 1 up to rounding
    ws = np.array([w for w, _ in s], float)
    # log-uniform: the median is near sqrt(128 * 2048) = 512
    assert 380 < np.median(ws) < 680


def test_batch_bytes_host_rule():
    """ubg_batch_bytes is host arithmetic (no device needed): it grows with
    the domains and rejects domains below 16 x 16."""
    a = B.batch_bytes([(128, 64)])
    b = B.batch_bytes([(128, 64), (256, 256)])
    assert b > a > 128 * 64 * 32
    with pytest.raises(L.UbgError) as e:
        B.batch_bytes([(8, 64)])
    assert e.value.code == L.E_DOMAIN_TOO_SMALL


# ------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_batch_equals_single_domain(gpu):
    probs = problems(MIXED_SIZES)
    bt = B.Batch([(P.W, P.H) for P in probs])
    for d, P in enumerate(probs):
        bt.set_problem(d, P)
        bt.set_guess(d)
    reps = bt.solve(tol=1e-6)
    assert reps >= 1
    for d, P in enumerate(probs):
        rb = bt.result(d)
        xb = bt.solution(d)
        rs, xs = single(P, 1e-6)
        assert rb["converged"] and rb["rel_residual"] <= 1e-6
        assert rb == rs, f"domain {d} {P.W}x{P.H}: {rb} vs {rs}"
        assert np.array_equal(xb, xs), f"domain {d} {P.W}x{P.H}"
        # independent of the GPU: the oracle's float64 residual of the
        # returned solution against ||r_0|| = ||F f|| of the zero start
        _, r0 = O.residual(np.zeros((P.H, P.W)), P.f, P.flag, float(P.h))
        _, rn = O.residual(xb, P.f, P.flag, float(P.h))
        assert rn <= 1.0001e-6 * r0, f"domain {d} {P.W}x{P.H}: {rn / r0:.3e}"
    bt.close()


def stream_units(w, h):
    """Units (124-column strips x 16-row blocks) of a level's stand-alone
    sweep; a batch streams a level when those of its eligible domains (>= 512
    wide, >= 2^18 cells) sum to at least 4096."""
    return -(-w // 124) * -(-(h - 2) // 16)


@pytest.mark.gpu
def test_batch_streamed_levels_equal_single(gpu):
    """Level 0 runs as row-streamed passes here (an 8192 x 1024 domain lifts
    the level's units over the gate): an odd width (unaligned loads) and
    mid-size domains stream beside it, a small domain tiles -- each domain
    bit-identical to the single-domain solve."""
    sizes = [(8192, 1024), (1023, 300), (2048, 512), (1024, 256), (100, 60)]
    assert sum(stream_units(w, h) for w, h in sizes[:4]) >= 4096
    probs = problems(sizes, seed=33)
    bt = B.Batch([(P.W, P.H) for P in probs])
    for d, P in enumerate(probs):
        bt.set_problem(d, P)
    bt.solve(tol=1e-6)
    for d, P in enumerate(probs):
        rs, xs = single(P, 1e-6)
        assert bt.result(d) == rs, f"domain {d} {P.W}x{P.H}"
        assert np.array_equal(bt.solution(d), xs), f"domain {d} {P.W}x{P.H}"
    bt.close()


@pytest.mark.gpu
def test_batch_streamed_short_levels_equal_single(gpu):
    """Sixteen 8192 x 64 domains: their 64-row level 0 (four row blocks per
    strip; shorter wide domains have no coarse level that fits in LDS) runs
    streamed over the batch, while the single-domain solve tiles it; the
    results are bit-identical."""
    assert 16 * stream_units(8192, 64) >= 4096
    probs = problems([(8192, 64)] * 16, seed=41)
    bt = B.Batch([(P.W, P.H) for P in probs])
    for d, P in enumerate(probs):
        bt.set_problem(d, P)
    bt.solve(tol=1e-6)
    for d in (0, 7, 15):
        rs, xs = single(probs[d], 1e-6)
        assert bt.result(d) == rs, f"domain {d}"
        assert np.array_equal(bt.solution(d), xs), f"domain {d}"
    bt.close()


@pytest.mark.gpu
def test_batch_alone_equals_in_mix_and_tuning_invariant(gpu):
    probs = problems(MIXED_SIZES[:8], seed=5)
    alone = B.Batch([(probs[3].W, probs[3].H)])
    alone.set_problem(0, probs[3])
    alone.solve(tol=1e-7)
    x_alone = alone.solution(0)
    r_alone = alone.result(0)
    alone.close()
    for shape, blocks, rows in (((32, 8), 0, 8), ((128, 4), 64, 32), ((64, 16), 7, 16),
                                ((64, 8), 0, 0)):
        bt = B.Batch([(P.W, P.H) for P in probs])
        bt.set_tuning(*shape, blocks, rows)
        for d, P in enumerate(probs):
            bt.set_problem(d, P)
        bt.solve(tol=1e-7)
        assert bt.result(3) == r_alone
        assert np.array_equal(bt.solution(3), x_alone), (shape, blocks)
        bt.close()


@pytest.mark.gpu
def test_batch_resolve_and_warm_start(gpu):
    """A second solve of the same batch (new graph state) from zero gives the
    same answer; a warm start reuses the previous x."""
    probs = problems(MIXED_SIZES[:5], seed=9)
    bt = B.Batch([(P.W, P.H) for P in probs])
    for d, P in enumerate(probs):
        bt.set_problem(d, P)
    bt.solve(tol=1e-6)
    x1 = [bt.solution(d) for d in range(len(probs))]
    r1 = [bt.result(d) for d in range(len(probs))]
    for d in range(len(probs)):
        bt.set_guess(d)
    bt.solve(tol=1e-6)
    for d in range(len(probs)):
        assert bt.result(d) == r1[d]
        assert np.array_equal(bt.solution(d), x1[d])
    bt.solve(tol=0.5)  # warm: r0 is already ~1e-6 of the cold one
    for d in range(len(probs)):
        assert bt.result(d)["r0_norm"] <= 2e-6 * r1[d]["r0_norm"]
    bt.close()


@pytest.mark.gpu
def test_batch_autotune_and_db(gpu, tmp_path):
    probs = problems(MIXED_SIZES[:6], seed=4)
    bt = B.Batch([(P.W, P.H) for P in probs])
    for d, P in enumerate(probs):
        bt.set_problem(d, P)
    db = str(tmp_path / "tune.db")
    t1 = bt.autotune(db)
    assert (t1["tile_w"], t1["tile_h"]) in B.TILE_SHAPES and t1["sweep_ms"] > 0
    assert t1["pass_rows"] in (0, 8, 16, 32)
    assert os.path.getsize(db) > 0
    t2 = bt.autotune(db)  # reused from the file
    keys = ("tile_w", "tile_h", "blocks", "pass_rows")
    assert tuple(t2[k] for k in keys) == tuple(t1[k] for k in keys)
    assert bt.tuning() == {k: t1[k] for k in keys}
    bt.solve(tol=1e-6)
    rs, xs = single(probs[0], 1e-6)
    assert np.array_equal(bt.solution(0), xs)
    bt.close()


@pytest.mark.gpu
def test_batch_error_paths(gpu):
    P = W.poisson_problem(64, 48, seed=3, solid=0.0)
    bt = B.Batch([(64, 48), (32, 32)])
    with pytest.raises(L.UbgError) as e:
        bt.solve()
    assert e.value.code == L.E_STATE
    walls = W.walls(64, 48).astype(np.float32)
    with pytest.raises(L.UbgError) as e:
        bt.set_domain(0, walls, P.f, P.h)
    assert e.value.code == L.E_FLAG_VALUE
    with pytest.raises(L.UbgError) as e:
        bt.set_tuning(48, 8)
    assert e.value.code == L.E_INVALID
    with pytest.raises(L.UbgError) as e:
        bt.set_tuning(64, 8, 0, 12)
    assert e.value.code == L.E_INVALID
    bt.set_problem(0, P)
    P2 = W.poisson_problem(32, 32, seed=4, solid=0.0)
    bt.set_problem(1, P2)
    with pytest.raises(L.UbgError) as e:
        bt.solve(tol=1e-12, vcycles_per_iter=1, max_iter=1)
    assert e.value.code == L.E_NOT_CONVERGED
    bt.solve(tol=1e-6)
    assert bt.result(0)["converged"] and bt.result(1)["converged"]
    bt.close()
    with pytest.raises(L.UbgError) as e:
        B.Batch([(8, 64)])
    assert e.value.code == L.E_DOMAIN_TOO_SMALL
