"""Fluid components and per-component mean removal (ubg_flag_components,
ubg_remove_component_means): the classification a pure-Neumann rhs needs
to be made compatible (each component's rhs summing to zero).

The oracle (oracle_components, oracle_remove_component_means) is pinned
on the CPU against scipy.ndimage.label under 4-connectivity (the same
numbering: raster order of each component's first cell) and against what
the arithmetic fixes (component sums of the result, a float64 mean
subtraction, components that are single cells, constant fields); the GPU
results are then held to the oracle bit for bit.
"""
import numpy as np
import pytest
from scipy import ndimage

import ubootgl_amd as U
from ubootgl_amd import workloads as W
from oracle import pyoracle as O


def spiral(n):
    """A one-cell-wide fluid corridor wound as a square spiral (walls one
    cell thick between turns): one component whose cells are linked in a
    chain as long as the domain allows."""
    fl = np.zeros((n, n), np.float32)
    k = 0
    while True:
        lo, hi = 2 * k, n - 1 - 2 * k
        if hi - lo < 2:
            if hi >= lo:
                fl[lo, lo:hi + 1] = 1
            break
        fl[lo, lo:hi + 1] = 1          # top row, left to right
        fl[lo:hi + 1, hi] = 1          # right column, down
        fl[hi, lo:hi + 1] = 1          # bottom row, right to left
        fl[lo + 2:hi + 1, lo] = 1      # left column, up to two below the top
        fl[lo + 2, lo + 1] = 1         # step in to the next turn's top row
        k += 1
    return fl


def cases():
    g = np.random.default_rng(3)
    out = [("all fluid", np.ones((17, 23), np.float32)),
           ("all solid", np.zeros((9, 11), np.float32)),
           ("checkerboard", (np.indices((20, 31)).sum(0) % 2).astype(np.float32)),
           ("row", (g.random((1, 70)) > 0.3).astype(np.float32)),
           ("column", (g.random((70, 1)) > 0.3).astype(np.float32)),
           ("spiral", spiral(61))]
    for k in range(6):
        h, w = g.integers(2, 90, 2)
        out.append((f"random {k}", (g.random((h, w)) > g.uniform(0.3, 0.6)).astype(np.float32)))
    d = W.small_domain(160, 96, seed=5, ndisks=8)
    out.append(("walls + disks", d.flag.astype(np.float32)))
    return out


CASES = cases()


# ------------------------------------------------------------------- CPU


@pytest.mark.parametrize("name,flag", CASES, ids=[c[0] for c in CASES])
def test_oracle_components_equal_scipy_labels(name, flag):
    lab, n = O.components(flag)
    ref, nr = ndimage.label(flag > 0.5)  # 4-connectivity, raster numbering
    assert n == nr
    np.testing.assert_array_equal(lab, ref.astype(np.int32) - 1)


def test_oracle_components_closed_forms():
    # a checkerboard: every fluid cell is its own component
    cb = (np.indices((12, 13)).sum(0) % 2).astype(np.float32)
    lab, n = O.components(cb)
    assert n == int(cb.sum())
    assert sorted(lab[cb > 0].tolist()) == list(range(n))
    # a spiral is one component
    assert O.components(spiral(61))[1] == 1
    # two disjoint bars, numbered in raster order of their first cell
    fl = np.zeros((5, 9), np.float32)
    fl[3, 0:4] = 1
    fl[1, 5:9] = 1
    lab, n = O.components(fl)
    assert n == 2 and lab[1, 5] == 0 and lab[3, 0] == 1


@pytest.mark.parametrize("name,flag", CASES, ids=[c[0] for c in CASES])
def test_oracle_mean_removal(name, flag):
    """Every component of the result sums to ~0 (within the float32
    rounding of its cells), the result is the float64 mean subtraction
    rounded to float32 (within one ulp), and solid cells are unchanged."""
    g = np.random.default_rng(len(name))
    f = (g.standard_normal(flag.shape) * 3.0 + 1.5).astype(np.float32)
    out, n = O.remove_component_means(flag, f)
    lab, n2 = O.components(flag)
    assert n == n2
    np.testing.assert_array_equal(out[lab < 0], f[lab < 0])
    ref = f.astype(np.float64).copy()
    for c in range(n):
        m = lab == c
        ref[m] -= f[m].astype(np.float64).mean()
        s = out[m].astype(np.float64).sum()
        assert abs(s) <= m.sum() * np.spacing(np.float32(np.abs(f[m]).max() + 4.0))
    fl = lab >= 0
    np.testing.assert_allclose(out[fl], ref[fl].astype(np.float32),
                               rtol=0, atol=float(np.spacing(np.float32(16.0))))
    # a constant field on a component becomes exactly 0 there
    c = np.full(flag.shape, 2.5, np.float32)
    out_c, _ = O.remove_component_means(flag, c)
    assert np.all(out_c[fl] == 0.0)


def test_component_entry_points_reject_bad_input():
    """Argument errors are found on the host, before any device work (so
    they hold without a GPU too)."""
    from ubootgl_amd import _lib as L
    bad = np.ones((8, 8), np.float32)
    bad[3, 3] = 0.5
    with pytest.raises(L.UbgError) as e:
        U.flag_components(bad)
    assert e.value.code == L.E_FLAG_VALUE
    f = np.zeros((8, 8), np.float32)
    f[2, 2] = np.nan
    with pytest.raises(L.UbgError) as e:
        U.remove_component_means(np.ones((8, 8), np.float32), f)
    assert e.value.code == L.E_NONFINITE


# ------------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("name,flag", CASES, ids=[c[0] for c in CASES])
def test_gpu_components_equal_oracle(gpu, name, flag):
    lab, n = U.flag_components(flag)
    ref, nr = O.components(flag)
    assert n == nr
    np.testing.assert_array_equal(lab, ref)
    g = np.random.default_rng(7)
    f = (g.standard_normal(flag.shape) * 3.0 + 1.5).astype(np.float32)
    out, n3 = U.remove_component_means(flag, f)
    oref, _ = O.remove_component_means(flag, f)
    assert n3 == nr
    np.testing.assert_array_equal(out, oref)


@pytest.mark.gpu
def test_gpu_components_large(gpu):
    """The 4096x2048 benchmark flag (walls, 400 disks) and a 2048x2048 random
    flag near the percolation threshold (tens of thousands of components,
    long irregular chains): labels and mean removal bit-identical to the
    oracle, and the same on a second call."""
    d = W.config1()
    g = np.random.default_rng(11)
    rnd = (g.random((2048, 2048)) > 0.41).astype(np.float32)
    for flag in (d.flag.astype(np.float32), rnd):
        lab, n = U.flag_components(flag)
        ref, nr = O.components(flag)
        assert n == nr
        np.testing.assert_array_equal(lab, ref)
        lab2, _ = U.flag_components(flag)
        np.testing.assert_array_equal(lab2, lab)
        f = g.standard_normal(flag.shape).astype(np.float32)
        out, _ = U.remove_component_means(flag, f)
        oref, _ = O.remove_component_means(flag, f)
        np.testing.assert_array_equal(out, oref)
        print(f"{flag.shape}: {n} components", flush=True)
