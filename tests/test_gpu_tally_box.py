"""The sessions pick their LDS tally box on the GPU (densest cell, total and greedy growth in one
workgroup, never copying the pilot's count grid to the host). Pilot counts are whole numbers, so
the device's double sums are exact and its box must equal the host search of tally_box.cpp
(`umcx_choose_tally_box`) cell for cell, ties included; its in-box and total counts must equal
numpy's sums.
"""
import ctypes as C

import numpy as np
import pytest

from umcx_amd._native import check, lib

pytestmark = pytest.mark.gpu


def host_box(c, budget):
    nt, nx, ny, nz = c.shape
    box = (C.c_int32 * 8)()
    check(lib().umcx_choose_tally_box(c.ctypes.data, nx, ny, nz, nt, budget, box))
    return list(box)


def device_box(c, budget):
    nt, nx, ny, nz = c.shape
    box = (C.c_int32 * 8)()
    sums = (C.c_double * 2)()
    check(lib().umcx_choose_tally_box_device(c.ctypes.data, nx, ny, nz, nt, budget, box, sums, 0))
    return list(box), sums[0], sums[1]


def in_box(c, b):
    x0, y0, z0, g0, sx, sy, sz, sg = b
    return float(c[g0:g0 + sg, x0:x0 + sx, y0:y0 + sy, z0:z0 + sz].sum(dtype=np.float64))


def pilot_like(rng, shape, spots):
    """Whole-number counts: Poisson draws around a few hot spots over a sparse background."""
    g = np.indices(shape).astype(np.float64)
    lam = rng.uniform(0.0, 0.3, shape)
    for _ in range(spots):
        centre = rng.uniform([0] * 4, shape)
        width = rng.uniform(0.5, 8.0)
        d2 = sum((g[i] - centre[i]) ** 2 for i in range(4))
        lam += rng.uniform(50, 5000) * np.exp(-d2 / width)
    return rng.poisson(lam).astype(np.float32)


def check_same(c, budget):
    c = np.ascontiguousarray(c, dtype=np.float32)
    want = host_box(c, budget)
    got, inside, total = device_box(c, budget)
    assert got == want, (c.shape, budget, got, want)
    assert inside == in_box(c, want) if want[4] else inside == 0.0
    assert total == float(c.sum(dtype=np.float64))


def test_device_box_search_equals_host_search():
    rng = np.random.default_rng(2024)
    for trial in range(60):
        shape = tuple(int(v) for v in rng.integers(1, [9, 41, 41, 41]))
        check_same(pilot_like(rng, shape, int(rng.integers(1, 4))), int(rng.integers(8, 6000)))
    # headline-sized grid: 50 gates of 60^3, a pencil-beam-like plume
    g = np.indices((50, 60, 60, 60), dtype=np.float32)
    lam = 3e4 * np.exp(-((g[1] - 29.5) ** 2 + (g[2] - 29.5) ** 2) / 40.0 - g[3] / 6.0 - g[0] / 4.0)
    check_same(rng.poisson(lam).astype(np.float32), 25992)


def test_device_box_search_ties_and_edges():
    check_same(np.zeros((3, 5, 6, 7)), 64)                        # empty grid: no box
    check_same(np.ones((2, 7, 7, 7)), 50)                         # one plateau: first cell, first slabs
    flat = np.zeros((1, 9, 9, 9), dtype=np.float32)
    flat[0, 2, 2, 2] = flat[0, 6, 6, 6] = 7.0                      # two equal maxima: the first wins
    check_same(flat, 27)
    check_same(np.full((1, 1, 1, 1), 3.0), 1)                     # one cell
    check_same(np.arange(4 * 5 * 6 * 7, dtype=np.float32).reshape(4, 5, 6, 7) % 5, 40)
    check_same(np.ones((1, 4, 4, 4)), 1)                          # budget of one cell
