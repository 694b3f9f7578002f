"""Fluence grids past 2^32 cells (HBM-sized outputs).

A 1024^3-voxel cube with 5 time gates has 5.4e9 fluence cells (21.5 GB of float32): every deposit
in gate 4 lands at cell 2^32 or beyond; the pilot's count grid is as large, and the LDS tile is
chosen from it on the device. The photons start at (64.5, 64.5, 64.5) and stay within 28 voxels
of it for the simulated 0.5 ns, so the same photons in a 128^3
cube (small enough for the oracle) follow the same trajectories: the final photon states must be
bitwise equal, the big grid's [0,128)^3 window must hold the oracle's fluence and every other
cell must stay zero.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cube_config(n, photons):
    return {"Session": {"ID": "bigcube", "Photons": photons, "RNGSeed": 7171, "DoMismatch": 0, "OutputType": "x"},
            "Forward": {"T0": 0, "T1": 5e-10, "Dt": 1e-10},
            "Domain": {"Dim": [n, n, n], "LengthUnit": 1,
                       "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.01, "mus": 20, "g": 0.9, "n": 1.37}]},
            "Optode": {"Source": {"Type": "pencil", "Pos": [64.5, 64.5, 64.5], "Dir": [0, 0, 1]}, "Detector": []},
            "Shapes": [{"Grid": {"Tag": 1, "Size": [n, n, n]}}]}


def test_fluence_grid_past_2_to_32_cells():
    small, big, photons = 128, 1024, 20000
    cfg = cube_config(small, photons)
    prob = O.problem_from_config(cfg, U.domain(cfg))
    say("oracle: 128^3 cube, 5 gates, 20,000 photons")
    o = O.run(prob, traces=True, nthreads=8)
    say("gpu: 1024^3 cube, 5 gates (5.4e9 fluence cells), forced LDS tile")
    with U.Simulation(cube_config(big, photons), trace=True, tally_mode=1) as sim:
        nx, ny, nz, nt = sim.dims
        assert (nt, nx, ny, nz) == (5, big, big, big) and nt * nx * ny * nz > 2 ** 32
        sim.run(0, photons)
        tr = sim.trace(photons)
        box = sim.tally_box()
        esc, tot, dep = sim.energies()
        say("gpu: reading back 21.5 GB of fluence")
        flu = sim.fluence()
    ot = o["traces"]
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len"):
        assert np.array_equal(exact(tr[key]), exact(ot[key])), key
    assert np.array_equal(tr["mediaid"], ot["mediaid"])
    # the tile sits where the deposits are: inside the oracle's window, which it matches
    assert box is not None and box[4] > 0
    assert box[0] + box[4] <= small and box[1] + box[5] <= small and box[2] + box[6] <= small
    oesc, otot, _ = o["energies"]
    assert tot == otot and abs(esc - oesc) <= 1e-10 * otot
    assert abs(dep + esc - tot) <= 1e-6 * tot
    say("checking the window and the cells outside it")
    win = flu[:, :small, :small, :small]
    ref = o["fluence"].reshape(nt, small, small, small)
    assert np.count_nonzero(win[4]) == np.count_nonzero(ref[4]) > 0       # gate 4: cells 2^32 and up
    rel = np.linalg.norm((win - ref).ravel()) / np.linalg.norm(ref.ravel())
    assert rel <= 1e-4, rel
    assert np.count_nonzero(flu) == np.count_nonzero(win)
