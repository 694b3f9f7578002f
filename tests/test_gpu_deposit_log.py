"""GPU parity of the deposit log: global fluence deposits appended to a device log and summed into
the grid by separate passes (histogram, scan, scatter by bucket, LDS reduction) instead of float
atomics. Only the float summation order of a cell changes, so every bar is the atomics path's:
fluence within 1e-4 relative L2 of the reference's volume, energies and detected rows unchanged,
per-photon states bitwise. Covered: forced on for golden cases with labels in LDS and in HBM,
with and without the per-lane deposit cache and the LDS tally box; runs split into chunks by a
small log; a log too small for its chunk (deposits past it go to atomics); the pilot's choice of the log and of the
scatter-round vote; every deposit on a few hot cells (the reduce pass's contended adds)."""
import copy
import sys

import numpy as np
import pytest

import bench
import oracle_lib as O
import umcx_amd as U

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def check_against_reference(g, prob, sim):
    """The reference-golden bars of the atomics path (tests/test_gpu_parity.py): energies, absorbed
    fraction, energy balance, fluence within 1e-4 relative L2 (whole grid, or the golden's sample
    and sum of squares for large grids) and per-gate sums."""
    normalize = bool(prob["isnormalized"]) and "-U" not in g["argv"]
    res = sim.result(normalize=normalize)
    assert res.energy_total == pytest.approx(float(g["energy"]), abs=0.006)
    assert abs(res.absorbed_percent - float(g["absorbed"])) <= 1e-3
    assert abs(res.energy_deposit + res.energy_escape - res.energy_total) <= 1e-6 * res.energy_total
    f = res.fluence.reshape(-1)
    if "fluence" in g:
        assert rel_l2(f, g["fluence"]) <= 1e-4
    else:
        assert rel_l2(f[g["sample_idx"]], g["sample_val"]) <= 1e-4
        ss = float((f.astype(np.float64) ** 2).sum())
        assert abs(ss - float(g["sumsq"])) / float(g["sumsq"]) <= 2e-4
    np.testing.assert_allclose(res.fluence.astype(np.float64).sum(axis=(1, 2, 3)), g["gate_sums"], rtol=1e-4,
                               atol=1e-30)
    return res


@pytest.mark.parametrize("tally", [0, 1])
@pytest.mark.parametrize("name", ["c3", "c2", "skinvessel", "gated", "cube60", "normx", "zipvol"])
def test_forced_deposit_log_matches_reference(name, tally):
    """tally 0: no LDS box, every deposit goes to the log; tally 1: the pilot's box in LDS, the
    rest (through the per-lane cache where the pilot turns it on) to the log."""
    g, prob = O.golden_problem(name)
    say(f"deposit log forced on {name} tally {tally}")
    with U.Simulation(g["cfg"], deposit_log=1, tally_mode=tally) as sim:
        assert sim.deposit_log()
        sim.run()
        res = check_against_reference(g, prob, sim)
    if "detp" in g:                           # detected rows do not depend on the tally path
        got = res.detp[np.lexsort(res.detp.T[::-1])] if res.detp.size else res.detp
        assert np.array_equal(got.view(np.uint32), g["detp"].view(np.uint32))


def test_log_in_chunks_equals_atomics_and_keeps_photon_states():
    """A log of 16,384 segments holds about a quarter of 2e5 skin-model photons' deposits: the run
    goes in four chunks (trace entries offset per chunk). Fluence equals the atomics run up to
    summation order; every photon's final state is bitwise the same."""
    g = O.golden("c3")
    cfg = copy.deepcopy(g["cfg"])
    n = 200_000
    cfg["Session"]["Photons"] = n
    say("deposit log in chunks, c3 2e5 photons")
    with U.Simulation(cfg, deposit_log=-1, trace=True) as sim:
        assert not sim.deposit_log()
        sim.run()
        flu_a = sim.fluence()
        tr_a = sim.trace(n)
        e_a = sim.energies()
    with U.Simulation(cfg, deposit_log=1, deposit_log_segments=16384, trace=True) as sim:
        assert sim.deposit_log()
        sim.run()
        flu_l = sim.fluence()
        tr_l = sim.trace(n)
        e_l = sim.energies()
    assert rel_l2(flu_l, flu_a) <= 1e-5
    assert abs(float(flu_l.astype(np.float64).sum()) - float(flu_a.astype(np.float64).sum())) <= 1e-5 * float(flu_a.sum())
    for key in ("pos", "vec", "len", "w0", "mediaid", "lastvoxelidx"):
        assert np.array_equal(np.ascontiguousarray(tr_l[key]).view(np.uint8),
                              np.ascontiguousarray(tr_a[key]).view(np.uint8)), key
    assert e_l[1] == e_a[1] and abs(e_l[0] - e_a[0]) <= 1e-9 * e_a[1]


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_full_log_falls_back_to_atomics(name):
    """64 segments for thousands of waves: most waves find the log full at once and send their
    deposits to float atomics; nothing is lost."""
    g, prob = O.golden_problem(name)
    say(f"deposit log of 64 segments, {name}")
    with U.Simulation(g["cfg"], deposit_log=1, deposit_log_segments=64) as sim:
        sim.run()
        check_against_reference(g, prob, sim)


def test_pilot_chooses_the_log_where_atomics_would_bind():
    """Skin model (about 430 global deposits per photon after the tally box): the pilot turns the
    log on; heterogeneous reflective cube (about 110): float atomics. With the log, 1e6 skin-model
    photons give the atomics path's fluence and energies."""
    skin = copy.deepcopy(bench.SKIN_CONFIG)
    skin["Session"]["Photons"] = 1_000_000
    say("pilot choice: skin model")
    with U.Simulation(skin) as sim:
        sim.run()
        assert sim.deposit_log()
        assert sim.global_deposits_per_photon() > 300
        assert sim.vote() == 16          # 0.11 scatters per crossing: scatter rounds wait for 16 lanes
        flu_l = sim.fluence()
        e_l = sim.energies()
    with U.Simulation(skin, deposit_log=-1) as sim:
        sim.run()
        assert not sim.deposit_log()
        flu_a = sim.fluence()
        e_a = sim.energies()
    assert rel_l2(flu_l, flu_a) <= 1e-5
    assert e_l[1] == e_a[1] and abs(e_l[0] - e_a[0]) <= 1e-9 * e_a[1]
    hetero = copy.deepcopy(bench.CONFIG)
    hetero["Session"]["Photons"] = 1_000_000
    say("pilot choice: heterogeneous reflective cube")
    with U.Simulation(hetero) as sim:
        sim.run()
        assert not sim.deposit_log()
        assert 50 < sim.global_deposits_per_photon() < 200
        assert sim.vote() == 0           # 0.83 scatters per crossing: scatter right away


def test_log_on_a_few_hot_cells():
    """Every deposit lands in one of 4^3 cells (a 4 mm cube of strong absorption and scattering, no
    LDS box): the reduce pass's compare-and-swap adds meet the same cells from many lanes at once,
    and its equal record shares split the grid's one bucket between many workgroups, each adding
    its share's sums to the same cells. The grid equals the atomics run's up to summation order
    (float32 sums of ~1e5 deposits per cell: the fluence bar, 1e-4 relative L2) and holds the
    deposited energy (energy output: each deposit is its weight drop)."""
    cfg = {"Session": {"ID": "hot", "Photons": 200_000, "RNGSeed": 12345, "DoMismatch": 0,
                       "DoSaveVolume": 1, "DoPartialPath": 0, "OutputType": "e"},
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
           "Domain": {"Dim": [4, 4, 4], "LengthUnit": 1,
                      "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.5, "mus": 20, "g": 0.5, "n": 1}]},
           "Optode": {"Source": {"Type": "pencil", "Pos": [2, 2, 0], "Dir": [0, 0, 1]}},
           "Shapes": [{"Grid": {"Tag": 1, "Size": [4, 4, 4]}}]}
    say("deposit log on 64 hot cells")
    out = {}
    for mode in (1, -1):
        with U.Simulation(cfg, deposit_log=mode, tally_mode=0) as sim:
            assert sim.deposit_log() == (mode == 1)
            sim.run()
            out[mode] = (sim.fluence().astype(np.float64), sim.energies())
    (flu_l, e_l), (flu_a, e_a) = out[1], out[-1]
    assert e_l[1] == e_a[1] and abs(e_l[2] - e_a[2]) <= 1e-9 * e_a[1]
    assert e_l[2] > 0.3 * e_l[1]                       # about half the weight is deposited (oracle: 47 %)
    assert rel_l2(flu_l, flu_a) <= 1e-4
    assert float(flu_l.sum()) == pytest.approx(e_l[2], rel=1e-4)
    assert float(flu_a.sum()) == pytest.approx(e_a[2], rel=1e-4)
