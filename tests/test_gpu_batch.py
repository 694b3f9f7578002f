"""GPU checks of the batched form (umcx_batch_*): every job of one persistent launch equals a
separate run of the same job (and the oracle), and graph replays are consistent."""
import json
import sys

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U
from umcx_amd import batch as B

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_batch_equals_separate_runs_and_graph_replays():
    jobs = B.uneven_jobs(seed=7, njobs=12, max_photons=30000)
    say("batch vs separate runs")
    with B.BatchSimulation(jobs) as bs:
        bs.run()
        first = [(bs.fluence(j), bs.energies(j)) for j in range(len(jobs))]
        # replay the graph twice more with eager work in between; every replay must agree
        for _ in range(2):
            with U.Simulation(jobs[0]) as s:
                s.run(0, 1000)
            bs.run()
            for j in range(len(jobs)):
                f, e = bs.fluence(j), bs.energies(j)
                assert e[1] == first[j][1][1]
                assert e[0] == pytest.approx(first[j][1][0], rel=1e-10)
                assert np.linalg.norm(f - first[j][0]) <= 1e-4 * np.linalg.norm(first[j][0])
    for j, cfg in enumerate(jobs):
        with U.Simulation(cfg, tally_mode=0) as s:
            s.run()
            f, e = s.fluence(), s.energies()
        bf, be = first[j]
        assert be[1] == e[1], j                                   # launched weight: exact
        assert be[0] == pytest.approx(e[0], rel=1e-10), j         # escaped: summation order
        assert abs(be[2] + be[0] - be[1]) <= 1e-6 * be[1]          # energy balance
        assert np.linalg.norm(bf - f) <= 1e-4 * np.linalg.norm(f), j   # summation order


def test_batch_job_matches_oracle():
    jobs = B.uneven_jobs(seed=7, njobs=3, max_photons=20000)
    say("batch vs oracle")
    with B.BatchSimulation(jobs) as bs:
        bs.run()
        res = [(bs.fluence(j), bs.energies(j)) for j in range(len(jobs))]
    for j, cfg in enumerate(jobs):
        prob = O.problem_from_config(cfg, U.domain(cfg))
        o = O.run(prob, nthreads=8)
        f, e = res[j]
        assert e[1] == o["energies"][1]
        assert e[0] == pytest.approx(o["energies"][0], rel=1e-10)
        of = o["fluence"].reshape(f.shape)
        assert np.linalg.norm(f - of) <= 1e-4 * np.linalg.norm(of)


def mixed_jobs():
    """Jobs that exercise every per-job switch of the batched kernel: several gates, reflection
    off, energy output, no volume output, a 4-bit-label volume and a tiny photon count."""
    base = B.uneven_jobs(seed=11, njobs=6, max_photons=20000)
    base[0]["Forward"] = {"T0": 0.0, "T1": 3e-9, "Dt": 1e-9}          # 3 gates
    base[1]["Session"]["DoMismatch"] = 0                               # no Fresnel
    base[2]["Session"]["OutputType"] = "e"                             # energy deposit
    base[3]["Session"]["DoSaveVolume"] = 0                             # energies only
    base[4]["Session"]["Photons"] = 37                                 # fewer photons than a wave
    return base


@pytest.mark.parametrize("block,blocks,tally", [(0, 0, -1), (256, 0, 0), (1024, 0, 1), (512, 7, 1)])
def test_batch_switches_and_shapes_equal_separate_runs(block, blocks, tally):
    jobs = mixed_jobs()
    say(f"mixed batch block {block} blocks {blocks} tally {tally}")
    with B.BatchSimulation(jobs, block=block, blocks=blocks, tally_mode=tally) as bs:
        bs.run()
        got = [(bs.fluence(j), bs.energies(j)) for j in range(len(jobs))]
        if tally == 0:
            assert all(t["tile_cells"] == 0 for t in bs.trials())
    for j, cfg in enumerate(jobs):
        with U.Simulation(cfg, tally_mode=0) as s:
            s.run()
            f, e = s.fluence(), s.energies()
        bf, be = got[j]
        assert be[1] == e[1], j
        assert be[0] == pytest.approx(e[0], rel=1e-10, abs=1e-12), j
        if f is None:
            assert bf.size == 0, j
        else:
            assert bf.shape == f.shape, j
            assert np.linalg.norm(bf - f) <= 1e-4 * max(np.linalg.norm(f), 1e-30), j
