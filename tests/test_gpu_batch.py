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


def srt(r):
    return r[np.lexsort(r.T[::-1])] if r.size else r


@pytest.mark.parametrize("block,tally", [(0, -1), (1024, 0)])
def test_batch_detected_photons_equal_separate_runs_and_reference(block, tally):
    """Jobs with detectors in one batch: each job's detected rows (with photon index and final
    weight) equal, bitwise as a multiset, those of a separate run of the same job; at the golden
    photon counts the rows equal the reference binary's. The jobs mix partial paths in registers
    and in LDS (3 and 5 media), exit position and direction columns, a DoSaveVolume-0 job, a job
    whose MaxDetPhoton overflows and a job without detectors."""
    names = ["c2", "sphshells", "cube60b", "overflow"]
    jobs = [O.golden(n)["cfg"] for n in names]
    nodet = json.loads(json.dumps(jobs[2]))
    nodet["Session"]["DoPartialPath"] = 0
    novol = json.loads(json.dumps(jobs[2]))
    novol["Session"]["DoSaveVolume"] = 0
    novol["Session"]["RNGSeed"] = 99
    jobs = jobs + [nodet, novol]
    say(f"batch with detectors, block {block} tally {tally}")
    with B.BatchSimulation(jobs, block=block, tally_mode=tally) as bs:
        bs.run()
        got = [(bs.detected(j), bs.detp(j), bs.energies(j)) for j in range(len(jobs))]
        bs.run()                                   # a graph replay restarts the counters
        again = [bs.detected(j) for j in range(len(jobs))]
    assert again == [g[0] for g in got]
    for j, cfg in enumerate(jobs):
        with U.Simulation(cfg, tally_mode=0) as s:
            s.run()
            det, rows_ids_w, e = s.detected(), s.detp(), s.energies()
        (bdet, (brows, bids, bw), be) = got[j]
        assert bdet == det, (j, bdet, det)
        assert be[1] == e[1], j
        key = lambda r, i, w: srt(np.column_stack([r, i.view(np.float32), w]))
        if names[j % len(names)] == "overflow" and j < len(names):
            # which detections get the 10 slots depends on the order of the atomics (in the
            # reference too): every kept row must be a genuine one
            full = json.loads(json.dumps(cfg))
            full["Session"]["MaxDetPhoton"] = 10 ** 6
            with U.Simulation(full, tally_mode=0) as s:
                s.run()
                allrows = {r.tobytes() for r in key(*s.detp())}
            assert all(r.tobytes() in allrows for r in key(brows, bids, bw))
            continue
        assert np.array_equal(key(brows, bids, bw).view(np.uint32), key(*rows_ids_w).view(np.uint32)), j
        if j < len(names):
            g = O.golden(names[j])
            assert bdet[1] == int(g["detected"]), j
            assert np.array_equal(srt(brows).view(np.uint32), g["detp"].view(np.uint32)), j
    assert got[3][0][0] > got[3][0][1]            # the overflow job dropped rows, as the reference does
    assert got[4][0] == (0, 0)
