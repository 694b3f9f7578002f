"""Pins the per-photon work model bench.py prices its roofline with: the mean numbers of steps,
voxel crossings, scatters, random draws and Fresnel tests per photon of the heterogeneous
reflective workload, measured with the oracle."""
import numpy as np
import pytest

import bench
import oracle_lib as O
import umcx_amd as U


def test_event_counts_per_photon():
    prob = O.problem_from_config(bench.CONFIG, U.domain(bench.CONFIG))
    t = O.run(prob, nphoton=40000, traces=True, nthreads=8)["traces"]
    got = {"S": t["nstep"].mean(), "X": t["ncross"].mean(), "K": t["nscat"].mean(), "R": t["nrand"].mean(),
           "F": t["nfresnel"].mean()}
    for k, v in bench.COUNTS.items():
        assert got[k] == pytest.approx(v, rel=0.03), (k, got[k], v)
    # R = 1 + 3K + Fresnel draws: one draw for the first free path, three per scatter
    assert np.all(t["nrand"] >= 1 + 3 * t["nscat"])
