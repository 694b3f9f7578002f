"""Pinning of the oracle's per-voxel replay (oracle_replay), the CPU reference of the Jacobian.

The replay credits each voxel segment of a photon's path (length L_v, scatterings n_v) to the
voxel the photon leaves. What fixes it, independently of its own code:

* per medium, the segment lengths summed in crossing order equal, bit for bit, the partial path
  lengths the reference binary saved for the same detected photons (golden c4src rows,
  src/umcx.cpp:286-288);
* a non-scattering (mus = 0) tilted pencil beam crosses the voxels along a straight line, so
  d log w / d mua_v = -L_v must equal the chord lengths of an independent ray/voxel sweep in
  float64, voxel by voxel (an off-by-one at a crossing -- the planted mutant 4 -- fails it);
* the scatterings credited over all voxels equal each photon's scatter count (vec.w) from the
  oracle's forward run, and the replayed final weights equal the forward ones bitwise.
"""
import copy

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U


def srt(r):
    return r[np.lexsort(r.T[::-1])] if r.size else r


def test_replay_segments_sum_to_reference_partial_paths():
    g, prob = O.golden_problem("c4src")
    o = O.run(prob, nthreads=8)
    ref = g["detp"]
    assert np.array_equal(srt(o["detp"]).view(np.uint32), srt(ref).view(np.uint32))
    ids = o["detp_id"]
    r = O.replay(prob, ids, np.zeros(len(ids), np.int32), np.ones(len(ids)), 1, nthreads=8)
    rows = np.column_stack([o["detp"][:, 0], r["mpath"]])
    assert np.array_equal(srt(rows).view(np.uint32), srt(ref).view(np.uint32))
    assert np.array_equal(r["final_w"].view(np.uint32), o["detp_w"].view(np.uint32))
    # every segment length is also a -d log w / d mua term (mm, unit 1 here)
    assert -r["gmua"].sum() == pytest.approx(float(ref[:, 1:].astype(np.float64).sum()), rel=1e-6)


def ballistic_cfg(unit=0.5):
    d = np.array([0.30, 0.21, 0.93])
    d = (d / np.linalg.norm(d)).astype(np.float32)
    cfg = {"Session": {"Photons": 3, "RNGSeed": 7, "DoMismatch": 0, "DoPartialPath": 0, "DoSaveVolume": 1,
                       "OutputType": "x"},
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
           "Domain": {"Dim": [12, 10, 16], "LengthUnit": unit,
                      "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.02, "mus": 0, "g": 0.9, "n": 1}]},
           "Optode": {"Source": {"Type": "pencil", "Pos": [2.3, 3.7, 0], "Dir": [float(v) for v in d]}},
           "Shapes": [{"Grid": {"Tag": 1, "Size": [12, 10, 16]}}]}
    return cfg, d


def chords(p0, d, dims):
    """Path length (in grid units) of the ray p0 + t d inside each voxel, by a float64 sweep over
    the plane crossings."""
    ts = [0.0]
    for ax in range(3):
        if d[ax] != 0:
            for k in range(0, dims[ax] + 1):
                t = (k - p0[ax]) / d[ax]
                if t > 0:
                    ts.append(t)
    ts = np.unique(np.array(ts))
    out = {}
    for a, b in zip(ts[:-1], ts[1:]):
        mid = p0 + 0.5 * (a + b) * d
        ijk = np.floor(mid).astype(int)
        if np.any(ijk < 0) or np.any(ijk >= dims):
            break
        v = (ijk[0] * dims[1] + ijk[1]) * dims[2] + ijk[2]
        out[v] = out.get(v, 0.0) + (b - a) * np.linalg.norm(d)
    return out


def ballistic_replay(unit=0.5):
    cfg, d = ballistic_cfg(unit)
    lab = U.domain(cfg)
    prob = O.problem_from_config(cfg, lab)
    ids = np.arange(3, dtype=np.uint32)
    r = O.replay(prob, ids, np.arange(3, dtype=np.int32), np.ones(3), 3, nthreads=2)
    exact = chords(np.array([2.3, 3.7, 0.0]), d.astype(np.float64), np.array(lab.shape))
    return r, exact, unit


def test_ballistic_beam_gives_chord_lengths_per_voxel():
    r, exact, unit = ballistic_replay()
    ref = np.zeros(r["gmua"].shape[1])
    for v, L in exact.items():
        ref[v] = L * unit
    assert len(exact) > 20
    for k in range(3):                       # every photon flies the same straight line
        np.testing.assert_allclose(-r["gmua"][k], ref, atol=2e-5)
        np.testing.assert_allclose(r["gmus"][k], r["gmua"][k], atol=0)    # no scattering: n_v = 0
        assert set(np.flatnonzero(r["gmua"][k])) == set(exact)


def test_planted_voxel_shift_fails_the_chord_check():
    O.lib().oracle_set_mutant(4)
    try:
        r, exact, unit = ballistic_replay()
    finally:
        O.lib().oracle_set_mutant(0)
    ref = np.zeros(r["gmua"].shape[1])
    for v, L in exact.items():
        ref[v] = L * unit
    assert np.abs(-r["gmua"][0] - ref).max() > 0.05


def test_scatterings_sum_to_the_scatter_count_and_weights_replay_bitwise():
    g, prob = O.golden_problem("c4src")
    n = 20
    o = O.run(prob, nphoton=n, traces=True, nthreads=4)
    ids = np.arange(n, dtype=np.uint32)
    r = O.replay(prob, ids, np.arange(n, dtype=np.int32), np.ones(n), n, nthreads=4)
    lab = prob["labels"].reshape(-1) & 0x7FFFFFFF
    mus_mm = prob["media"][lab, 1].astype(np.float64) / float(prob["unitinmm"])
    with np.errstate(divide="ignore", invalid="ignore"):
        nv = np.where(mus_mm > 0, (r["gmus"] - r["gmua"]) * mus_mm, 0.0)
    np.testing.assert_allclose(nv.sum(axis=1), o["traces"]["vec"][:, 3].astype(np.float64), rtol=1e-9, atol=1e-6)
    # whole numbers of scatterings in every voxel
    np.testing.assert_allclose(nv, np.round(nv), atol=1e-6)
    assert np.array_equal(r["final_w"].view(np.uint32), o["traces"]["pos"][:, 3].copy().view(np.uint32))


def test_sources_table_replays_each_source():
    """A photon replayed from a source table entry equals the same photon of a run whose config
    holds that source."""
    g, prob = O.golden_problem("c4src")
    prob2 = copy.deepcopy(prob)
    prob2["srcpos"] = [np.float32(40.5), np.float32(20.5), np.float32(0), np.float32(1)]
    o2 = O.run(prob2, nphoton=50, traces=True, nthreads=2)
    table = np.array([[24.5, 24.5, 0, 0, 0, 1], [40.5, 20.5, 0, 0, 0, 1]], dtype=np.float32)
    ids = np.arange(50, dtype=np.uint32)
    r = O.replay(prob, ids, np.zeros(50, np.int32), np.ones(50), 1, sources=table,
                 src_index=np.ones(50, np.int32), nthreads=2)
    assert np.array_equal(r["final_w"].view(np.uint32), o2["traces"]["pos"][:, 3].copy().view(np.uint32))
