"""GPU parity: the gfx950 kernel (through libumcx.so's C ABI) against the oracle and the
reference's golden vectors.

Bars:
* per photon: position, direction, scatter count, time of flight, path, medium and weight bitwise
  equal to the oracle;
* detected-photon rows: bitwise equal to the reference's rows as a multiset;
* fluence: relative L2 <= 1e-4 against the reference's volume;
* energy balance from fp64 tallies: |deposited + escaped - launched| <= 1e-6 launched;
* absorbed fraction within 1e-3 percentage points of the reference's printed value.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U
from umcx_amd import bjdata

pytestmark = pytest.mark.gpu

TRACE_CASES = ["c1", "cube60b", "c2", "c3", "isotropic", "cone", "cube60planar", "fourier", "sphshells",
               "spherebox", "zipvol", "c4src", "skinvessel", "gated", "longx"]
DETP_CASES = ["cube60", "cube60b", "cubesph60b", "cube60planar", "sphshells", "spherebox", "c2", "c4src", "zipvol",
              "fourier", "save2pt0", "gated"]
FLUENCE_CASES = ["c1", "cube60", "cube60b", "cubesph60b", "cube60planar", "sphshells", "spherebox", "skinvessel",
                 "c2", "c3", "c4src", "zipvol", "isotropic", "cone", "fourier", "gated", "normx", "longx"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sorted_rows(rows):
    return rows[np.lexsort(rows.T[::-1])] if rows.size else rows


@pytest.mark.parametrize("which", [0, 1, 2, 3, 4])
def test_short_correctly_rounded_sequences_are_bitwise(which):
    """The short sqrt, reciprocal and division sequences the photon loop uses away from special
    operands equal the compiler's correctly rounded ones: every float for sqrt (1.9e9 inputs), the
    direction reciprocal (2.0e9) and the reciprocal's whole range (4.2e9), 2^32 hashed pairs and
    every b in [1, 2) against eight edge mantissas of a (1.3e8) for the division."""
    import ctypes as C
    from umcx_amd._native import lib
    say(f"arith self-check {which}")
    tested, bad = C.c_uint64(0), C.c_uint64(0)
    assert lib().umcx_arith_check_device(which, C.byref(tested), C.byref(bad), 0) == 0, U.last_error()
    assert tested.value > 100_000_000
    assert bad.value == 0


def test_device_libm_bitwise_on_reachable_inputs():
    say("device libm replicas vs host libm")
    for fn, x in O.reachable_inputs().items():
        dev = U.libm_eval(fn, x, device=0)
        ref = O.host_libm(fn, x)
        assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32)), f"fn {fn}"


@pytest.mark.parametrize("label_cache", [0, -1])
@pytest.mark.parametrize("name", TRACE_CASES)
def test_per_photon_state_matches_oracle(name, label_cache):
    """label_cache 0: media ids staged in LDS where the volume fits; -1: labels read from HBM."""
    g, prob = O.golden_problem(name)
    n = 3000 if name not in ("c3", "skinvessel") else 1500
    say(f"trace parity {name} n={n} label_cache {label_cache}")
    o = O.run(prob, nphoton=n, traces=True, nthreads=8)["traces"]
    with U.Simulation(g["cfg"], trace=True, vote=32 if name in ("c2", "cube60b", "c3") else 0,
                      label_cache=label_cache) as sim:
        sim.run(0, n)
        t = sim.trace(n)
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    assert np.array_equal(exact(t["pos"][:, :3]), exact(o["pos"][:, :3]))
    assert np.array_equal(exact(t["vec"]), exact(o["vec"]))
    assert np.array_equal(exact(t["len"][:, :3]), exact(o["len"][:, :3]))
    assert np.array_equal(t["mediaid"], o["mediaid"])
    assert np.array_equal(t["lastvoxelidx"], o["lastvoxelidx"])
    assert np.array_equal(exact(t["w0"]), exact(o["w0"]))
    assert np.array_equal(exact(t["pos"][:, 3]), exact(o["pos"][:, 3]))
    assert np.array_equal(exact(t["len"][:, 3]), exact(o["len"][:, 3]))


@pytest.mark.parametrize("name,param,value", [("fourier", ("Param1", 3), 37.3), ("fourier", ("Param2", 3), 41.6),
                                              ("cone", ("Param1", 0), 200.5)])
def test_source_angles_past_120_match_oracle(name, param, value):
    """User-set source angles reach sinf/cosf arguments of 120 and more, where the replicas take
    the large-argument reduction: fourier phases (2 pi (kx u + ky v + frac)) with kx or ky near 40
    and a cone half-angle of 200.5 rad; per-photon states, launch weights included, bitwise equal
    to the oracle's."""
    import copy
    g, _ = O.golden_problem(name)
    cfg = copy.deepcopy(g["cfg"])
    cfg["Optode"]["Source"][param[0]][param[1]] = value
    prob = O.problem_from_config(cfg, U.domain(cfg))
    n = 3000
    say(f"trace parity {name} {param} = {value}")
    o = O.run(prob, nphoton=n, traces=True, nthreads=8)["traces"]
    with U.Simulation(cfg, trace=True) as sim:
        sim.run(0, n)
        t = sim.trace(n)
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len", "w0"):
        assert np.array_equal(exact(t[key]), exact(o[key])), key
    assert np.array_equal(t["lastvoxelidx"], o["lastvoxelidx"])
    if name == "fourier":                     # the launch weights spread over (0, 1]: the phase matters
        assert len(np.unique(t["w0"])) > 100


@pytest.mark.parametrize("label_cache", [0, -1])
@pytest.mark.parametrize("name", DETP_CASES)
def test_detected_rows_match_reference(name, label_cache):
    g = O.golden(name)
    say(f"detp parity {name} label_cache {label_cache}")
    with U.Simulation(g["cfg"], label_cache=label_cache) as sim:
        sim.run()
        rows, ids, w = sim.detp()
        detected, saved = sim.detected()
    assert saved == int(g["detected"])
    got = sorted_rows(rows)
    assert got.shape == g["detp"].shape
    assert np.array_equal(got.view(np.uint32), g["detp"].view(np.uint32))
    assert len(np.unique(ids)) >= 1 and ids.max() < int(float(g["cfg"]["Session"]["Photons"]))


def test_label_staging_choice():
    """Volumes whose media ids pack into <= 96 KiB are staged in LDS (2 bits for a run of at most 4
    ids, 4 bits for 16: sphshells uses ids 1-4 of its 5 media, 2 bits); larger ones, and
    label_cache -1, are not."""
    expect = {"c2": 2, "c4src": 2, "cube60": 2, "zipvol": 2, "spherebox": 2, "sphshells": 2, "c3": 0}
    for name, bits in expect.items():
        g = O.golden(name)
        with U.Simulation(g["cfg"]) as sim:
            assert sim.info["lds_label_bits"] == bits, name
        with U.Simulation(g["cfg"], label_cache=-1) as sim:
            assert sim.info["lds_label_bits"] == 0, name


@pytest.mark.parametrize("name", ["c2", "c4src", "zipvol", "cube60planar", "gated", "cube60b"])
def test_one_bit_labels_after_the_pilot_match_oracle(name):
    """A volume holding at most two media ids also has a 1-bit label form; after the pilot, when
    most deposits fall outside the tally tile, the session stages that form and grows the tile.
    Every photon must still match the oracle."""
    g, prob = O.golden_problem(name)
    n = 3000
    say(f"trace parity after the pilot {name}")
    o = O.run(prob, nphoton=n, traces=True, nthreads=8)["traces"]
    with U.Simulation(g["cfg"], trace=True, tally_mode=1) as sim:
        assert sim.info["lds_label_bits"] == 2
        sim.run(0, n)
        t = sim.trace(n)
        bits = sim.lds_label_bits()
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len"):
        assert np.array_equal(exact(t[key]), exact(o[key])), key
    assert np.array_equal(t["mediaid"], o["mediaid"])
    assert np.array_equal(t["lastvoxelidx"], o["lastvoxelidx"])
    if name == "c2":
        assert bits == 1     # 29 % of the pilot's deposits in the tile: switched
    assert bits in (1, 2)


def layered_config(nphoton=2000):
    """40^3 voxels in five layers of different media and refractive index (4-bit LDS labels),
    Fresnel reflection on, two detectors."""
    return {"Session": {"ID": "layers5", "Photons": nphoton, "RNGSeed": 29012, "DoMismatch": 1, "SaveDetFlag": 5},
            "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-10},
            "Domain": {"Dim": [40, 40, 40], "LengthUnit": 1,
                       "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1},
                                 {"mua": 0.01, "mus": 2.0, "g": 0.8, "n": 1.33},
                                 {"mua": 0.02, "mus": 5.0, "g": 0.9, "n": 1.45},
                                 {"mua": 0.005, "mus": 1.0, "g": 0.7, "n": 1.37},
                                 {"mua": 0.05, "mus": 10.0, "g": 0.95, "n": 1.40},
                                 {"mua": 0.001, "mus": 0.5, "g": 0.5, "n": 1.50}]},
            "Optode": {"Source": {"Type": "pencil", "Pos": [20, 20, 0], "Dir": [0, 0, 1]},
                       "Detector": [{"Pos": [14, 20, 0], "R": 2}, {"Pos": [26, 20, 0], "R": 2}]},
            "Shapes": [{"Grid": {"Tag": 1, "Size": [40, 40, 40]}},
                       {"ZLayers": [[8, 16, 2], [16, 24, 3], [24, 32, 4], [32, 40, 5]]}]}


def test_four_bit_lds_labels_match_oracle():
    cfg = layered_config()
    prob = O.problem_from_config(cfg, U.domain(cfg))
    n = cfg["Session"]["Photons"]
    say("trace parity layered 5 media (4-bit LDS labels)")
    o = O.run(prob, nphoton=n, traces=True, nthreads=8)
    with U.Simulation(cfg, trace=True) as sim:
        assert sim.info["lds_label_bits"] == 4
        sim.run(0, n)
        t = sim.trace(n)
        rows = sorted_rows(sim.detp()[0])
    ot = o["traces"]
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len"):
        assert np.array_equal(exact(t[key]), exact(ot[key])), key
    assert np.array_equal(t["mediaid"], ot["mediaid"])
    assert np.array_equal(rows.view(np.uint32), sorted_rows(o["detp"]).view(np.uint32))


def many_media_config(nmed: int, seed: int = 5):
    """28 x 24 x 20 voxels in 4^3 blocks of `nmed` media (each block a random medium, refractive
    indices 1.2-1.5, Fresnel on), pencil source, three detectors saving every partial path."""
    rng = np.random.default_rng(seed)
    dims = [28, 24, 20]
    blocks = rng.integers(1, nmed + 1, size=(7, 6, 5))
    vol = np.repeat(np.repeat(np.repeat(blocks, 4, 0), 4, 1), 4, 2).astype(np.int32)
    media = [{"mua": 0.0, "mus": 0.0, "g": 1.0, "n": 1.0}]
    for _ in range(nmed):
        media.append({"mua": float(rng.uniform(0.0, 0.05)), "mus": float(rng.uniform(0.5, 4.0)),
                      "g": float(rng.uniform(-0.5, 0.9)), "n": float(rng.uniform(1.2, 1.5))})
    return {"Session": {"ID": f"media{nmed}", "Photons": 1500, "RNGSeed": 4711 + nmed, "DoMismatch": 1,
                        "DoPartialPath": 1, "SaveDetFlag": 5, "OutputType": "x", "DoNormalize": 0},
            "Forward": {"T0": 0, "T1": 2e-9, "Dt": 5e-10},
            "Domain": {"Dim": dims, "LengthUnit": 1, "Media": media},
            "Optode": {"Source": {"Type": "pencil", "Pos": [14, 12, 0], "Dir": [0, 0, 1]},
                       "Detector": [{"Pos": [10, 12, 0], "R": 3}, {"Pos": [18, 12, 0], "R": 3},
                                    {"Pos": [14, 6, 0], "R": 3}]},
            "Shapes": {"_ArraySize_": dims, "_ArrayType_": "int32", "_ArrayData_": vol.reshape(-1).tolist()}}


@pytest.mark.parametrize("nmed", [20, 140])
def test_many_media_match_oracle(nmed):
    """20 media: byte labels read from HBM (too many ids to pack), partial paths of 20 media per
    lane in LDS; 140 media: 32-bit labels. Per-photon states, detected rows (with a partial path
    per medium) and the fluence equal the oracle's."""
    cfg = many_media_config(nmed)
    prob = O.problem_from_config(cfg, U.domain(cfg))
    n = cfg["Session"]["Photons"]
    say(f"trace parity {nmed} media")
    o = O.run(prob, nphoton=n, traces=True, nthreads=8)
    with U.Simulation(cfg, trace=True) as sim:
        assert sim.info["lds_label_bits"] == 0
        assert sim.info["label_bytes"] == (4 if nmed > 128 else 1)
        assert sim.info["ppath_columns"] == nmed + 1      # media 0..nmed (:157)
        sim.run(0, n)
        t = sim.trace(n)
        rows = sorted_rows(sim.detp()[0])
        flu = sim.fluence().astype(np.float64)
    ot = o["traces"]
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len", "w0"):
        assert np.array_equal(exact(t[key]), exact(ot[key])), key
    assert np.array_equal(t["mediaid"], ot["mediaid"])
    assert np.array_equal(t["lastvoxelidx"], ot["lastvoxelidx"])
    assert len(rows) > 10 and rows.shape[1] == 2 + nmed
    assert np.array_equal(rows.view(np.uint32), sorted_rows(o["detp"]).view(np.uint32))
    of = np.asarray(o["fluence"], dtype=np.float64).reshape(flu.shape)
    assert np.linalg.norm(flu - of) <= 1e-5 * np.linalg.norm(of)


def test_unusual_refractive_indices_match_oracle():
    """Layers with refractive indices far from tissue values (0.45 and 4.5 beside 1.37): total
    internal reflection at most angles and large transmission bends; every photon matches the
    oracle."""
    cfg = layered_config()
    cfg["Session"]["ID"] = "layers_n"
    cfg["Domain"]["Media"][2]["n"] = 4.5
    cfg["Domain"]["Media"][4]["n"] = 0.45
    prob = O.problem_from_config(cfg, U.domain(cfg))
    n = cfg["Session"]["Photons"]
    say("trace parity with refractive indices 0.45 and 4.5")
    o = O.run(prob, nphoton=n, traces=True, nthreads=8)
    with U.Simulation(cfg, trace=True) as sim:
        sim.run(0, n)
        t = sim.trace(n)
        rows = sorted_rows(sim.detp()[0])
    ot = o["traces"]
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len"):
        assert np.array_equal(exact(t[key]), exact(ot[key])), key
    assert np.array_equal(t["mediaid"], ot["mediaid"])
    assert np.array_equal(rows.view(np.uint32), sorted_rows(o["detp"]).view(np.uint32))


def test_detector_overflow_keeps_genuine_rows():
    g, prob = O.golden_problem("overflow")
    with U.Simulation(g["cfg"]) as sim:
        sim.run()
        rows, ids, w = sim.detp()
        detected, saved = sim.detected()
    full = O.run(dict(prob, maxdetphotons=10 ** 6), nthreads=8)
    assert saved == 10 and detected == full["detected"]
    allrows = {r.tobytes() for r in full["detp"]}
    assert all(r.tobytes() in allrows for r in rows)


@pytest.mark.parametrize("tally", [0, 1])
@pytest.mark.parametrize("name", FLUENCE_CASES)
def test_fluence_and_energy_match_reference(name, tally):
    """tally 0: every deposit is a global atomic; tally 1: deposits inside the pilot-chosen box
    accumulate in LDS first."""
    g, prob = O.golden_problem(name)
    say(f"fluence parity {name} tally {tally}")
    with U.Simulation(g["cfg"], tally_mode=tally) as sim:
        sim.run()
        if tally == 1 and prob["issavevol"]:
            assert sim.tally_box() is not None
        normalize = bool(prob["isnormalized"]) and "-U" not in g["argv"]
        res = sim.result(normalize=normalize)
    assert res.energy_total == pytest.approx(float(g["energy"]), abs=0.006)
    assert abs(res.absorbed_percent - float(g["absorbed"])) <= 1e-3
    if prob["srctype"] == 5:
        # fourier sources launched outside the grid: skip() restarts every photon at weight 1
        # (src/umcx.cpp:353) while the launch weight is what energytotal counts (:762, :767)
        launched = float(int(float(g["cfg"]["Session"]["Photons"])))
    else:
        launched = res.energy_total
    assert abs(res.energy_deposit + res.energy_escape - launched) <= 1e-6 * launched
    f = res.fluence.reshape(-1)
    if "fluence" in g:
        ref = g["fluence"]
        assert np.linalg.norm(f - ref) / np.linalg.norm(ref) <= 1e-4
    else:
        idx, ref = g["sample_idx"], g["sample_val"]
        assert np.linalg.norm(f[idx] - ref) / np.linalg.norm(ref) <= 1e-4
        ss = float((f.astype(np.float64) ** 2).sum())
        assert abs(ss - float(g["sumsq"])) / float(g["sumsq"]) <= 2e-4
    np.testing.assert_allclose(res.fluence.astype(np.float64).sum(axis=(1, 2, 3)), g["gate_sums"], rtol=1e-4,
                               atol=1e-30)


@pytest.mark.parametrize("name", ["cube60", "cube60b", "c2", "c3"])
def test_command_line_outputs_are_byte_compatible(name, tmp_path):
    g = O.golden(name)
    argv = list(g["argv"])
    if "INPUT.json" in argv:
        (tmp_path / "INPUT.json").write_text(json.dumps(g["cfg"]))
    say(f"CLI outputs {name}")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        assert U.run_cmd(argv) == 0, U.last_error()
    finally:
        os.chdir(cwd)
    bnii = (tmp_path / f"{name}.bnii").read_bytes()
    assert len(bnii) == int(g["bnii_size"])
    head = bytes(g["bnii_head"])
    hlen = bnii.find(b"NIFTIData") + 30       # header up to the raw payload
    assert bnii[:hlen] == head[:hlen]
    if "detp" in g:
        raw = (tmp_path / f"{name}_detp.jdb").read_bytes()
        assert len(raw) == int(g["detp_size"])
        dh = bytes(g["detp_head"])
        payload = len(raw) - int(g["detp"].size) * 4      # rows follow the header; their order varies
        assert raw[:payload] == dh[:payload]
        info, rows = bjdata.load_detp(str(tmp_path / f"{name}_detp.jdb"))
        assert np.array_equal(sorted_rows(rows).view(np.uint32), g["detp"].view(np.uint32))


def test_zero_photons_write_the_reference_files(tmp_path, capfd):
    """`-n 0`: the reference still writes both files. Its normaliser is 1/0 = inf, so every
    fluence value is 0 * inf, a NaN with the host's bit pattern; the detected-photon file holds no
    rows; the summary line reports zero energy and a NaN absorbed fraction."""
    g = O.golden("zerophotons")
    argv = list(g["argv"])
    say("CLI outputs with zero photons")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        assert U.run_cmd(argv) == 0, U.last_error()
    finally:
        os.chdir(cwd)
    line = capfd.readouterr().out.strip().splitlines()[-1]
    ref = str(g["line"])
    cut = lambda s: (s.split(", duration")[0], s.split(" ms, ")[1])       # the duration differs
    assert cut(line) == cut(ref), (line, ref)
    raw = (tmp_path / "zerophotons.bnii").read_bytes()
    assert len(raw) == int(g["bnii_size"]) and raw[:128] == bytes(g["bnii_head"])
    vals = bjdata.load_bnii(str(tmp_path / "zerophotons.bnii"))
    flat = vals.transpose(3, 0, 1, 2).reshape(-1)
    assert np.isnan(flat).all()
    assert np.array_equal(flat.view(np.uint32), g["fluence"].view(np.uint32))
    detp = (tmp_path / "zerophotons_detp.jdb").read_bytes()
    assert len(detp) == int(g["detp_size"]) and detp == bytes(g["detp_head"])[:len(detp)]


def test_photon_ranges_compose():
    """Photon i depends only on i: two half runs give the same detections as one full run."""
    g = O.golden("c2")
    cfg = dict(g["cfg"])
    n = 400000
    say("range composition c2")
    with U.Simulation(cfg) as sim:
        sim.run(0, n)
        full_rows = sorted_rows(sim.detp()[0])
        full_e = sim.energies()
        sim.reset()
        sim.run(0, n // 2)
        sim.run(n // 2, n - n // 2)
        split_rows = sorted_rows(sim.detp()[0])
        split_e = sim.energies()
    assert np.array_equal(full_rows.view(np.uint32), split_rows.view(np.uint32))
    assert split_e[1] == pytest.approx(full_e[1], rel=1e-12)
    assert split_e[0] == pytest.approx(full_e[0], rel=1e-9)


def test_sessions_run_concurrently_from_threads():
    """Separate sessions used from several host threads at once (the C calls release the GIL):
    every session's detected rows, energies and fluence equal a run of the same session alone."""
    from concurrent.futures import ThreadPoolExecutor
    names = ["cube60b", "c2", "spherebox", "c4src"]

    def one(name):
        g = O.golden(name)
        with U.Simulation(g["cfg"]) as sim:
            out = []
            for _ in range(3):
                sim.reset()
                sim.run()
                out.append((sorted_rows(sim.detp()[0]), sim.energies(), sim.fluence()))
            return out

    say("sessions alone")
    alone = {n: one(n) for n in names}
    say("sessions from 4 threads at once")
    with ThreadPoolExecutor(max_workers=4) as ex:
        together = dict(zip(names, ex.map(one, names)))
    for n in names:
        rows0, e0, f0 = alone[n][0]
        for rows, e, f in alone[n][1:] + together[n]:
            assert np.array_equal(rows.view(np.uint32), rows0.view(np.uint32)), n
            assert e[1] == e0[1] and e[0] == pytest.approx(e0[0], rel=1e-10), n
            # float32 sums in another order: the 1e-4 fluence bar (c2 without a tile reached 1.4e-5)
            assert np.linalg.norm((f - f0).ravel()) <= 1e-4 * np.linalg.norm(f0.ravel()), n


def test_launch_shape_does_not_change_results():
    g = O.golden("cube60b")
    outs = []
    for block, blocks, vote, lc in ((64, 7, 0, 0), (256, 0, 0, -1), (128, 33, 0, 0), (256, 0, 24, -1), (256, 0, 48, 0),
                                    (64, 5, 64, 0), (1024, 0, 0, 0), (512, 3, 0, 0)):
        with U.Simulation(g["cfg"], block=block, blocks=blocks, vote=vote, label_cache=lc) as sim:
            sim.run()
            outs.append((sorted_rows(sim.detp()[0]), sim.energies()))
    for rows, e in outs[1:]:
        assert np.array_equal(rows.view(np.uint32), outs[0][0].view(np.uint32))
        assert e[1] == outs[0][1][1]


@pytest.mark.parametrize("label_cache", [0, -1])
def test_short_runs_use_half_the_lanes_with_equal_results(label_cache):
    """A run under 12 photons per lane of the full shape launches half the lanes (half-size
    workgroups when one workgroup per CU holds the staged labels, else half the workgroups);
    every photon's final state equals the full shape's bitwise."""
    g = O.golden("cube60b")
    n = 3000
    say(f"short-run shape, label_cache {label_cache}")
    with U.Simulation(g["cfg"], trace=True, label_cache=label_cache) as sim:
        block, blocks = sim.info["block"], sim.info["blocks"]
        sim.run(0, n)
        short = sim.last_launch()
        t_short = sim.trace(n)
    if sim.info["lds_label_bits"]:
        assert short == (block // 2, blocks)
    else:
        assert short == (block, blocks // 2)
    with U.Simulation(g["cfg"], trace=True, label_cache=label_cache, block=block, blocks=blocks) as sim:
        sim.run(0, n)
        assert sim.last_launch() == (block, blocks)
        t_full = sim.trace(n)
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len", "w0"):
        assert np.array_equal(exact(t_short[key]), exact(t_full[key])), key
    assert np.array_equal(t_short["lastvoxelidx"], t_full["lastvoxelidx"])


def test_deposit_cache_changes_only_summation_order():
    """The per-lane cache of global deposits (on, off): same detections and energies, fluence
    equal up to float summation order (float32 sums of ~1e5 deposits in the hottest cells differ
    by ~1e-5 relative between orders, so the bar is the fluence bar, 1e-4 relative L2)."""
    g = O.golden("c2")
    n = 200000
    say("deposit cache on/off c2")
    out = {}
    for mode in (1, -1):
        with U.Simulation(g["cfg"], deposit_cache=mode, tally_mode=1) as sim:
            assert sim.deposit_cache() == (mode == 1)
            sim.run(0, n)
            out[mode] = (sorted_rows(sim.detp()[0]), sim.energies(), sim.fluence().astype(np.float64))
    assert np.array_equal(out[1][0].view(np.uint32), out[-1][0].view(np.uint32))
    assert out[1][1][1] == out[-1][1][1]
    a, b = out[1][2], out[-1][2]
    assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b)
    # nothing lost in eviction or the final flush: each deposit is the weight drop dw divided by
    # the voxel's mua (fluence output, every medium here has mua >= FLT_EPSILON), so the cached
    # tallies times mua sum to the kernel's fp64 deposited weight
    lab = U.domain(g["cfg"]) & 0x7FFFFFFF
    mua = O.problem_from_config(g["cfg"], U.domain(g["cfg"]))["media"][:, 0].astype(np.float64)
    per_vox = a.sum(axis=0).reshape(-1)
    assert float((per_vox * mua[lab.reshape(-1)]).sum()) == pytest.approx(out[1][1][2], rel=1e-5)


def test_deposit_cache_fluence_matches_reference_golden():
    """Forced on, the deposit cache's fluence equals the reference binary's at the fluence bar."""
    g, prob = O.golden_problem("c2")
    say("deposit cache on vs reference golden c2")
    with U.Simulation(g["cfg"], deposit_cache=1) as sim:
        assert sim.deposit_cache()
        sim.run()
        res = sim.result(normalize=bool(prob["isnormalized"]))
    f = res.fluence.reshape(-1)
    ref = g["fluence"] if "fluence" in g else None
    if ref is not None:
        assert np.linalg.norm(f - ref) / np.linalg.norm(ref) <= 1e-4
    else:
        idx, val = g["sample_idx"], g["sample_val"]
        assert np.linalg.norm(f[idx] - val) / np.linalg.norm(val) <= 1e-4
    np.testing.assert_allclose(res.fluence.astype(np.float64).sum(axis=(1, 2, 3)), g["gate_sums"], rtol=1e-4,
                               atol=1e-30)


def test_full_size_heterogeneous_reflective_properties():
    """The 1e8-photon heterogeneous reflective workload at 1e7 photons: energy balance, and the
    detections of a 2e5-photon prefix equal the oracle's bitwise."""
    g, prob = O.golden_problem("c2")
    cfg = json.loads(json.dumps(g["cfg"]))
    cfg["Session"]["Photons"] = 10_000_000
    say("full-size c2 1e7")
    with U.Simulation(cfg) as sim:
        sim.run()
        assert sim.last_launch() == (sim.info["block"], sim.info["blocks"])   # full shape
        assert sim.deposit_cache()            # the LDS tile takes under half of the deposits here
        esc, tot, dep = sim.energies()
        detected, saved = sim.detected()
        sim.reset()
        sim.run(0, 200000)
        rows = sorted_rows(sim.detp()[0])
    assert tot == pytest.approx(1e7)
    assert abs(dep + esc - tot) <= 1e-6 * tot
    assert 4000e-6 < detected / 1e7 < 5000e-6
    o = O.run(prob, nphoton=200000, nthreads=8)
    assert np.array_equal(rows.view(np.uint32), sorted_rows(o["detp"]).view(np.uint32))
