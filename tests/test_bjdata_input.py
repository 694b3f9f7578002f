"""BJData config input (MCX_run_simulation with argn < 0, src/umcx.cpp:475-478: argv[0] is a
BJData buffer of -argn bytes, as the reference's MATLAB binding passes it).

The fixtures (tests/golden/bjdata_inputs.npz, written by tests/golden/make_bjdata.py) hold configs
encoded by the reference's vendored JSON library and hand-built buffers over the rest of the BJData
grammar, each with the value that library's json::from_bjdata reads from it (or its rejection)."""
import json

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U

FIX = np.load(O.os.path.join(O.ROOT, "tests", "golden", "bjdata_inputs.npz"), allow_pickle=False)
NAMES = [str(n) for n in FIX["names"]]


def buf(name):
    return FIX[f"bjd_{name}"].tobytes()


@pytest.mark.parametrize("name", NAMES)
def test_reads_like_the_reference_json_library(name):
    expected = str(FIX[f"json_{name}"])
    if not expected:
        with pytest.raises(U.UmcxError, match="BJData"):
            U.config_from_bjdata(buf(name))
        return
    got = U.config_from_bjdata(buf(name))
    # same values, same key order, same integer/float kinds
    assert json.dumps(got) == json.dumps(json.loads(expected))


@pytest.mark.parametrize("name", [n for n in NAMES if n.startswith("cfg_")])
def test_configs_give_the_json_path_domain(name):
    src = json.loads(str(FIX[f"src_{name}"]))
    got = U.config_from_bjdata(buf(name))
    assert np.array_equal(U.domain(got), U.domain(src))


def test_truncated_and_corrupted_buffers_fail_cleanly():
    rng = np.random.default_rng(5)
    for name in ("cfg_c2_11", "cfg_jdatavol_11", "cfg_zipbytes_00", "gram_noops_and_scalars", "gram_nd_int32_2x2x2"):
        b = buf(name)
        for k in range(1, len(b)):
            with pytest.raises(U.UmcxError):
                U.config_from_bjdata(b[:k])
        # byte flips: either a clean error or some value, never a crash
        for _ in range(300):
            m = bytearray(b)
            for _ in range(int(rng.integers(1, 4))):
                m[int(rng.integers(0, len(m)))] = int(rng.integers(0, 256))
            try:
                U.config_from_bjdata(bytes(m))
            except U.UmcxError:
                pass


def test_empty_buffer_is_rejected():
    with pytest.raises(U.UmcxError):
        U.config_from_bjdata(b"")


def sorted_rows(rows):
    return rows[np.lexsort(rows.T[::-1])] if rows.size else rows


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cfg_c2_11", "cfg_jdatavol_00", "cfg_zipbytes_10"])
def test_bjdata_session_runs_like_the_json_session(name):
    """Every photon's final state from a BJData-built session equals the JSON-built one's."""
    src = json.loads(str(FIX[f"src_{name}"]))
    n = 3000
    print(f"BJData session {name}", flush=True)
    outs = []
    for kw in ({"bjdata": buf(name)}, {"cfg": src}):
        with U.Simulation(trace=True, **kw) as sim:
            sim.run(0, n)
            outs.append(sim.trace(n))
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len", "w0"):
        assert np.array_equal(exact(outs[0][key]), exact(outs[1][key])), key
    assert np.array_equal(outs[0]["lastvoxelidx"], outs[1]["lastvoxelidx"])


@pytest.mark.gpu
def test_run_simulation_with_a_bjdata_buffer_writes_the_reference_outputs(tmp_path):
    """MCX_run_simulation(argv, -len) on the heterogeneous reflective config encoded by the
    reference's JSON library: the .bnii header and the detected rows equal the reference's files
    for the same config (1e6 photons)."""
    from umcx_amd import bjdata
    g = O.golden("c2")
    print("MCX_run_simulation BJData c2", flush=True)
    cwd = O.os.getcwd()
    O.os.chdir(tmp_path)
    try:
        assert U.run_bjdata(buf("cfg_c2_11")) == 0, U.last_error()
    finally:
        O.os.chdir(cwd)
    bnii = (tmp_path / "c2.bnii").read_bytes()
    assert len(bnii) == int(g["bnii_size"])
    hlen = bnii.find(b"NIFTIData") + 30
    assert bnii[:hlen] == bytes(g["bnii_head"])[:hlen]
    raw = (tmp_path / "c2_detp.jdb").read_bytes()
    assert len(raw) == int(g["detp_size"])
    payload = len(raw) - int(g["detp"].size) * 4
    assert raw[:payload] == bytes(g["detp_head"])[:payload]
    info, rows = bjdata.load_detp(str(tmp_path / "c2_detp.jdb"))
    assert np.array_equal(sorted_rows(rows).view(np.uint32), g["detp"].view(np.uint32))
