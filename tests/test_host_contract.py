"""Host side of the drop-in boundary, checked without a GPU.

* libumcx.so loads and exports every function include/umcx.h declares;
* the label volume (shapes, JData volumes incl. zlib+base64, detector masks) equals the
  reference's own --dumpmask output for every golden case, and the --dumpmask file is
  byte-identical;
* command-line flags, presets and --json merges expand to the reference's --dumpjson config;
* the libm replicas equal the host C library bit for bit on every argument the photon loop can
  form (run on the CPU here; the GPU build is checked in test_gpu_parity.py);
* seed words equal glibc srand()/rand();
* failures come back as error codes with the reference's messages, never as exceptions;
* without a GPU the simulation path fails loudly instead of falling back to the CPU.
"""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U
from umcx_amd import _native, bjdata

ROOT = O.ROOT
GOLDEN = ["c1", "cube60", "cube60b", "cubesph60b", "cube60planar", "sphshells", "spherebox", "skinvessel", "c2",
          "c3", "c4src", "zipvol", "isotropic", "cone", "fourier", "save2pt0", "overflow", "gated", "normx",
          "zerophotons", "longx"]
BENCH_ARGV = ["c1", "cube60", "cube60b", "cubesph60b", "cube60planar", "sphshells", "spherebox", "skinvessel",
              "isotropic", "cone", "fourier", "save2pt0", "overflow", "gated", "normx", "zerophotons"]


def header_functions():
    with open(os.path.join(ROOT, "include", "umcx.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    c_part = text.split("#ifdef __cplusplus\n}")[0]
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", c_part, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert "MCX_run_cmd" in names and "MCX_run_json" in names and len(names) >= 20
    L = _native.lib()
    for n in names:
        assert hasattr(L, n), n
    cxx = "_Z18MCX_run_simulationPPciiPPv"     # the C++ linkage entry, bound by its mangled name
    assert set(names) == set(_native.SIGNATURES) - {cxx}
    nm = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True, text=True).stdout
    assert cxx in nm     # C++ linkage entry of src/umcx.cpp:783


def test_binding_structs_match_the_header(tmp_path):
    """The ctypes mirrors of the C structs have the header's sizes and field offsets."""
    checks = {"umcx_launch": _native.UmcxLaunch, "umcx_info": _native.UmcxInfo, "umcx_batch_trial": _native.UmcxBatchTrial}
    src = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{os.path.join(ROOT, "include", "umcx.h")}"',
           "int main(void) {"]
    for cname, py in checks.items():
        src.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            src.append(f'    printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    src += ["    return 0;", "}"]
    (tmp_path / "layout.c").write_text("\n".join(src))
    subprocess.run(["gcc", "-o", str(tmp_path / "layout"), str(tmp_path / "layout.c")], check=True)
    out = subprocess.run([str(tmp_path / "layout")], capture_output=True, text=True, check=True).stdout.split()
    got = dict(zip(out[::2], (int(v) for v in out[1::2])))
    for cname, py in checks.items():
        assert got[cname] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[f"{cname}.{f}"] == getattr(py, f).offset, f"{cname}.{f}"


@pytest.mark.parametrize("name", GOLDEN)
def test_domain_matches_reference_dumpmask(name):
    g = O.golden(name)
    lab = U.domain(g["cfg"])
    assert lab.shape == g["mask"].shape
    assert np.array_equal(lab, g["mask"])


@pytest.mark.parametrize("name", BENCH_ARGV)
def test_cli_expansion_matches_reference_dumpjson(name):
    g = O.golden(name)
    cfg = U.config_from_argv(g["argv"])
    assert cfg == g["cfg"]


@pytest.mark.parametrize("name", ["cube60b", "c3", "skinvessel", "zipvol"])
def test_dumpmask_file_is_byte_compatible(name, tmp_path):
    g = O.golden(name)
    argv = list(g["argv"])
    if "INPUT.json" in argv:
        (tmp_path / "INPUT.json").write_text(json.dumps(g["cfg"]))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        assert U.run_cmd(argv + ["--dumpmask"]) == 0, U.last_error()
    finally:
        os.chdir(cwd)
    path = tmp_path / f"{name}_vol.bnii"
    raw = path.read_bytes()
    assert len(raw) == int(g["mask_size"])
    assert raw[:128] == bytes(g["mask_head"])
    vol = bjdata.load_bnii(str(path))[..., 0]
    assert np.array_equal(vol, g["mask"])


def test_libm_replicas_match_host_libm_on_reachable_inputs():
    inputs = O.reachable_inputs()
    for fn, x in inputs.items():
        mine = U.libm_eval(fn, x)
        ref = O.host_libm(fn, x)
        assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32)), f"fn {fn}"
    # sin/cos also take acos outputs (src/umcx.cpp:372): every float in [0, pi]
    th = np.arange(0, np.float32(np.pi).view(np.uint32) + 1, dtype=np.uint32).view(np.float32)
    for fn in (1, 2):
        assert np.array_equal(U.libm_eval(fn, th).view(np.uint32), O.host_libm(fn, th).view(np.uint32))


def test_sin_cos_replicas_on_a_sample_of_every_float():
    """Every 1009th bit pattern of all finite floats (both signs, arguments of 120 and more
    included, where the large-argument reduction runs) against the host library; the GPU test
    covers every float."""
    for lo, hi in ((0x00000000, 0x7F800000), (0x80000000, 0xFF800000)):
        x = np.arange(lo, hi + 1, 1009, dtype=np.uint64).astype(np.uint32).view(np.float32)
        for fn in (1, 2):
            mine, ref = U.libm_eval(fn, x), O.host_libm(fn, x)
            diff = (mine.view(np.uint32) != ref.view(np.uint32)) & ~(np.isnan(mine) & np.isnan(ref))
            assert not diff.any(), (fn, x[diff][:3])


def test_replica_check_has_teeth():
    """(float)log(double) differs from glibc's logf on ~61k reachable inputs; the replica must not."""
    x = O.reachable_inputs()[0]
    cr = np.log(x.astype(np.float64)).astype(np.float32)
    ref = O.host_libm(0, x)
    assert np.count_nonzero(cr.view(np.uint32) != ref.view(np.uint32)) > 50000


def test_seed_words():
    with open(os.path.join(ROOT, "tests", "golden", "seed_words.json")) as f:
        kat = json.load(f)
    for seed, words in kat.items():
        assert U.seed_words(int(seed)) == words


def test_error_paths_return_codes(tmp_path):
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        assert U.run_cmd(["--bench", "nosuch"]) == 1
        assert U.last_error() == "benchmark nosuch is not found"
        assert U.run_cmd(["--bench", "cube60", "-n"]) == 1
        assert "incomplete input parameter: -n" in U.last_error()
        assert U.run_cmd(["a.json", "b.json"]) == 1
        assert U.last_error() == "must use -flag or --flag to use more than 1 input"
        bad_shape = U.config_from_argv(["--bench", "cube60"])
        bad_shape["Shapes"].append({"Torus": {"R": 3}})
        assert U.run_json(bad_shape) == 1
        assert U.last_error() == "shape constructTorus is not supported"
        nophotons = U.config_from_argv(["--bench", "cube60"])
        del nophotons["Session"]["Photons"]
        assert U.run_json(nophotons) == 1
        assert "type must be number" in U.last_error()
        assert U.run_json("{not json") == 1
        assert "parse_error" in U.last_error()
        huge = U.config_from_argv(["--bench", "cube60"])
        huge["Domain"]["Dim"] = [32767, 32767, 3]       # 3.2e9 voxels: rejected before allocation
        assert U.run_json(huge) == 1
        assert U.last_error() == "Domain must hold fewer than 2^31 voxels"
    finally:
        os.chdir(cwd)


def test_dumpjson_and_help_succeed(capfd):
    assert U.run_cmd(["--bench", "cube60", "--json", '{"Optode":{"Source":{"Type":"isotropic"}}}', "--dumpjson"]) == 0
    out = capfd.readouterr().out
    assert '"ID": "cube60"' in out and "isotropic" in out
    assert U.run_cmd(["-h"]) == 0
    assert "cube60" in capfd.readouterr().out


@pytest.mark.skipif(U.device_count() > 0, reason="checks the no-GPU behaviour")
def test_simulation_without_gpu_fails_loudly():
    with pytest.raises(U.UmcxError, match="no HIP device"):
        U.Simulation(U.config_from_argv(["--bench", "cube60", "-n", "100"]))
    assert U.run_cmd(["--bench", "cube60", "-n", "100"]) == 1
    assert "no HIP device" in U.last_error()


@pytest.mark.skipif(U.device_count() > 0, reason="checks the no-GPU behaviour")
def test_batch_without_gpu_fails_loudly():
    from umcx_amd import batch as B
    with pytest.raises(U.UmcxError, match="no HIP device"):
        B.BatchSimulation(B.uneven_jobs(seed=7, njobs=2, max_photons=100))


def test_batch_job_errors_name_the_job():
    """Jobs are parsed and their domains built in parallel; a bad job's error names it (the first
    bad one when several are), before any GPU work."""
    import ctypes as C
    import json as _json
    from umcx_amd import batch as B
    from umcx_amd._native import UmcxLaunch, lib
    jobs = B.uneven_jobs(seed=7, njobs=48, max_photons=100)
    jobs[17]["Domain"]["Dim"] = [40000, 10, 10]
    del jobs[31]["Session"]["Photons"]
    texts = [_json.dumps(j).encode() for j in jobs]
    arr = (C.c_char_p * len(texts))(*texts)
    h = C.c_void_p()
    assert lib().umcx_batch_create(arr, len(texts), C.byref(UmcxLaunch()), C.byref(h)) == 1
    assert U.last_error() == "job 17: Domain dimensions must be within 1..32767 voxels per axis"
    arr[17] = None
    assert lib().umcx_batch_create(arr, len(texts), C.byref(UmcxLaunch()), C.byref(h)) == 1
    assert U.last_error() == "job 17: umcx_batch_create: null job config"


def test_batch_rejects_an_empty_job_list():
    import ctypes as C
    from umcx_amd._native import UmcxLaunch, lib
    h = C.c_void_p()
    assert lib().umcx_batch_create(None, 0, C.byref(UmcxLaunch()), C.byref(h)) != 0
    assert "at least one job" in U.last_error()


def test_null_handles_return_error_codes():
    """Every session and batch entry point turns a null handle into an error code and a message
    (no exception crosses the C ABI, nothing is dereferenced); the checks run before any GPU call."""
    import math
    from umcx_amd._native import UmcxInfo, lib
    L = lib()
    nul = C.c_void_p()
    f = (C.c_float * 4)()
    d = (C.c_double * 3)()
    i32 = (C.c_int32 * 8)()
    calls = {
        "umcx_session_info": lambda: L.umcx_session_info(nul, C.byref(UmcxInfo())),
        "umcx_session_reset": lambda: L.umcx_session_reset(nul),
        "umcx_session_run": lambda: L.umcx_session_run(nul, 0, 1, None),
        "umcx_session_fluence": lambda: L.umcx_session_fluence(nul, f, 4),
        "umcx_session_detected": lambda: L.umcx_session_detected(nul, None, None),
        "umcx_session_detp": lambda: L.umcx_session_detp(nul, None, None, None),
        "umcx_session_energies": lambda: L.umcx_session_energies(nul, d),
        "umcx_session_tally_box": lambda: L.umcx_session_tally_box(nul, i32),
        "umcx_session_set_source": lambda: L.umcx_session_set_source(nul, f, f),
        "umcx_batch_run": lambda: L.umcx_batch_run(nul, None),
        "umcx_batch_info": lambda: L.umcx_batch_info(nul, 0, C.byref(UmcxInfo())),
        "umcx_batch_energies": lambda: L.umcx_batch_energies(nul, 0, d),
        "umcx_batch_fluence": lambda: L.umcx_batch_fluence(nul, 0, f, 4),
        "umcx_batch_tally_box": lambda: L.umcx_batch_tally_box(nul, 0, i32, None),
        "umcx_batch_trials": lambda: L.umcx_batch_trials(nul, None, 0, None),
    }
    for name, call in calls.items():
        assert call() == 1, name
        assert U.last_error().startswith(name + ": null"), (name, U.last_error())
    assert math.isnan(L.umcx_session_normalizer(nul, 1.0))
    assert "null session" in U.last_error()
    assert L.umcx_seed_words(1, None) == 1
    assert L.umcx_libm_host(0, None, None, 4) == 1
    assert L.umcx_libm_host(0, None, None, 0) == 0
    L.umcx_session_free(nul)
    L.umcx_batch_free(nul)


def test_reference_structural_cli_checks():
    """The structural checks of the reference's test/testumcx.sh:13-53 against this build's
    binary: it exists and is executable, its libraries resolve, the help lists at least five long
    flags and the cube60 benchmark, and --dumpjson of that benchmark carries its ID."""
    import re
    import subprocess
    umcx = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "umcx_amd", "bin", "umcx")
    assert os.path.isfile(umcx) and os.access(umcx, os.X_OK)
    ldd = subprocess.run(["ldd", umcx], capture_output=True, text=True).stdout
    assert "not found" not in ldd
    helptext = subprocess.run([umcx], capture_output=True, text=True).stdout
    assert len(set(re.findall(r"--[a-z]+", helptext))) >= 5
    assert "cube60" in helptext
    dump = subprocess.run([umcx, "--bench", "cube60", "--dumpjson"], capture_output=True, text=True).stdout
    assert re.search(r'"ID":\s*"cube60"', dump)
    dump = subprocess.run([umcx, "--bench", "cube60", "--json",
                           '{"Optode":{"Source":{"Type":"isotropic","Pos":[29,29,29]}}}', "--dumpjson"],
                          capture_output=True, text=True).stdout
    assert "isotropic" in dump


def _box_sum(c, b):
    x0, y0, z0, g0, sx, sy, sz, sg = b
    return c[g0:g0 + sg, x0:x0 + sx, y0:y0 + sy, z0:z0 + sz].sum()


def test_tally_box_choice():
    """The pilot-driven LDS tally box: within budget, contains the densest cell, never worse than
    the densest cell alone, empty for an empty grid, and on small grids within a few percent of
    the best box of the same budget found by exhaustive search."""
    import ctypes as C
    from umcx_amd._native import check, lib

    def choose(c, budget):
        nt, nx, ny, nz = c.shape
        box = (C.c_int32 * 8)()
        arr = np.ascontiguousarray(c, dtype=np.float32)
        check(lib().umcx_choose_tally_box(arr.ctypes.data, nx, ny, nz, nt, budget, box))
        return list(box)

    assert choose(np.zeros((2, 4, 4, 4)), 16) == [0] * 8
    rng = np.random.default_rng(3)
    for trial in range(12):
        nt, nx, ny, nz = 2, 6, 5, 7
        g = np.indices((nt, nx, ny, nz)).astype(np.float64)
        centre = rng.uniform([0, 0, 0, 0], [nt, nx, ny, nz])
        d2 = sum((g[i] - centre[i]) ** 2 for i in range(4))
        c = (np.exp(-d2 / rng.uniform(1.0, 6.0)) * 1000 + rng.uniform(0, 5, d2.shape)).astype(np.float32)
        budget = int(rng.integers(8, 60))
        b = choose(c, budget)
        x0, y0, z0, g0, sx, sy, sz, sg = b
        assert 0 < sx * sy * sz * sg <= budget
        assert 0 <= g0 and g0 + sg <= nt and 0 <= x0 and x0 + sx <= nx and 0 <= y0 and y0 + sy <= ny and 0 <= z0 and z0 + sz <= nz
        am = np.unravel_index(np.argmax(c), c.shape)
        assert g0 <= am[0] < g0 + sg and x0 <= am[1] < x0 + sx and y0 <= am[2] < y0 + sy and z0 <= am[3] < z0 + sz
        got = _box_sum(c, b)
        best = 0.0
        for SG in range(1, nt + 1):
            for SX in range(1, nx + 1):
                for SY in range(1, ny + 1):
                    for SZ in range(1, nz + 1):
                        if SG * SX * SY * SZ > budget:
                            continue
                        for G0 in range(nt - SG + 1):
                            for X0 in range(nx - SX + 1):
                                for Y0 in range(ny - SY + 1):
                                    for Z0 in range(nz - SZ + 1):
                                        best = max(best, _box_sum(c, (X0, Y0, Z0, G0, SX, SY, SZ, SG)))
        assert got >= 0.9 * best, (trial, got, best)
