"""ctypes access to the CPU oracle (oracle/umcx_oracle.c) for the parity tests.

Test infrastructure only. ``problem_from_config`` restates how MCX_run_simulation derives its
constants from the JSON config (src/umcx.cpp:785-810) in float32 arithmetic, so the oracle can
be driven from the same configs and label volumes as the reference.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "libumcx_oracle.so")
F32 = np.float32

SRC_TYPES = ["pencil", "isotropic", "cone", "disk", "planar", "fourier"]   # src/umcx.cpp:61
OUT_TYPES = "xfe"                                                          # src/umcx.cpp:59


class OracleProblem(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("maxgate", C.c_int32),
        ("tstart", C.c_float), ("tend", C.c_float), ("rtstep", C.c_float), ("unitinmm", C.c_float),
        ("isreflect", C.c_int32), ("issavevol", C.c_int32), ("issavedet", C.c_int32), ("savedetflag", C.c_int32),
        ("mediumnum", C.c_int32), ("outputtype", C.c_int32), ("detnum", C.c_int32), ("srctype", C.c_int32),
        ("maxdetphotons", C.c_uint32),
        ("srcpos", C.c_float * 4), ("srcdir", C.c_float * 4), ("srcparam1", C.c_float * 4), ("srcparam2", C.c_float * 4),
        ("seeds", C.c_uint32 * 4),
        ("labels", C.POINTER(C.c_int32)), ("media", C.POINTER(C.c_float)), ("detpos", C.POINTER(C.c_float)),
    ]


class OracleTrace(C.Structure):
    _fields_ = [("pos", C.c_float * 4), ("vec", C.c_float * 4), ("len", C.c_float * 4),
                ("mediaid", C.c_int32), ("lastvoxelidx", C.c_int32), ("w0", C.c_float),
                ("nstep", C.c_uint32), ("ncross", C.c_uint32), ("nscat", C.c_uint32),
                ("nrand", C.c_uint32), ("nfresnel", C.c_uint32)]


TRACE_DTYPE = np.dtype([("pos", "<f4", 4), ("vec", "<f4", 4), ("len", "<f4", 4), ("mediaid", "<i4"),
                        ("lastvoxelidx", "<i4"), ("w0", "<f4"), ("nstep", "<u4"), ("ncross", "<u4"),
                        ("nscat", "<u4"), ("nrand", "<u4"), ("nfresnel", "<u4")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
        _lib = C.CDLL(LIB)
        _lib.oracle_run.argtypes = [C.POINTER(OracleProblem), C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double),
                                    C.c_void_p, C.c_int]
        _lib.oracle_run.restype = C.c_int
        _lib.oracle_seed_words.argtypes = [C.c_int32, C.POINTER(C.c_uint32)]
        _lib.oracle_detp_len.argtypes = [C.POINTER(OracleProblem)]
        _lib.oracle_host_libm_array.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint64]
        _lib.oracle_set_mutant.argtypes = [C.c_int]
        _lib.oracle_replay.argtypes = [C.POINTER(OracleProblem), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int]
        _lib.oracle_replay.restype = C.c_int
        for fn in ("logf", "sinf", "cosf", "acosf", "expf"):
            f = getattr(_lib, "oracle_host_" + fn)
            f.argtypes = [C.c_float]
            f.restype = C.c_float
    return _lib


def host_libm(fn: int, x: np.ndarray) -> np.ndarray:
    """The host C library's logf/sinf/cosf/acosf/expf (fn 0..4) over an array."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib().oracle_host_libm_array(fn, x.ctypes.data, out.ctypes.data, x.size)
    return out


def reachable_inputs(g_values=(0.9, 0.01, 0.89, 0.95, 0.7)) -> dict:
    """Every argument the photon loop can hand to a trajectory-affecting libm call: u on the
    2^23 grid of rand01 (src/umcx.cpp:209-210); logf(u+eps) (:213); sin/cos(2 pi u) (:366-367);
    acosf of the Henyey-Greenstein cosine per medium g (:370-372) and of 2u-1 (:374)."""
    F = np.float32
    k = np.arange(1 << 23, dtype=np.uint32)
    u = (k | np.uint32(0x3F800000)).view(np.float32) - F(1)
    out = {0: [u + F(np.finfo(np.float32).eps)], 1: [], 2: [], 3: []}
    phi = F(6.2831855) * u
    out[1].append(phi)
    out[2].append(phi)
    out[3].append(F(2) * u - F(1))
    for g in g_values:
        g = F(g)
        t = (F(1) - g * g) / (F(1) - g + F(2) * g * u)
        t = (F(1) + g * g - t * t) / (F(2) * g)
        out[3].append(np.maximum(F(-1), np.minimum(F(1), t)))
    # expf(-mua*d): every 61st float of [-104, 0] plus the two hardest rounding cases found by an
    # exhaustive scan of all floats in [-104, 88]
    neg = np.arange(0x80000000, 0xC2D00001, 61, dtype=np.uint64).astype(np.uint32).view(np.float32)
    hard = np.array([float.fromhex("-0x1.f8cbb2p+5"), float.fromhex("0x1.04845ep+5")], dtype=np.float32)
    out[4] = [neg, hard]
    return {fn: np.concatenate(v) for fn, v in out.items()}


def seed_words(seed: int) -> list[int]:
    out = (C.c_uint32 * 4)()
    lib().oracle_seed_words(int(seed), out)
    return list(out)


def _f(v) -> np.float32:
    return F32(float(v))


def problem_from_config(cfg: dict, labels: np.ndarray) -> dict:
    """Derive MCX_param and the kernel inputs from an (expanded) umcx JSON config.

    Follows src/umcx.cpp:785-810 (constants, media scaling, detectors) and :725-729 (seed, source).
    ``labels`` is the [x,y,z] int32 label volume with detector bits already set.
    """
    S = cfg.get("Session", {})
    Fw = cfg["Forward"]
    D = cfg["Domain"]
    O = cfg["Optode"]
    T0, T1, Dt = _f(Fw["T0"]), _f(Fw["T1"]), _f(Fw["Dt"])
    unit = _f(D.get("LengthUnit", 1.0))
    media = []
    for m in D["Media"]:
        if isinstance(m, list):
            media.append([_f(v) for v in m])
        else:
            media.append([_f(m["mua"]) * unit, _f(m["mus"]) * unit, _f(m["g"]), _f(m["n"])])
    src = O["Source"]
    p1 = src.get("Param1", [0, 0, 0, 0])
    p2 = src.get("Param2", [0, 0, 0, 0])
    dets = O.get("Detector", [])
    seed = int(S.get("RNGSeed", 1648335518))
    return dict(
        dims=tuple(int(v) for v in labels.shape),
        tstart=T0, tend=T1, rtstep=F32(1.0) / Dt, unitinmm=unit,
        maxgate=int((T1 - T0) / Dt + F32(0.5)),
        isreflect=int(S.get("DoMismatch", 1)), isnormalized=int(S.get("DoNormalize", 1)),
        issavevol=int(S.get("DoSaveVolume", 1)), issavedet=int(S.get("DoPartialPath", 1)),
        savedetflag=int(S.get("SaveDetFlag", 5)), outputtype=OUT_TYPES.find(S.get("OutputType", "x")[0]),
        maxdetphotons=int(S.get("MaxDetPhoton", 1000000)),
        srctype=SRC_TYPES.index(src.get("Type", "pencil")) if src.get("Type", "pencil") in SRC_TYPES else 6,
        srcpos=[_f(v) for v in src["Pos"][:3]] + [F32(1)], srcdir=[_f(v) for v in src["Dir"][:3]] + [F32(0)],
        srcparam1=[_f(v) for v in p1], srcparam2=[_f(v) for v in p2],
        seeds=seed_words(seed), photons=int(float(S["Photons"])),
        media=np.asarray(media, dtype=np.float32).reshape(-1, 4),
        detpos=np.asarray([[_f(d["Pos"][0]), _f(d["Pos"][1]), _f(d["Pos"][2]), _f(d["R"])] for d in dets],
                          dtype=np.float32).reshape(-1, 4),
        labels=np.ascontiguousarray(labels, dtype=np.int32),
    )


def _oracle_problem(prob: dict) -> OracleProblem:
    nx, ny, nz = prob["dims"]
    op = OracleProblem()
    op.nx, op.ny, op.nz = nx, ny, nz
    op.maxgate = prob["maxgate"]
    op.tstart, op.tend, op.rtstep, op.unitinmm = prob["tstart"], prob["tend"], prob["rtstep"], prob["unitinmm"]
    op.isreflect, op.issavevol, op.issavedet, op.savedetflag = (prob["isreflect"], prob["issavevol"],
                                                                 prob["issavedet"], prob["savedetflag"])
    media, detpos, labels = prob["media"], prob["detpos"], prob["labels"]
    op.mediumnum, op.outputtype, op.detnum, op.srctype = len(media), prob["outputtype"], len(detpos), prob["srctype"]
    op.maxdetphotons = prob["maxdetphotons"]
    for name in ("srcpos", "srcdir", "srcparam1", "srcparam2"):
        arr = getattr(op, name)
        for i, v in enumerate(prob[name][:4]):
            arr[i] = float(v)
    for i, v in enumerate(prob["seeds"]):
        op.seeds[i] = v
    op.labels = labels.ctypes.data_as(C.POINTER(C.c_int32))
    op.media = media.ctypes.data_as(C.POINTER(C.c_float))
    op.detpos = detpos.ctypes.data_as(C.POINTER(C.c_float)) if len(detpos) else C.POINTER(C.c_float)()
    return op


def replay(prob: dict, ids, slices, coef, nslices: int, sources=None, src_index=None, vprops=None,
           nthreads: int = 0) -> dict:
    """Per-voxel perturbation-MC replay of photons ``ids`` (oracle_replay): returns gmua, gmus
    (nslices x nvox float64: sum of coef x (-L_v) and coef x (n_v/mus_v - L_v)), amua (sums of the
    mua terms' absolute values), amus (sums of |coef| (n_v/mus_v + L_v), the scale of the mus
    terms' rounding), final_w and mpath (per-medium float32 segment sums in voxels,
    one row per record). ``sources`` is a (S, 6) {pos, dir} table indexed by ``src_index``;
    ``vprops`` per-voxel (nvox, 2) {mua, mus} in per-voxel-length units."""
    op = _oracle_problem(prob)
    nvox = int(np.prod(prob["dims"]))
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    n = len(ids)
    slices = np.ascontiguousarray(slices, dtype=np.int32)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    gmua = np.zeros((nslices, nvox)); gmus = np.zeros((nslices, nvox))
    amua = np.zeros((nslices, nvox)); amus = np.zeros((nslices, nvox))
    final_w = np.zeros(max(n, 1), dtype=np.float32)
    mpath = np.zeros((max(n, 1), len(prob["media"])), dtype=np.float32)
    srcs = np.ascontiguousarray(sources, dtype=np.float32).reshape(-1) if sources is not None else None
    sidx = np.ascontiguousarray(src_index, dtype=np.int32) if src_index is not None else None
    vp = np.ascontiguousarray(vprops, dtype=np.float32).reshape(-1) if vprops is not None else None
    rc = lib().oracle_replay(C.byref(op), ids.ctypes.data, sidx.ctypes.data if sidx is not None else None,
                             srcs.ctypes.data if srcs is not None else None, slices.ctypes.data, coef.ctypes.data, n,
                             vp.ctypes.data if vp is not None else None, gmua.ctypes.data, gmus.ctypes.data,
                             amua.ctypes.data, amus.ctypes.data, nvox, final_w.ctypes.data, mpath.ctypes.data,
                             nthreads)
    if rc != 0:
        raise RuntimeError("oracle_replay rejected its input")
    return dict(gmua=gmua, gmus=gmus, amua=amua, amus=amus, final_w=final_w[:n], mpath=mpath[:n])


def run(prob: dict, nphoton: int | None = None, first: int = 0, traces: bool = False, nthreads: int = 0):
    """Run the oracle; returns dict(fluence [t][x][y][z] raw sums, detp, detp_id, detp_w, detected,
    energies=(escape, total, deposit), traces)."""
    L = lib()
    n = prob["photons"] if nphoton is None else int(nphoton)
    nx, ny, nz = prob["dims"]
    op = OracleProblem()
    op.nx, op.ny, op.nz = nx, ny, nz
    op.maxgate = prob["maxgate"]
    op.tstart, op.tend, op.rtstep, op.unitinmm = prob["tstart"], prob["tend"], prob["rtstep"], prob["unitinmm"]
    op.isreflect, op.issavevol, op.issavedet, op.savedetflag = (prob["isreflect"], prob["issavevol"],
                                                                 prob["issavedet"], prob["savedetflag"])
    media, detpos, labels = prob["media"], prob["detpos"], prob["labels"]
    op.mediumnum, op.outputtype, op.detnum, op.srctype = len(media), prob["outputtype"], len(detpos), prob["srctype"]
    op.maxdetphotons = prob["maxdetphotons"]
    for name in ("srcpos", "srcdir", "srcparam1", "srcparam2"):
        arr = getattr(op, name)
        for i, v in enumerate(prob[name][:4]):
            arr[i] = float(v)
    for i, v in enumerate(prob["seeds"]):
        op.seeds[i] = v
    op.labels = labels.ctypes.data_as(C.POINTER(C.c_int32))
    op.media = media.ctypes.data_as(C.POINTER(C.c_float))
    op.detpos = detpos.ctypes.data_as(C.POINTER(C.c_float)) if len(detpos) else C.POINTER(C.c_float)()
    flu = np.zeros(prob["maxgate"] * nx * ny * nz, dtype=np.float32) if prob["issavevol"] else None
    detlen = L.oracle_detp_len(C.byref(op))
    maxdet = prob["maxdetphotons"] if prob["issavedet"] else 0
    # a photon is saved once per detector whose disk holds its exit point (detectors may overlap)
    cap = min(maxdet, max(n, 1) * max(len(detpos), 1))
    detp = np.zeros((cap, max(detlen, 1)), dtype=np.float32)
    detid = np.zeros(cap, dtype=np.uint32)
    detw = np.zeros(cap, dtype=np.float32)
    ndet = C.c_uint32(0)
    en = (C.c_double * 3)()
    tr = np.zeros(n, dtype=TRACE_DTYPE) if traces else None
    # the oracle writes at most maxdetphotons rows; size the buffer by photons x detectors instead
    saved_max = op.maxdetphotons
    op.maxdetphotons = cap
    rc = L.oracle_run(C.byref(op), first, n, flu.ctypes.data if flu is not None else None,
                      detp.ctypes.data, detid.ctypes.data, detw.ctypes.data, C.byref(ndet), en,
                      tr.ctypes.data if tr is not None else None, nthreads)
    op.maxdetphotons = saved_max
    if rc != 0:
        raise RuntimeError("oracle_run rejected the problem")
    saved = min(ndet.value, cap)
    return dict(fluence=flu, detp=detp[:saved, :detlen], detp_id=detid[:saved], detp_w=detw[:saved],
                detected=ndet.value, energies=(en[0], en[1], en[2]), traces=tr)


def normalizer(prob: dict, energytotal: float) -> np.float32:
    """src/umcx.cpp:820 -- double arithmetic, stored as float."""
    ot, rt, u = prob["outputtype"], float(prob["rtstep"]), float(prob["unitinmm"])
    if ot == 2:
        return F32(1.0 / energytotal)
    if ot == 0:
        return F32(rt / (energytotal * u * u))
    return F32(1.0 / (energytotal * u * u))


def golden(name: str) -> dict:
    d = np.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz"), allow_pickle=False)
    out = {k: d[k] for k in d.files}
    out["cfg"] = json.loads(str(out["dumpjson"]))
    out["argv"] = json.loads(str(out["argv"]))
    return out


def golden_problem(name: str) -> tuple[dict, dict]:
    g = golden(name)
    prob = problem_from_config(g["cfg"], g["mask"])
    return g, prob
