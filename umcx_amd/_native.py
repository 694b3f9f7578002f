"""ctypes binding of libumcx.so (C ABI declared in include/umcx.h).

The library is built in-tree (``umcx_amd/lib/libumcx.so``, see ``umcx_amd/build.py``). There is
no fallback: when it is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libumcx.so")


class UmcxLaunch(C.Structure):
    _fields_ = [("device", C.c_int32), ("block", C.c_int32), ("blocks", C.c_int32),
                ("tally_mode", C.c_int32), ("trace", C.c_int32), ("vote", C.c_int32), ("label_cache", C.c_int32),
                ("deposit_cache", C.c_int32), ("deposit_log", C.c_int32),
                ("deposit_log_segments", C.c_int32)]


class UmcxInfo(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 4), ("nphoton", C.c_uint64),
                ("mediumnum", C.c_int32), ("detnum", C.c_int32), ("detp_columns", C.c_int32),
                ("ppath_columns", C.c_int32), ("outputtype", C.c_int32), ("isnormalized", C.c_int32),
                ("issavevol", C.c_int32), ("issavedet", C.c_int32), ("isreflect", C.c_int32),
                ("srctype", C.c_int32), ("tstart", C.c_float), ("tend", C.c_float), ("rtstep", C.c_float),
                ("unitinmm", C.c_float), ("seeds", C.c_uint32 * 4), ("block", C.c_int32), ("blocks", C.c_int32),
                ("tally_mode", C.c_int32), ("label_bytes", C.c_int32), ("lds_label_bits", C.c_int32),
                ("last_block", C.c_int32), ("last_blocks", C.c_int32), ("deposit_cache", C.c_int32),
                ("deposit_cache_hits", C.c_double), ("deposit_log", C.c_int32),
                ("global_deposits_per_photon", C.c_double), ("vote", C.c_int32)]


class UmcxBatchTrial(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks", C.c_int32), ("tile_cells", C.c_int32), ("chosen", C.c_int32),
                ("ms", C.c_double), ("tile_share", C.c_double)]


class UmcxDeviceProblem(C.Structure):
    """umcx_device_problem of include/umcx.h (every pointer a device pointer)."""
    _fields_ = [("dims", C.c_uint32 * 4), ("mediumnum", C.c_int32), ("detnum", C.c_int32),
                ("isreflect", C.c_int32), ("issavedet", C.c_int32), ("savedetflag", C.c_int32),
                ("outputtype", C.c_int32), ("srctype", C.c_int32),
                ("tstart", C.c_float), ("tend", C.c_float), ("rtstep", C.c_float), ("unitinmm", C.c_float),
                ("srcpos", C.c_float * 4), ("srcdir", C.c_float * 4), ("srcparam1", C.c_float * 4),
                ("srcparam2", C.c_float * 4), ("seeds", C.c_uint32 * 4), ("maxdetphotons", C.c_uint32),
                ("deposit_cache", C.c_int32),
                ("labels", C.c_void_p), ("media", C.c_void_p), ("detpos", C.c_void_p), ("fluence", C.c_void_p),
                ("detp", C.c_void_p), ("detid", C.c_void_p), ("detw", C.c_void_p), ("ndet", C.c_void_p),
                ("energies", C.c_void_p)]


# every exported symbol and its signature: (restype, argtypes)
SIGNATURES = {
    "MCX_run_cmd": (C.c_int, [C.POINTER(C.c_char_p), C.c_int]),
    "MCX_run_json": (C.c_int, [C.c_char_p]),
    # C++ linkage entry of src/umcx.cpp:783, MCX_run_simulation(char**, int, int, void**)
    "_Z18MCX_run_simulationPPciiPPv": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_void_p]),
    "umcx_last_error": (C.c_char_p, []),
    "umcx_version": (C.c_char_p, []),
    "umcx_device_count": (C.c_int, []),
    "umcx_session_from_json": (C.c_int, [C.c_char_p, C.POINTER(UmcxLaunch), C.POINTER(C.c_void_p)]),
    "umcx_session_from_argv": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.POINTER(UmcxLaunch), C.POINTER(C.c_void_p)]),
    "umcx_session_info": (C.c_int, [C.c_void_p, C.POINTER(UmcxInfo)]),
    "umcx_session_reset": (C.c_int, [C.c_void_p]),
    "umcx_session_run": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]),
    "umcx_session_fluence": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "umcx_session_fluence_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "umcx_session_detected": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "umcx_session_detp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "umcx_session_energies": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "umcx_session_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "umcx_session_normalizer": (C.c_float, [C.c_void_p, C.c_double]),
    "umcx_session_tally_box": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "umcx_session_set_voxel_props": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    "umcx_session_set_source": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "umcx_session_detp_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_void_p)]),
    "umcx_session_replay": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                      C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_double)]),
    "umcx_session_set_sources": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64]),
    "umcx_session_replay_sources": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint64,
                                              C.c_void_p, C.POINTER(C.c_double)]),
    "umcx_session_free": (None, [C.c_void_p]),
    "umcx_batch_create": (C.c_int, [C.POINTER(C.c_char_p), C.c_int32, C.POINTER(UmcxLaunch), C.POINTER(C.c_void_p)]),
    "umcx_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "umcx_batch_info": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(UmcxInfo)]),
    "umcx_batch_fluence": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64]),
    "umcx_batch_energies": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "umcx_batch_detected": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "umcx_batch_detp": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "umcx_batch_trials": (C.c_int, [C.c_void_p, C.POINTER(UmcxBatchTrial), C.c_int32, C.POINTER(C.c_int32)]),
    "umcx_batch_tally_box": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "umcx_batch_free": (None, [C.c_void_p]),
    "umcx_run_device_noalloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]),
    "umcx_domain_from_json": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_void_p)]),
    "umcx_config_from_argv": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.POINTER(C.c_void_p)]),
    "umcx_choose_tally_box": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                        C.POINTER(C.c_int32)]),
    "umcx_choose_tally_box_device": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                               C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int32]),
    "umcx_seed_words": (C.c_int, [C.c_int32, C.POINTER(C.c_uint32)]),
    "umcx_libm_host": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64]),
    "umcx_libm_device": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32]),
    "umcx_arith_check_device": (C.c_int, [C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int32]),
    "umcx_free": (None, [C.c_void_p]),
}

_lib = None


class UmcxError(RuntimeError):
    pass


def _bind_torch_runtime() -> None:
    """Load PyTorch's HIP runtime before libumcx.so when PyTorch is installed.

    libumcx.so and PyTorch's ROCm build both need libamdhip64.so.7, and a process holds one
    library per soname: whichever loads first serves both. PyTorch only works on the runtime
    it ships with, so it has to come first; libumcx.so runs on either.
    """
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib() -> C.CDLL:
    """Load libumcx.so once; raises when the native library is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UmcxError(f"native library missing: {LIB_PATH} (build it with `python -m umcx_amd.build`)")
        _bind_torch_runtime()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise UmcxError(lib().umcx_last_error().decode("utf-8", "replace"))


def buffer_argv(buf: bytes):
    """argv for argn < 0: argv[0] points at a BJData buffer (the array keeps `buf` alive)."""
    if not isinstance(buf, (bytes, bytearray)) or len(buf) == 0:
        raise UmcxError("a BJData input needs a non-empty bytes buffer")
    arr = (C.c_char_p * 2)()
    arr[0] = bytes(buf)
    arr[1] = None
    return arr


def argv_array(args: list[str]):
    arr = (C.c_char_p * (len(args) + 1))()
    for i, a in enumerate(args):
        arr[i] = a.encode()
    arr[len(args)] = None
    return arr
