"""Python front end of the MI355X umcx library.

``Simulation`` wraps one GPU session (include/umcx.h): it takes the same JSON configuration as
umcx (a dict, a JSON string, or command-line arguments), runs photon ranges of the MCX_kernel
loop (src/umcx.cpp:722-781) on the GPU, and returns the fluence grid, detected photons and
energies as numpy arrays. ``run_cmd``/``run_json`` are the reference's C exports
(src/umcx.cpp:850-856), writing the same output files.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

from ._native import UmcxError, UmcxInfo, UmcxLaunch, argv_array, buffer_argv, check, lib

TRACE_DTYPE = np.dtype([("pos", "<f4", 4), ("vec", "<f4", 4), ("len", "<f4", 4), ("mediaid", "<i4"),
                        ("lastvoxelidx", "<i4"), ("w0", "<f4"), ("pad", "<u4")])
SOURCE_TYPES = ["pencil", "isotropic", "cone", "disk", "planar", "fourier"]


@dataclass
class Result:
    fluence: np.ndarray | None       # [t, x, y, z] raw tallies (or normalised, see Simulation.result)
    detp: np.ndarray                 # [saved, columns] rows as umcx saves them
    detp_id: np.ndarray              # photon index of each row
    detp_w: np.ndarray               # final weight of each row
    detected: int                    # all detections, incl. dropped beyond MaxDetPhoton
    energy_escape: float
    energy_total: float
    energy_deposit: float
    normalizer: float
    kernel_ms: float

    @property
    def absorbed_percent(self) -> float:
        return (self.energy_total - self.energy_escape) / self.energy_total * 100.0


def _cfg_text(cfg) -> str:
    if isinstance(cfg, (dict, list)):
        return json.dumps(cfg)
    return str(cfg)


class Simulation:
    """One prepared problem on the GPU."""

    def __init__(self, cfg=None, argv: list[str] | None = None, device: int = 0, block: int = 0,
                 blocks: int = 0, tally_mode: int = -1, trace: bool = False, vote: int = -1, label_cache: int = 0,
                 bjdata: bytes | None = None, deposit_cache: int = 0, deposit_log: int = 0,
                 deposit_log_segments: int = 0):
        L = lib()
        launch = UmcxLaunch(device, block, blocks, tally_mode, 1 if trace else 0, vote, label_cache, deposit_cache,
                            deposit_log, deposit_log_segments)
        h = C.c_void_p()
        if bjdata is not None:      # a BJData config buffer, as MATLAB passes it (src/umcx.cpp:475-478)
            check(L.umcx_session_from_argv(buffer_argv(bjdata), -len(bjdata), C.byref(launch), C.byref(h)))
        elif argv is not None:
            args = ["umcx"] + list(argv)
            check(L.umcx_session_from_argv(argv_array(args), len(args), C.byref(launch), C.byref(h)))
        else:
            check(L.umcx_session_from_json(_cfg_text(cfg).encode(), C.byref(launch), C.byref(h)))
        self._h = h
        info = UmcxInfo()
        check(L.umcx_session_info(h, C.byref(info)))
        self.info = {f: (list(getattr(info, f)) if f in ("dims", "seeds") else getattr(info, f))
                     for f, _ in UmcxInfo._fields_}
        self.trace_enabled = trace
        self.kernel_ms = 0.0

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().umcx_session_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- running -------------------------------------------------------------------------------
    def reset(self):
        check(lib().umcx_session_reset(self._h))
        self.kernel_ms = 0.0

    def last_launch(self) -> tuple[int, int]:
        """(lanes per workgroup, workgroups) of the last run's launch."""
        info = UmcxInfo()
        check(lib().umcx_session_info(self._h, C.byref(info)))
        return info.last_block, info.last_blocks

    def lds_label_bits(self) -> int:
        """Bits per voxel of the labels staged in LDS now (0: read from HBM); the pilot may switch a
        volume of at most two media ids to its 1-bit form."""
        info = UmcxInfo()
        check(lib().umcx_session_info(self._h, C.byref(info)))
        return int(info.lds_label_bits)

    def deposit_cache(self) -> bool:
        """Whether runs use the per-lane cache of global fluence deposits."""
        info = UmcxInfo()
        check(lib().umcx_session_info(self._h, C.byref(info)))
        return bool(info.deposit_cache)

    def deposit_log(self) -> bool:
        """Whether runs append global fluence deposits to the deposit log (summed into the grid by
        separate passes) instead of issuing float atomics."""
        info = UmcxInfo()
        check(lib().umcx_session_info(self._h, C.byref(info)))
        return bool(info.deposit_log)

    def global_deposits_per_photon(self) -> float:
        """The pilot's estimate of the global fluence deposits per photon of the runs (0 before a pilot)."""
        info = UmcxInfo()
        check(lib().umcx_session_info(self._h, C.byref(info)))
        return float(info.global_deposits_per_photon)

    def vote(self) -> int:
        """Lanes of a wave that must wait to scatter before a scatter round runs (0: none), as the
        runs use it now (the pilot picks it unless the session was given one)."""
        info = UmcxInfo()
        check(lib().umcx_session_info(self._h, C.byref(info)))
        return int(info.vote)

    def deposit_cache_hits(self) -> float:
        """Share of the pilot's global deposits that hit the per-lane deposit cache (0 before a pilot)."""
        info = UmcxInfo()
        check(lib().umcx_session_info(self._h, C.byref(info)))
        return float(info.deposit_cache_hits)

    def run(self, first: int = 0, count: int | None = None) -> float:
        """Run photons [first, first+count); returns the kernel time in ms."""
        n = self.info["nphoton"] if count is None else int(count)
        ms = C.c_double(0.0)
        check(lib().umcx_session_run(self._h, int(first), n, C.byref(ms)))
        self.kernel_ms += ms.value
        return ms.value

    # -- outputs -------------------------------------------------------------------------------
    @property
    def dims(self):
        return tuple(self.info["dims"])

    def fluence(self) -> np.ndarray | None:
        """Raw tallies with shape [t, x, y, z] (umcx memory order)."""
        if not self.info["issavevol"]:
            return None
        nx, ny, nz, nt = self.dims
        out = np.empty(nt * nx * ny * nz, dtype=np.float32)
        check(lib().umcx_session_fluence(self._h, out.ctypes.data, out.size))
        return out.reshape(nt, nx, ny, nz)

    def tally_box(self):
        """The LDS tally tile chosen for this session, {x0,y0,z0,g0,sx,sy,sz,sg}, or None."""
        b = (C.c_int32 * 8)()
        check(lib().umcx_session_tally_box(self._h, b))
        return list(b) if b[4] > 0 else None

    def fluence_device_ptr(self) -> int:
        p = C.c_void_p()
        check(lib().umcx_session_fluence_device(self._h, C.byref(p)))
        return p.value or 0

    def detected(self) -> tuple[int, int]:
        d, s = C.c_uint32(), C.c_uint32()
        check(lib().umcx_session_detected(self._h, C.byref(d), C.byref(s)))
        return d.value, s.value

    def detp(self):
        _, saved = self.detected()
        cols = self.info["detp_columns"]
        rows = np.empty((saved, cols), dtype=np.float32)
        ids = np.empty(saved, dtype=np.uint32)
        w = np.empty(saved, dtype=np.float32)
        check(lib().umcx_session_detp(self._h, rows.ctypes.data if rows.size else None,
                                      ids.ctypes.data if saved else None, w.ctypes.data if saved else None))
        return rows, ids, w

    def energies(self) -> tuple[float, float, float]:
        e = (C.c_double * 3)()
        check(lib().umcx_session_energies(self._h, e))
        return e[0], e[1], e[2]

    def normalizer(self, energytotal: float) -> float:
        return float(lib().umcx_session_normalizer(self._h, float(energytotal)))

    def trace(self, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=TRACE_DTYPE)
        check(lib().umcx_session_trace(self._h, out.ctypes.data, n))
        return out

    def result(self, normalize: bool | None = None) -> Result:
        esc, tot, dep = self.energies()
        norm = self.normalizer(tot) if tot > 0 else 1.0
        flu = self.fluence()
        if normalize is None:
            normalize = bool(self.info["isnormalized"])
        if flu is not None and normalize and norm != 1.0:
            flu = flu * np.float32(norm)
        rows, ids, w = self.detp()
        return Result(flu, rows, ids, w, self.detected()[0], esc, tot, dep, norm, self.kernel_ms)


def simulate(cfg, nphoton: int | None = None, **launch) -> Result:
    """Run a whole simulation on the GPU and return its outputs (raw tallies)."""
    with Simulation(cfg, **launch) as sim:
        sim.run(0, nphoton)
        return sim.result(normalize=False)


def run_cmd(args: list[str]) -> int:
    """The reference's MCX_run_cmd: run like the umcx command line; writes output files."""
    argv = ["umcx"] + list(args)
    return lib().MCX_run_cmd(argv_array(argv), len(argv))


def run_json(cfg) -> int:
    """The reference's MCX_run_json: one JSON document as the whole configuration."""
    return lib().MCX_run_json(_cfg_text(cfg).encode())


def run_bjdata(buf: bytes) -> int:
    """The reference's MCX_run_simulation with argn < 0: argv[0] is a BJData config buffer (the
    MATLAB input, src/umcx.cpp:475-478); runs it and writes the output files."""
    argv = buffer_argv(buf)
    return lib()._Z18MCX_run_simulationPPciiPPv(argv, -len(buf), 0, None)


def config_from_bjdata(buf: bytes) -> dict:
    """The configuration a BJData input buffer holds, as umcx reads it (no GPU needed)."""
    L = lib()
    out = C.c_void_p()
    check(L.umcx_config_from_argv(buffer_argv(buf), -len(buf), C.byref(out)))
    text = C.cast(out, C.c_char_p).value.decode("utf-8", "surrogateescape")   # strings are raw bytes
    L.umcx_free(out)
    return json.loads(text)


def last_error() -> str:
    return lib().umcx_last_error().decode("utf-8", "replace")


def domain(cfg) -> np.ndarray:
    """Label volume [x, y, z] int32 with detector bits, as umcx builds it (no GPU needed)."""
    L = lib()
    dims = (C.c_uint32 * 3)()
    ptr = C.c_void_p()
    check(L.umcx_domain_from_json(_cfg_text(cfg).encode(), dims, C.byref(ptr)))
    n = dims[0] * dims[1] * dims[2]
    arr = np.ctypeslib.as_array((C.c_int32 * n).from_address(ptr.value)).copy()
    L.umcx_free(ptr)
    return arr.reshape(dims[0], dims[1], dims[2])


def config_from_argv(args: list[str]) -> dict:
    """The expanded configuration umcx would run for these command-line arguments."""
    L = lib()
    argv = ["umcx"] + list(args)
    out = C.c_void_p()
    check(L.umcx_config_from_argv(argv_array(argv), len(argv), C.byref(out)))
    text = C.cast(out, C.c_char_p).value.decode()
    L.umcx_free(out)
    return json.loads(text)


def seed_words(seed: int) -> list[int]:
    out = (C.c_uint32 * 4)()
    check(lib().umcx_seed_words(int(seed), out))
    return list(out)


def libm_eval(fn: int, x: np.ndarray, device: int | None = None) -> np.ndarray:
    """Evaluate a libm replica (0 logf, 1 sinf, 2 cosf, 3 acosf, 4 expf) on the CPU or the GPU."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    if device is None:
        check(lib().umcx_libm_host(fn, x.ctypes.data, out.ctypes.data, x.size))
    else:
        check(lib().umcx_libm_device(fn, x.ctypes.data, out.ctypes.data, x.size, device))
    return out


def device_count() -> int:
    return int(lib().umcx_device_count())


__all__ = ["Simulation", "Result", "simulate", "run_cmd", "run_json", "domain", "config_from_argv",
           "seed_words", "libm_eval", "device_count", "last_error", "UmcxError", "TRACE_DTYPE"]
