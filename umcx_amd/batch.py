"""Batched uneven simulations: many independent umcx jobs in one persistent GPU launch.

``BatchSimulation`` wraps ``umcx_batch_*`` (include/umcx.h). ``uneven_jobs`` generates the
batched workload: 256 jobs from seed 7 with volume edges uniform in [32, 128] per axis, 2-6
labels as random axis-aligned boxes, per-label mua ~ U(0.001, 0.05)/mm, mus ~ U(0.5, 20)/mm,
g ~ U(0.7, 0.95), n ~ U(1.33, 1.45), Fresnel reflection on, a pencil source at the centre of the
z = 0 face, photon counts log-uniform in [1e5, 1e7] and one 0-5 ns gate. Volumes travel as
JData ``_ArrayZipData_`` (zlib + base64), which umcx reads (src/umcx.cpp:578-597).
"""
from __future__ import annotations

import base64
import ctypes as C
import json
import time
import zlib

import numpy as np

from ._native import UmcxBatchTrial, UmcxInfo, UmcxLaunch, check, lib
from .api import Simulation


def uneven_jobs(seed: int = 7, njobs: int = 256, max_photons: int | None = None) -> list[dict]:
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        dims = [int(v) for v in rng.integers(32, 129, size=3)]
        nlab = int(rng.integers(2, 7))
        media = [{"mua": 0.0, "mus": 0.0, "g": 1.0, "n": 1.0}]
        for _ in range(nlab):
            media.append({"mua": float(rng.uniform(0.001, 0.05)), "mus": float(rng.uniform(0.5, 20.0)),
                          "g": float(rng.uniform(0.7, 0.95)), "n": float(rng.uniform(1.33, 1.45))})
        vol = np.ones(dims, dtype=np.uint8)
        for lab in range(2, nlab + 1):
            lo = [int(rng.integers(0, d - 4)) for d in dims]
            hi = [int(rng.integers(l + 4, d + 1)) for l, d in zip(lo, dims)]
            vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = lab
        photons = int(round(10 ** rng.uniform(5.0, 7.0)))
        if max_photons:
            photons = min(photons, max_photons)
        jseed = int(rng.integers(1, 2 ** 31 - 1))
        jobs.append({
            "Session": {"ID": f"job{j:03d}", "Photons": photons, "RNGSeed": jseed, "DoMismatch": 1,
                        "DoSaveVolume": 1, "DoPartialPath": 0, "OutputType": "x"},
            "Forward": {"T0": 0.0, "T1": 5e-9, "Dt": 5e-9},
            "Domain": {"Dim": dims, "LengthUnit": 1.0, "Media": media},
            "Optode": {"Source": {"Type": "pencil", "Pos": [dims[0] / 2, dims[1] / 2, 0.0], "Dir": [0.0, 0.0, 1.0]}},
            "Shapes": {"_ArraySize_": dims, "_ArrayType_": "uint8", "_ArrayZipType_": "zlib",
                       "_ArrayZipData_": base64.b64encode(zlib.compress(vol.reshape(-1).tobytes(), 1)).decode("ascii")},
        })
    return jobs


def roi_sum(fluence: np.ndarray, cfg: dict, half: int = 2) -> float:
    """Fluence summed over the (2*half+1)^3 voxels under the source (the per-job reported
    quantity of the batched comparison)."""
    x, y, _ = (int(v) for v in cfg["Optode"]["Source"]["Pos"])
    return float(fluence[..., max(x - half, 0):x + half + 1, max(y - half, 0):y + half + 1, 0:2 * half + 1]
                 .astype(np.float64).sum())


class BatchSimulation:
    """All jobs prepared on the GPU; ``run()`` replays the captured graph."""

    def __init__(self, cfgs: list, device: int = 0, block: int = 0, blocks: int = 0, tally_mode: int = -1):
        texts = [c if isinstance(c, str) else json.dumps(c) for c in cfgs]
        arr = (C.c_char_p * len(texts))(*[t.encode() for t in texts])
        launch = UmcxLaunch(device, block, blocks, tally_mode, 0, 0)
        h = C.c_void_p()
        check(lib().umcx_batch_create(arr, len(texts), C.byref(launch), C.byref(h)))
        self._h = h
        self.njobs = len(texts)
        self.infos = []
        for j in range(self.njobs):
            info = UmcxInfo()
            check(lib().umcx_batch_info(h, j, C.byref(info)))
            self.infos.append({f: (list(getattr(info, f)) if f in ("dims", "seeds") else getattr(info, f))
                               for f, _ in UmcxInfo._fields_})

    def trials(self) -> list[dict]:
        """The autotuner's timed launch shapes (see umcx_batch_trials)."""
        n = C.c_int32(0)
        check(lib().umcx_batch_trials(self._h, None, 0, C.byref(n)))
        arr = (UmcxBatchTrial * max(n.value, 1))()
        check(lib().umcx_batch_trials(self._h, arr, n.value, C.byref(n)))
        return [{f: getattr(arr[i], f) for f, _ in UmcxBatchTrial._fields_} for i in range(n.value)]

    def tally_box(self, j: int):
        """({x0, y0, z0, gate0, nx, ny, nz, ngates}, pilot deposits per photon) of job j."""
        box = (C.c_int32 * 8)()
        dep = C.c_double(0.0)
        check(lib().umcx_batch_tally_box(self._h, j, box, C.byref(dep)))
        return list(box), dep.value

    def run(self) -> float:
        ms = C.c_double(0.0)
        check(lib().umcx_batch_run(self._h, C.byref(ms)))
        return ms.value

    def fluence(self, j: int) -> np.ndarray:
        """Raw tallies of job j, shape [t, x, y, z]; empty when the job saves no volume."""
        if not self.infos[j]["issavevol"]:
            return np.empty(0, dtype=np.float32)
        nx, ny, nz, nt = self.infos[j]["dims"]
        out = np.empty(nt * nx * ny * nz, dtype=np.float32)
        check(lib().umcx_batch_fluence(self._h, j, out.ctypes.data, out.size))
        return out.reshape(nt, nx, ny, nz)

    def detected(self, j: int) -> tuple[int, int]:
        """(all detections, rows kept) of job j."""
        d, sv = C.c_uint32(), C.c_uint32()
        check(lib().umcx_batch_detected(self._h, j, C.byref(d), C.byref(sv)))
        return d.value, sv.value

    def detp(self, j: int):
        """Detected-photon rows of job j [saved, detp_columns], their photon indices and final
        weights, as Simulation.detp gives them for a separate run."""
        _, saved = self.detected(j)
        cols = self.infos[j]["detp_columns"]
        rows = np.empty((saved, cols), dtype=np.float32)
        ids = np.empty(saved, dtype=np.uint32)
        w = np.empty(saved, dtype=np.float32)
        check(lib().umcx_batch_detp(self._h, j, rows.ctypes.data if rows.size else None,
                                    ids.ctypes.data if saved else None, w.ctypes.data if saved else None))
        return rows, ids, w

    def energies(self, j: int):
        e = (C.c_double * 3)()
        check(lib().umcx_batch_energies(self._h, j, e))
        return e[0], e[1], e[2]

    def close(self):
        if getattr(self, "_h", None):
            lib().umcx_batch_free(self._h)
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


def plain_loop(cfgs: list, device: int = 0, keep: bool = True):
    """The comparison baseline: one plain session per job, run one after another. Returns
    (per-job results, summed kernel ms, wall ms of the run calls)."""
    out, kms, wall = [], 0.0, 0.0
    for c in cfgs:
        with Simulation(c, device=device) as sim:
            t0 = time.perf_counter()
            kms += sim.run()
            wall += (time.perf_counter() - t0) * 1e3
            if keep:
                out.append((sim.fluence(), sim.energies()))
    return out, kms, wall
