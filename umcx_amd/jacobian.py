"""Differentiable detected intensities: d log I_sd / d(per-voxel mua, mus) on the GPU.

For sources s and detectors d, I_sd is the sum of the final weights of the photons launched from
s and detected by d (the weight decays as exp(-sum_v mua_v L_v)). Its derivatives follow from
perturbation Monte Carlo on the detected photons' own paths:

    d I / d mua_v = - sum_p w_p L_pv
    d I / d mus_v =   sum_p w_p (n_pv / mus_v - L_pv)

with L_pv the path length (mm) and n_pv the number of scatterings of photon p in voxel v. Photon
i is a pure function of i (src/umcx.cpp:761), so the forward keeps only (photon index, detector,
final weight) per detection and the backward *replays* exactly those photons with a hand-written
kernel (``umcx_session_replay_sources``: the records of every source in one launch) that adds
coef_p x the two terms along the path: coef = w/I for the Jacobian of log I, coef = w g_sd / I_sd
for a vector-Jacobian product. Under common random
numbers a mua perturbation leaves every trajectory unchanged, so central differences of this
estimator match the replay gradient up to O(h^2).

``LogIntensity`` exposes it as a ``torch.autograd.Function`` over per-voxel mua/mus tensors.
"""
from __future__ import annotations

import copy
import ctypes as C

import numpy as np

from ._native import check, lib
from .api import Simulation


def _torch():
    import torch
    return torch


class MultiSourceModel:
    """One volume, several pencil sources and the configuration's detectors.

    ``cfg`` is a umcx JSON config (its Optode.Source gives the type and parameters; positions
    and directions come from ``sources``); ``nphoton`` photons are launched per source.
    """

    def __init__(self, cfg: dict, sources, nphoton: int, device: int = 0, max_records: int | None = None):
        cfg = copy.deepcopy(cfg)
        S = cfg.setdefault("Session", {})
        S["DoSaveVolume"] = 0
        S["DoPartialPath"] = 1
        S["SaveDetFlag"] = 1
        S["Photons"] = int(nphoton)
        S["MaxDetPhoton"] = int(max_records or len(sources) * max(1_000_000, nphoton // 4))
        self.cfg = cfg
        self.sources = [(list(map(float, p)), list(map(float, d))) for p, d in sources]
        self.nphoton = int(nphoton)
        self.device = device
        self.sim = Simulation(cfg, device=device, tally_mode=0)
        nx, ny, nz, _ = self.sim.dims
        self.shape = (nx, ny, nz)
        self.nvox = nx * ny * nz
        self.ndet = int(self.sim.info["detnum"])
        self.records = []          # per source: (ids int32, det int64, w float32) host arrays
        self.I = None              # (S, D) float64

    # -- forward ---------------------------------------------------------------------------------
    def _set_props(self, mua, mus):
        L = lib()
        if mua is None:
            check(L.umcx_session_set_voxel_props(self.sim._h, None, None, 0))
            return
        torch = _torch()
        if isinstance(mua, torch.Tensor):
            mua = mua.detach().to(device=f"cuda:{self.device}", dtype=torch.float32).contiguous()
            mus = mus.detach().to(device=f"cuda:{self.device}", dtype=torch.float32).contiguous()
            torch.cuda.synchronize()
            check(L.umcx_session_set_voxel_props(self.sim._h, C.c_void_p(mua.data_ptr()),
                                                 C.c_void_p(mus.data_ptr()), 1))
        else:
            mua = np.ascontiguousarray(mua, dtype=np.float32)
            mus = np.ascontiguousarray(mus, dtype=np.float32)
            check(L.umcx_session_set_voxel_props(self.sim._h, mua.ctypes.data, mus.ctypes.data, 0))

    def forward(self, mua=None, mus=None) -> np.ndarray:
        """Simulate every source; returns log I as a (sources, detectors) float64 array."""
        self._set_props(mua, mus)
        self.records = []
        S, N = len(self.sources), self.nphoton
        I = np.zeros((S, self.ndet), dtype=np.float64)
        L = lib()
        # every source in one launch: photon slot p is photon p % N of source p // N
        table = np.array([list(p) + list(d) for p, d in self.sources], dtype=np.float32).reshape(-1)
        check(L.umcx_session_set_sources(self.sim._h, table.ctypes.data, S, N))
        self.sim.reset()
        self.forward_ms = self.sim.run(0, S * N)
        check(L.umcx_session_set_sources(self.sim._h, None, 0, 0))
        detected, saved = self.sim.detected()
        if saved < detected:
            raise RuntimeError(f"{detected - saved} detected photons did not fit MaxDetPhoton; raise max_records")
        rows, slots, w = self.sim.detp()
        det = rows[:, 0].astype(np.int64) - 1
        src = slots.astype(np.int64) // N
        order = np.argsort(src, kind="stable")
        src, slots, det, w = src[order], slots[order], det[order], w[order]
        bounds = np.searchsorted(src, np.arange(S + 1))
        for s in range(S):
            a, b = bounds[s], bounds[s + 1]
            np.add.at(I[s], det[a:b], w[a:b].astype(np.float64))
            self.records.append(((slots[a:b] - s * N).astype(np.int32), det[a:b], w[a:b]))
        self.I = I
        with np.errstate(divide="ignore"):
            return np.log(I)

    # -- backward -------------------------------------------------------------------------------
    def _replay(self, s, coef, slices, gmua, gmus, final_w=None):
        torch = _torch()
        ids, det, w = self.records[s]
        n = len(ids)
        if n == 0:
            return 0.0
        dev = f"cuda:{self.device}"
        t_ids = torch.from_numpy(ids).to(dev)
        t_sl = torch.from_numpy(np.ascontiguousarray(slices, dtype=np.int32)).to(dev)
        t_cf = torch.from_numpy(np.ascontiguousarray(coef, dtype=np.float32)).to(dev)
        torch.cuda.synchronize()
        L = lib()
        pos, dirv = self.sources[s]
        check(L.umcx_session_set_source(self.sim._h, (C.c_float * 3)(*pos), (C.c_float * 3)(*dirv)))
        ms = C.c_double(0.0)
        check(L.umcx_session_replay(self.sim._h, C.c_void_p(t_ids.data_ptr()), C.c_void_p(t_sl.data_ptr()),
                                    C.c_void_p(t_cf.data_ptr()), n, C.c_void_p(gmua.data_ptr()),
                                    C.c_void_p(gmus.data_ptr()), self.nvox,
                                    C.c_void_p(final_w.data_ptr()) if final_w is not None else None, C.byref(ms)))
        return ms.value

    def _replay_all(self, coefs, slices, gmua, gmus):
        """Replay the records of every source in one launch (umcx_session_replay_sources)."""
        torch = _torch()
        dev = f"cuda:{self.device}"
        ids = np.concatenate([r[0] for r in self.records]) if self.records else np.zeros(0, np.int32)
        n = len(ids)
        if n == 0:
            return 0.0
        src = np.concatenate([np.full(len(r[0]), s, np.int32) for s, r in enumerate(self.records)])
        table = np.array([list(p) + list(d) for p, d in self.sources], dtype=np.float32).reshape(-1)
        t_ids = torch.from_numpy(ids.astype(np.int32)).to(dev)
        t_sl = torch.from_numpy(np.ascontiguousarray(np.concatenate(slices), dtype=np.int32)).to(dev)
        t_cf = torch.from_numpy(np.ascontiguousarray(np.concatenate(coefs), dtype=np.float32)).to(dev)
        t_src = torch.from_numpy(src).to(dev)
        t_tab = torch.from_numpy(table).to(dev)
        torch.cuda.synchronize()
        ms = C.c_double(0.0)
        check(lib().umcx_session_replay_sources(
            self.sim._h, C.c_void_p(t_ids.data_ptr()), C.c_void_p(t_sl.data_ptr()), C.c_void_p(t_cf.data_ptr()),
            C.c_void_p(t_src.data_ptr()), C.c_void_p(t_tab.data_ptr()), n, C.c_void_p(gmua.data_ptr()),
            C.c_void_p(gmus.data_ptr()), self.nvox, None, C.byref(ms)))
        return ms.value

    def vjp(self, grad_logI):
        """sum_sd grad[s,d] * d log I_sd / d(mua, mus): two (nx, ny, nz) float32 cuda tensors."""
        torch = _torch()
        g = np.asarray(grad_logI.detach().cpu().numpy() if hasattr(grad_logI, "detach") else grad_logI,
                       dtype=np.float64)
        dev = f"cuda:{self.device}"
        gmua = torch.zeros(self.nvox, dtype=torch.float32, device=dev)
        gmus = torch.zeros(self.nvox, dtype=torch.float32, device=dev)
        coefs, slices = [], []
        for s in range(len(self.sources)):
            ids, det, w = self.records[s]
            Isd = self.I[s, det]
            coefs.append(np.where(Isd > 0, w * g[s, det] / np.where(Isd > 0, Isd, 1.0), 0.0))
            slices.append(np.zeros(len(ids), np.int32))
        self.replay_ms = self._replay_all(coefs, slices, gmua, gmus)
        return gmua.view(self.shape), gmus.view(self.shape)

    def jacobian(self):
        """Full d log I_sd / d(mua_v, mus_v): (sources*detectors, nx*ny*nz, 2) float32 cuda tensor."""
        torch = _torch()
        dev = f"cuda:{self.device}"
        S, D = len(self.sources), self.ndet
        J = torch.zeros((2, S * D, self.nvox), dtype=torch.float32, device=dev)
        coefs, slices = [], []
        for s in range(S):
            ids, det, w = self.records[s]
            Isd = self.I[s, det]
            coefs.append(np.where(Isd > 0, w / np.where(Isd > 0, Isd, 1.0), 0.0))
            slices.append((s * D + det).astype(np.int32))
        self.replay_ms = self._replay_all(coefs, slices, J[0], J[1])
        return J.permute(1, 2, 0)

    def replay_weights(self, s: int):
        """Final weights of source s's detected photons as the replay kernel recomputes them."""
        torch = _torch()
        dev = f"cuda:{self.device}"
        n = len(self.records[s][0])
        out = torch.zeros(max(n, 1), dtype=torch.float32, device=dev)
        scratch = torch.zeros(self.nvox, dtype=torch.float32, device=dev)
        self._replay(s, np.zeros(n, np.float32), np.zeros(n, np.int32), scratch, scratch, out)
        return out[:n].cpu().numpy()

    def close(self):
        self.sim.close()


def _make_function():
    torch = _torch()

    class LogIntensity(torch.autograd.Function):
        """log I_sd for per-voxel mua, mus (1/mm, shape of the volume); backward replays the
        detected photons on the GPU."""

        @staticmethod
        def forward(ctx, mua, mus, model: MultiSourceModel):
            logI = model.forward(mua, mus)
            ctx.model = model
            return torch.from_numpy(logI).to(mua.device, torch.float32)

        @staticmethod
        def backward(ctx, grad):
            gmua, gmus = ctx.model.vjp(grad)
            return gmua, gmus, None

    return LogIntensity


_FN = None


def log_intensity(mua, mus, model: MultiSourceModel):
    """Differentiable log detected intensity, shape (sources, detectors)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(mua, mus, model)


def slab_config(seed: int = 11, nphoton: int = 10_000_000, inclusion_mua: float = 0.035) -> dict:
    """The differentiable workload: a 64x64x32 slab of 1 mm voxels (mua 0.01/mm, mus 1.0/mm,
    g 0.9, n 1.37) with one 6x6x6 inclusion (label 2, mua ``inclusion_mua``) at a seeded
    position, and 16 disk detectors of radius 1.5 mm on a 4x4 grid (pitch 8 mm from (20, 20)) on
    the z=0 face. Sources come from ``grid_sources(64, 64, 4, 16.5)``."""
    rng = np.random.default_rng(seed)
    vol = np.ones((64, 64, 32), dtype=np.int32)
    x0, y0, z0 = (int(rng.integers(8, 64 - 14)), int(rng.integers(8, 64 - 14)), int(rng.integers(4, 32 - 10)))
    vol[x0:x0 + 6, y0:y0 + 6, z0:z0 + 6] = 2
    return {"Session": {"ID": "slab", "Photons": int(nphoton), "RNGSeed": 1648335518, "DoMismatch": 1,
                        "DoPartialPath": 1, "SaveDetFlag": 5, "OutputType": "x"},
            "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
            "Domain": {"Dim": [64, 64, 32], "LengthUnit": 1,
                       "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1},
                                 {"mua": 0.01, "mus": 1.0, "g": 0.9, "n": 1.37},
                                 {"mua": float(inclusion_mua), "mus": 1.0, "g": 0.9, "n": 1.37}]},
            "Optode": {"Source": {"Type": "pencil", "Pos": [24.5, 24.5, 0], "Dir": [0, 0, 1]},
                       "Detector": [{"Pos": [20 + 8 * i, 20 + 8 * j, 0], "R": 1.5} for j in range(4) for i in range(4)]},
            "Shapes": {"_ArraySize_": [64, 64, 32], "_ArrayType_": "int32",
                       "_ArrayData_": vol.reshape(-1).tolist()}}


def grid_sources(nx, ny, n_side, margin, z=0.0, offset=0.0):
    """Pencil sources (+z) on an n_side x n_side grid inside [margin, n-margin]."""
    xs = np.linspace(margin, nx - margin, n_side) + offset
    ys = np.linspace(margin, ny - margin, n_side) + offset
    return [((float(x), float(y), float(z)), (0.0, 0.0, 1.0)) for y in ys for x in xs]
