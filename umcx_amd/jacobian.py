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


_HIP = None


def _device_copy(dst: int, src: int, nbytes: int) -> None:
    """Device-to-device hipMemcpy through PyTorch's HIP runtime (the one libumcx.so runs on)."""
    global _HIP
    if _HIP is None:
        import os
        torch = _torch()
        h = C.CDLL(os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so"))
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        _HIP = h
    rc = _HIP.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    if rc != 0:
        raise RuntimeError(f"hipMemcpy device to device failed ({rc})")


def _pair_sums(pair, w, n: int):
    """sum of w and of w^2 per (source, detector) pair index in [0, n), on the GPU: sorted by pair,
    prefix sums, differences at the segment bounds (float atomics into a few hundred bins would
    serialise on the same addresses)."""
    torch = _torch()
    order = torch.argsort(pair)
    ps, ws = pair[order], w[order]
    zero = torch.zeros(1, dtype=w.dtype, device=w.device)
    c1 = torch.cat([zero, torch.cumsum(ws, 0)])
    c2 = torch.cat([zero, torch.cumsum(ws * ws, 0)])
    bounds = torch.searchsorted(ps, torch.arange(n + 1, dtype=ps.dtype, device=ps.device))
    return c1[bounds[1:]] - c1[bounds[:-1]], c2[bounds[1:]] - c2[bounds[:-1]]


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
        self._records = None       # per source host arrays, built from the device records on use
        self._dev = None           # device records of the last forward (ids, source, pair, weight, I)
        self.nrecords = 0
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
        """Simulate every source; returns log I as a (sources, detectors) float64 array. The
        detection records stay on the GPU: they are copied device to device out of the session's
        buffers, and I_sd, the record slices and the coefficients are formed there."""
        torch = _torch()
        self._set_props(mua, mus)
        S, N, D = len(self.sources), self.nphoton, self.ndet
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
        dev = f"cuda:{self.device}"
        cols = int(self.sim.info["detp_columns"])
        d_ids, d_w, d_rows = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(L.umcx_session_detp_device(self.sim._h, C.byref(d_ids), C.byref(d_w), C.byref(d_rows)))
        slots = torch.empty(max(saved, 1), dtype=torch.int32, device=dev)
        w = torch.empty(max(saved, 1), dtype=torch.float32, device=dev)
        rows = torch.empty(max(saved, 1) * cols, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        if saved:
            _device_copy(slots.data_ptr(), d_ids.value, 4 * saved)
            _device_copy(w.data_ptr(), d_w.value, 4 * saved)
            _device_copy(rows.data_ptr(), d_rows.value, 4 * saved * cols)
        slots, w = slots[:saved], w[:saved]
        slot64 = slots.to(torch.int64) & 0xFFFFFFFF
        src = slot64 // N
        det = rows[:saved * cols].view(saved, cols)[:, 0].to(torch.int64) - 1
        pair = src * D + det
        I, w2 = _pair_sums(pair, w.double(), S * D)
        self._dev = {"ids": (slot64 - src * N).to(torch.int32), "src": src.to(torch.int32), "pair": pair,
                     "w": w, "I": I, "w2": w2, "table": torch.from_numpy(table).to(dev)}
        self.nrecords = int(saved)
        self.I = I.view(S, D).cpu().numpy()
        self._records = None
        with np.errstate(divide="ignore"):
            return np.log(self.I)

    @property
    def records(self):
        """Per source (photon ids int32, detector int64, final weight float32) host arrays of the
        last forward, copied from the GPU on first use."""
        if self._records is None:
            d = self._dev
            src = d["src"].cpu().numpy().astype(np.int64)
            ids, det, w = d["ids"].cpu().numpy(), (d["pair"].cpu().numpy() - src * self.ndet), d["w"].cpu().numpy()
            order = np.argsort(src, kind="stable")
            src, ids, det, w = src[order], ids[order], det[order], w[order]
            bounds = np.searchsorted(src, np.arange(len(self.sources) + 1))
            self._records = [(ids[a:b], det[a:b], w[a:b]) for a, b in zip(bounds[:-1], bounds[1:])]
        return self._records

    def noise(self) -> np.ndarray:
        """Relative standard deviation of each I_sd, sqrt(sum w^2) / sum w (NaN where nothing was
        detected), shape (sources, detectors)."""
        d = self._dev
        with np.errstate(divide="ignore", invalid="ignore"):
            return (d["w2"].sqrt() / d["I"]).view(len(self.sources), self.ndet).cpu().numpy()

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

    def _replay_all(self, coef, slices, gmua, gmus):
        """Replay the records of every source in one launch (umcx_session_replay_sources);
        coef (float32) and slices (int32) are device tensors, one entry per record."""
        torch = _torch()
        d = self._dev
        n = self.nrecords
        if n == 0:
            return 0.0
        coef = coef.to(torch.float32).contiguous()
        slices = slices.to(torch.int32).contiguous()
        torch.cuda.synchronize()
        ms = C.c_double(0.0)
        check(lib().umcx_session_replay_sources(
            self.sim._h, C.c_void_p(d["ids"].data_ptr()), C.c_void_p(slices.data_ptr()), C.c_void_p(coef.data_ptr()),
            C.c_void_p(d["src"].data_ptr()), C.c_void_p(d["table"].data_ptr()), n, C.c_void_p(gmua.data_ptr()),
            C.c_void_p(gmus.data_ptr()), self.nvox, None, C.byref(ms)))
        return ms.value

    def _coef(self, weights=None):
        """w / I_sd per record (times weights[s, d] when given), 0 where I_sd = 0; on the GPU."""
        torch = _torch()
        d = self._dev
        Ipair = d["I"][d["pair"]]
        num = d["w"].double()
        if weights is not None:
            num = num * weights.reshape(-1)[d["pair"]]
        return torch.where(Ipair > 0, num / torch.where(Ipair > 0, Ipair, torch.ones_like(Ipair)),
                           torch.zeros_like(Ipair))

    def vjp(self, grad_logI):
        """sum_sd grad[s,d] * d log I_sd / d(mua, mus): two (nx, ny, nz) float32 cuda tensors."""
        torch = _torch()
        dev = f"cuda:{self.device}"
        if hasattr(grad_logI, "detach"):
            g = grad_logI.detach().to(device=dev, dtype=torch.float64)
        else:
            g = torch.as_tensor(np.asarray(grad_logI, dtype=np.float64), device=dev)
        gmua = torch.zeros(self.nvox, dtype=torch.float32, device=dev)
        gmus = torch.zeros(self.nvox, dtype=torch.float32, device=dev)
        coef = self._coef(g)
        self.replay_ms = self._replay_all(coef, torch.zeros_like(self._dev["src"]), gmua, gmus)
        return gmua.view(self.shape), gmus.view(self.shape)

    def jacobian(self):
        """Full d log I_sd / d(mua_v, mus_v): (sources*detectors, nx*ny*nz, 2) float32 cuda tensor."""
        torch = _torch()
        dev = f"cuda:{self.device}"
        S, D = len(self.sources), self.ndet
        J = torch.zeros((2, S * D, self.nvox), dtype=torch.float32, device=dev)
        self.replay_ms = self._replay_all(self._coef(), self._dev["pair"], J[0], J[1])
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
