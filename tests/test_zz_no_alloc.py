"""umcx_run_device_noalloc: the photon kernel on caller-owned device buffers, with no allocation.

Its results are checked against umcx_session_run's on the same configs (the session in the same
tally mode: global atomics, no LDS tile), on two calls in a row, the second with a smaller
volume, fewer photons, more media (4-bit staged labels) and more partial-path columns than the
first: detected rows, photon ids and final weights bitwise equal as multisets, detection counts
and launched energy exact, escaped and deposited energy to fp64 summation order, fluence within
the 1e-4 relative L2 parity bar (float32 atomics in another order: about 2e-5 measured at 2e5
photons, where the hottest cells sum a quarter million deposits) with the same nonzero cells. A volume whose packed
labels exceed the static scratch buffer is rejected before anything runs.
"""
import copy
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U
from umcx_amd._native import UmcxDeviceProblem, check, lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def srt(r):
    return r[np.lexsort(r.T[::-1])] if r.size else r


def shells40():
    """The five-media spherical-shell config cropped to a 40^3 volume (4-bit staged labels)."""
    c = copy.deepcopy(O.golden("sphshells")["cfg"])
    c["Domain"]["Dim"] = [40, 40, 40]
    for sh in c["Shapes"]:
        if "Grid" in sh:
            sh["Grid"]["Size"] = [40, 40, 40]
    return c


class DeviceRun:
    """A config's inputs and zeroed outputs in torch device memory, as umcx_device_problem."""

    def __init__(self, cfg, deposit_cache):
        import torch
        lab = U.domain(cfg)
        prob = O.problem_from_config(cfg, lab)
        u = lab.view(np.uint32).reshape(-1)
        lab8 = ((u & 0x7F) | (((u >> 31) & 1) << 7)).astype(np.uint8)
        nx, ny, nz = lab.shape
        nt = prob["maxgate"]
        cols = (1 if prob["savedetflag"] & 1 else 0) + (len(prob["media"]) if prob["savedetflag"] & 4 else 0) + \
            3 * ((1 if prob["savedetflag"] & 16 else 0) + (1 if prob["savedetflag"] & 32 else 0))
        self.cols = cols
        cap = prob["maxdetphotons"]
        dev = torch.device("cuda:0")
        self.t = dict(
            labels=torch.from_numpy(lab8).to(dev),
            media=torch.from_numpy(np.ascontiguousarray(prob["media"].reshape(-1))).to(dev),
            detpos=torch.from_numpy(np.ascontiguousarray(prob["detpos"].reshape(-1))).to(dev)
            if len(prob["detpos"]) else torch.zeros(4, dtype=torch.float32, device=dev),
            fluence=torch.zeros(nt * nx * ny * nz, dtype=torch.float32, device=dev),
            detp=torch.zeros(cap * max(cols, 1), dtype=torch.float32, device=dev),
            detid=torch.zeros(cap, dtype=torch.int32, device=dev),
            detw=torch.zeros(cap, dtype=torch.float32, device=dev),
            ndet=torch.zeros(1, dtype=torch.int32, device=dev),
            energies=torch.zeros(3, dtype=torch.float64, device=dev))
        q = UmcxDeviceProblem()
        for i, v in enumerate((nx, ny, nz, nt)):
            q.dims[i] = v
        q.mediumnum, q.detnum = len(prob["media"]), len(prob["detpos"])
        q.isreflect, q.issavedet, q.savedetflag = prob["isreflect"], prob["issavedet"], prob["savedetflag"]
        q.outputtype, q.srctype = prob["outputtype"], prob["srctype"]
        q.tstart, q.tend, q.rtstep, q.unitinmm = prob["tstart"], prob["tend"], prob["rtstep"], prob["unitinmm"]
        for name in ("srcpos", "srcdir", "srcparam1", "srcparam2"):
            arr = getattr(q, name)
            for i, v in enumerate(prob[name][:4]):
                arr[i] = float(v)
        for i, v in enumerate(prob["seeds"]):
            q.seeds[i] = v
        q.maxdetphotons = cap
        q.deposit_cache = deposit_cache
        for name, ten in self.t.items():
            setattr(q, name, ten.data_ptr())
        self.q = q

    def run(self, first, count):
        import torch
        stream = torch.cuda.current_stream().cuda_stream
        rc = lib().umcx_run_device_noalloc(C.byref(self.q), first, count, C.c_void_p(stream))
        torch.cuda.synchronize()
        return rc

    def results(self):
        n = int(self.t["ndet"].cpu().numpy()[0])
        saved = min(n, self.q.maxdetphotons)
        rows = self.t["detp"].cpu().numpy()[:saved * self.cols].reshape(saved, self.cols)
        ids = self.t["detid"].cpu().numpy()[:saved].view(np.uint32)
        w = self.t["detw"].cpu().numpy()[:saved]
        return dict(fluence=self.t["fluence"].cpu().numpy(), rows=rows, ids=ids, w=w, detected=n,
                    energies=tuple(self.t["energies"].cpu().numpy().tolist()))


def session_results(cfg, count, deposit_cache):
    with U.Simulation(cfg, tally_mode=0, deposit_cache=1 if deposit_cache else -1) as sim:
        sim.run(0, count)
        rows, ids, w = sim.detp()
        return dict(fluence=sim.fluence().reshape(-1), rows=rows, ids=ids, w=w, detected=sim.detected()[0],
                    energies=sim.energies(), lds_bits=sim.lds_label_bits())


def check_same(a, b):
    assert a["detected"] == b["detected"]
    key = lambda r: srt(np.column_stack([r["rows"], r["ids"].view(np.float32), r["w"]]))
    assert np.array_equal(key(a).view(np.uint32), key(b).view(np.uint32))
    assert a["energies"][1] == b["energies"][1]
    for k in (0, 2):
        assert a["energies"][k] == pytest.approx(b["energies"][k], rel=1e-11)
    fa, fb = a["fluence"], b["fluence"]
    assert fa.shape == fb.shape and np.isfinite(fa).all()
    assert np.linalg.norm(fa - fb) <= 1e-4 * np.linalg.norm(fb)
    assert np.array_equal(fa != 0, fb != 0)


def test_noalloc_run_equals_session_run_on_two_calls():
    import torch
    assert torch.cuda.is_available()
    c2 = O.golden("c2")["cfg"]
    say("noalloc call 1: heterogeneous reflective 60^3, 3 media, 200000 photons, no deposit cache")
    r1 = DeviceRun(c2, deposit_cache=0)
    check(r1.run(0, 150000))
    check(r1.run(150000, 50000))                     # a range split over two calls accumulates
    a1 = r1.results()
    say(f"noalloc call 1 done: {a1['detected']} detections; session run")
    b1 = session_results(c2, 200000, False)
    assert b1["lds_bits"] == 2
    check_same(a1, b1)
    del r1

    small = shells40()
    say("noalloc call 2: five-media shells 40^3, 4-bit labels, 5 path columns, 50000 photons, deposit cache on")
    r2 = DeviceRun(small, deposit_cache=1)
    check(r2.run(0, 50000))
    a2 = r2.results()
    say(f"noalloc call 2 done: {a2['detected']} detections; session run")
    b2 = session_results(small, 50000, True)
    assert b2["lds_bits"] == 2          # the session packs ids 1-4 as 2-bit codes; this entry point, the ids at 4 bits
    check_same(a2, b2)
    assert a2["detected"] > 0 and a2["rows"].shape[1] == 1 + 5 + 6


def test_noalloc_rejects_labels_beyond_the_scratch_buffer():
    import torch
    big = O.golden("sphshells")["cfg"]                # 60^3 voxels x 4 bits = 108,000 bytes > 96 KiB
    say("noalloc rejection: five-media 60^3")
    r = DeviceRun(big, deposit_cache=0)
    rc = r.run(0, 1000)
    assert rc != 0
    msg = U.last_error()
    assert "scratch" in msg and "108000" in msg, msg
    assert float(r.t["energies"].abs().sum().cpu()) == 0.0 and int(r.t["ndet"].cpu()[0]) == 0
    q = UmcxDeviceProblem()                            # an empty problem: an error, nothing launched
    assert lib().umcx_run_device_noalloc(C.byref(q), 0, 10, None) != 0
    assert lib().umcx_run_device_noalloc(None, 0, 10, None) != 0
    torch.cuda.synchronize()
