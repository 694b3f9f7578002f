#!/usr/bin/env python3
"""Benchmark of the umcx photon-propagation hot path on one MI355X.

Workload (the headline configuration): heterogeneous reflective cube -- 60^3 voxels of 1 mm,
label 1 (mua 0.005/mm, mus 1.0/mm, g 0.01, n 1.37) with a centred sphere of radius 15 voxels
(label 2: mua 0.002, mus 5.0, g 0.9, n 1.0), Fresnel reflection at the n=1 exterior, pencil
source at (29.5, 29.5, 0) along +z, 50 time gates of 0.1 ns, 4 disk detectors of radius 1 on the
z=0 face, 1e8 photons per step. A step is one full simulation of those 1e8 photons on device-
resident inputs (the label volume and media are uploaded once, outside the timed region).

Prints one JSON line. ``roofline`` prices the photon kernel's algorithmic FP32 work (per-photon
op model of the photon loop, counted from src/umcx.cpp:315-428: 34 per step, 9 per voxel crossing,
62 per scatter, 1 per random draw, 35 per Fresnel test) against the MI355X FP32 peak, using the
kernel's HIP-event time on its own stream. ``cpu_baseline`` times the reference umcx binary built
from source (or, when it is absent, the CPU oracle) on this machine's host cores over a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/ms on 1 GPU + Jacobian & batched speedup vs plain at matched noise"
PHOTONS = 100_000_000
# per-photon event counts of this workload (oracle-measured; pinned by tests/test_workload_model.py)
COUNTS = {"S": 346.63, "X": 189.12, "K": 157.29, "R": 475.06, "F": 3.649}
OPS_PER_PHOTON = 34 * COUNTS["S"] + 9 * COUNTS["X"] + 62 * COUNTS["K"] + COUNTS["R"] + 35 * COUNTS["F"]
FP32_PEAK_TFLOPS = 157.3          # MI355X published FP32 (vector/matrix) peak

CONFIG = {
    "Session": {"ID": "hetero_reflective", "Photons": PHOTONS, "RNGSeed": 1648335518, "DoMismatch": 1,
                "DoSaveVolume": 1, "DoPartialPath": 1, "SaveDetFlag": 5, "OutputType": "x"},
    "Forward": {"T0": 0, "T1": 5e-9, "Dt": 1e-10},
    "Domain": {"Dim": [60, 60, 60], "LengthUnit": 1,
               "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.005, "mus": 1.0, "g": 0.01, "n": 1.37},
                         {"mua": 0.002, "mus": 5.0, "g": 0.9, "n": 1.0}]},
    "Optode": {"Source": {"Type": "pencil", "Pos": [29.5, 29.5, 0], "Dir": [0, 0, 1]},
               "Detector": [{"Pos": [20, 29, 0], "R": 1}, {"Pos": [40, 29, 0], "R": 1},
                            {"Pos": [29, 20, 0], "R": 1}, {"Pos": [29, 40, 0], "R": 1}]},
    "Shapes": [{"Grid": {"Tag": 1, "Size": [60, 60, 60]}}, {"Sphere": {"O": [30, 30, 30], "R": 15, "Tag": 2}}],
}
LINE = re.compile(r"speed ([0-9.eE+-]+) photon/ms")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> int:
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0)))
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(target_s: float = 12.0) -> dict:
    """Time the reference CPU build (or the oracle) on a bounded sample of the same workload."""
    threads = host_threads()
    ref = os.path.join(ROOT, "oracle", "_ref", "umcx")
    if os.path.exists(ref):
        def run(n):
            cfg = json.loads(json.dumps(CONFIG))
            cfg["Session"]["Photons"] = int(n)
            cfg["Session"]["ID"] = "cpu_sample"
            with tempfile.TemporaryDirectory() as td:
                path = os.path.join(td, "in.json")
                with open(path, "w") as f:
                    json.dump(cfg, f)
                out = subprocess.run([ref, "-f", path], cwd=td, capture_output=True, text=True,
                                     env=dict(os.environ, OMP_NUM_THREADS=str(threads)), timeout=600).stdout
            m = LINE.search(out)
            if not m:
                raise RuntimeError(f"reference run printed no speed line: {out!r}")
            return float(m.group(1))
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O  # the checker, used here only as the timed CPU baseline

        import numpy as np
        prob = O.problem_from_config(CONFIG, np.asarray(O_domain(), dtype=np.int32))

        def run(n):
            t0 = time.perf_counter()
            O.run(prob, nphoton=int(n), nthreads=threads)
            return n / ((time.perf_counter() - t0) * 1e3)
        kind = "port"
    probe = 50_000
    rate = run(probe)
    n = int(min(max(rate * target_s * 1e3, probe), 2e8))
    rate = run(n)
    return {"value": round(rate, 2), "unit": "photons/ms", "cores": threads, "kind": kind,
            "sample": f"{n} photons of the same workload (fluence and detectors on), {threads} OpenMP threads; "
                      "rate from the binary's own timing window"}


def O_domain():
    import umcx_amd as U
    return U.domain(CONFIG)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=int, default=PHOTONS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(photon ranges would shard across GPUs with one fluence all-reduce, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import umcx_amd as U
    try:
        import torch
        sync = torch.cuda.synchronize
    except Exception:
        torch, sync = None, (lambda: None)

    cfg = json.loads(json.dumps(CONFIG))
    cfg["Session"]["Photons"] = args.photons
    log("bench: building session")
    sim = U.Simulation(cfg)
    log(f"bench: block {sim.info['block']} x {sim.info['blocks']} workgroups")
    for w in range(args.warmup):
        sim.reset()
        ms = sim.run()
        log(f"bench: warmup {w} {ms:.1f} ms, LDS tally box {sim.tally_box()}")
    kernel_ms = []
    sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        sim.reset()
        kernel_ms.append(sim.run())
        log(f"bench: step {s} kernel {kernel_ms[-1]:.1f} ms")
    sync()
    wall_ms = (time.perf_counter() - t0) * 1e3
    esc, tot, dep = sim.energies()
    detected, saved = sim.detected()
    sim_box = sim.tally_box()
    sim.close()
    ms_per_step = wall_ms / args.steps
    value = args.photons * args.steps / wall_ms
    kavg = sum(kernel_ms) / len(kernel_ms)
    achieved_tflops = OPS_PER_PHOTON * args.photons / (kavg * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "photons/ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "heterogeneous reflective: 60^3 1 mm voxels, sphere r=15 (n=1.0) in n=1.37 medium, "
                               "Fresnel on, pencil (29.5,29.5,0)+z, 50 gates x 0.1 ns, 4 disk detectors r=1",
                   "photons_per_step": args.photons, "seed": 1648335518},
        "roofline": {"bound": "valu", "achieved": round(achieved_tflops, 3), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 5), "traffic": None,
                     "ops_per_photon": round(OPS_PER_PHOTON, 1), "kernel_ms": round(kavg, 3)},
        "tally_box": sim_box,
        "detected_per_photon": detected / args.photons,
        "absorbed_percent": (tot - esc) / tot * 100.0,
    }
    if not args.no_cpu:
        log("bench: CPU baseline")
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # report, never fake
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
