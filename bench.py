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
kernel's HIP-event time on its own stream. Beside it are the two ceilings the kernel runs
against, each priced with per-photon rocprofv3 counters read from the kernel's committed profile
summary (``counters.build_match`` says whether they were measured on the sources built now):
``roofline.valu_issue`` -- busy VALU quad-cycles per second against 256 CUs x 4 SIMDs x 2.4 GHz / 4,
the ruler the issue-rate probes of umcx_amd/bin/roofline_probe calibrate -- and
``roofline.atomics``, memory-side float atomics per second against the rate the same probe
measures; ``roofline.bound`` names the larger. ``cpu_baseline`` times the reference umcx binary
built from source (or, when it is absent, the CPU oracle, and says so) on this machine's host
cores over bounded samples of the same workload and of the small cube and the skin model, with
the host's CPU model and glibc and the binary's provenance.
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
# VALU issue capacity: one quad-cycle (a wave64 instruction of the 4-cycle classes) per SIMD every
# 4 clocks, 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock = 6.144e11 busy quad-cycles/s. The
# issue-rate probes of umcx_amd/bin/roofline_probe --valu (FP32 FMA and add, int32, compare+select,
# move, FP64 FMA, reciprocal) each fill 98-99.5 % of it as rocprofv3 counts it: busy quad-cycles =
# SQ_ACTIVE_INST_VALU - SQ_ACTIVE_INST_VALU2 (dual-issued pairs counted once), against
# GRBM_GUI_ACTIVE per XCD x 256 CUs (profiles/roofline_probe_gfx950.md). The kernels' busy
# quad-cycles per photon are read from their committed counter summaries.
VALU_BUSY_PEAK = 256 * 4 * 2.4e9 / 4
# memory-side float atomics per second that roofline_probe measures for hashed addresses
ATOMIC_PEAK = 2.11e10
PHOTON_PROFILE = "profiles/photon_kernel_hetero_reflective_1e8.md"
BATCH_PROFILE = "profiles/batch_kernel_uneven256.md"


def counter_summary(profile: str) -> dict:
    """The machine-readable counter block of a committed profile summary, and whether the kernel
    sources it was measured on are the ones built now (sha256 over the listed files)."""
    import hashlib
    text = open(os.path.join(ROOT, profile)).read()
    head = text.index("## Counter summary (read by bench.py)")
    a = text.index("```json", head) + len("```json")
    b = text.index("```", a)
    c = json.loads(text[a:b])
    h = hashlib.sha256()
    for f in c["kernel_sources"]:
        h.update(open(os.path.join(ROOT, f), "rb").read())
    c["profile"] = profile
    c["build_match"] = h.hexdigest() == c["kernel_sources_sha256"]
    return c


def ceilings(c: dict, photons: float, kernel_s: float) -> dict:
    """VALU-issue and memory-side-atomic ceilings of a kernel from its per-photon counters."""
    busy = c["valu_busy_quad_cycles_per_photon"] * photons / kernel_s
    atom = c["memory_side_atomics_per_photon"] * photons / kernel_s
    out = {"valu_issue": {"busy_quad_cycles_per_photon": c["valu_busy_quad_cycles_per_photon"],
                          "instr_per_photon": c["valu_instr_per_photon"],
                          "achieved_per_s": busy, "peak_per_s": VALU_BUSY_PEAK, "frac": round(busy / VALU_BUSY_PEAK, 4),
                          "active_lane_frac": c["active_lane_frac"]},
           "atomics": {"per_photon": c["memory_side_atomics_per_photon"], "achieved_per_s": atom,
                       "peak_per_s": ATOMIC_PEAK, "frac": round(atom / ATOMIC_PEAK, 4)},
           "counters": {"profile": c["profile"], "build_match": c["build_match"],
                        "kernel_sources_sha256": c["kernel_sources_sha256"]}}
    out["binding"] = max(("valu_issue", "atomics"), key=lambda k: out[k]["frac"])
    return out


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


def host_info() -> dict:
    """The host the CPU baseline runs on: CPU model, sockets, cores, threads per core, glibc."""
    info = {"threads_available": host_threads()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and keys[k.strip()] not in info:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except Exception as e:  # lscpu absent: say so
        info["lscpu_error"] = str(e)
    try:
        info["glibc"] = os.confstr("CS_GNU_LIBC_VERSION")
    except Exception:
        info["glibc"] = None
    return info


def reference_provenance() -> dict | None:
    """Source hashes, flags and compiler of oracle/_ref/umcx, written beside it by oracle/Makefile."""
    path = os.path.join(ROOT, "oracle", "_ref", "umcx.provenance")
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        k, _, v = line.strip().partition(" ")
        if k in ("flags", "compiler"):
            out[k] = v
        elif v.startswith("sha256 "):
            out[k + "_sha256"] = v.split()[1]
    out["build"] = "oracle/Makefile target ref: g++ on the reference's src/umcx.cpp, the flags of its CPU target " \
                   "(src/Makefile:8,64-65)"
    return out


def cpu_rate(cfg: dict | None, argv: list[str] | None, n: int, threads: int) -> float:
    """photons/ms of the reference binary on n photons of a config (JSON) or a command line, from
    its own timing window (src/umcx.cpp:812-840)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "umcx")
    with tempfile.TemporaryDirectory() as td:
        if cfg is not None:
            c = json.loads(json.dumps(cfg))
            c["Session"]["Photons"] = int(n)
            c["Session"]["ID"] = "cpu_sample"
            path = os.path.join(td, "in.json")
            with open(path, "w") as f:
                json.dump(c, f)
            cmd = [ref, "-f", path]
        else:
            cmd = [ref] + list(argv) + ["-n", str(int(n))]
        out = subprocess.run(cmd, cwd=td, capture_output=True, text=True,
                             env=dict(os.environ, OMP_NUM_THREADS=str(threads)), timeout=600).stdout
    m = LINE.search(out)
    if not m:
        raise RuntimeError(f"reference run printed no speed line: {out!r}")
    return float(m.group(1))


def bounded_sample(run, target_s: float, probe: int = 50_000) -> tuple[float, int]:
    rate = run(probe)
    n = int(min(max(rate * target_s * 1e3, probe), 2e8))
    return run(n), n


def cpu_baseline(target_s: float = 12.0, others_s: float = 6.0) -> dict:
    """Time the reference CPU build on bounded samples of the headline workload (and of the small
    homogeneous cube and the skin model). Without the reference binary the oracle restatement is
    timed instead, and the result says so."""
    threads = host_threads()
    ref = os.path.join(ROOT, "oracle", "_ref", "umcx")
    out = {"unit": "photons/ms", "cores": threads, "host": host_info()}
    if os.path.exists(ref):
        rate, n = bounded_sample(lambda k: cpu_rate(CONFIG, None, k, threads), target_s)
        out.update(value=round(rate, 2), kind="reference", reference=reference_provenance(),
                   sample=f"{n} photons of the same workload (fluence and detectors on), {threads} OpenMP threads; "
                          "rate from the binary's own timing window")
        others = {}
        for name, cfg, argv in (("cube60_1e6", None, ["--bench", "cube60", "-d", "0"]),
                                ("skinvessel_200cube_1e8", SKIN_CONFIG, None)):
            r, k = bounded_sample(lambda k: cpu_rate(cfg, argv, k, threads), others_s, probe=20_000)
            others[name] = {"value": round(r, 2), "photons": k}
        out["other_configs"] = others
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O  # the checker, used here only as the timed CPU baseline

        import numpy as np
        prob = O.problem_from_config(CONFIG, np.asarray(O_domain(), dtype=np.int32))

        def run(k):
            t0 = time.perf_counter()
            O.run(prob, nphoton=int(k), nthreads=threads)
            return k / ((time.perf_counter() - t0) * 1e3)
        rate, n = bounded_sample(run, target_s)
        out.update(value=round(rate, 2), kind="port",
                   note="reference binary oracle/_ref/umcx absent: the oracle's C restatement was timed instead",
                   sample=f"{n} photons of the same workload through the oracle, {threads} OpenMP threads")
    return out


def batched_leg(sync) -> dict:
    """256 uneven jobs (seed 7) in one persistent, graph-replayed launch vs a loop of plain
    launches of the same jobs. Both trace exactly the same photons (per-job seeds), so their
    Monte Carlo noise is matched by construction; the noise of the reported per-job quantity
    (fluence summed over the 5x5x5 voxels under the source) is measured from a second batch with
    every job's seed changed."""
    import copy

    import numpy as np

    from umcx_amd import batch as B
    jobs = B.uneven_jobs(seed=7, njobs=256)
    photons = sum(j["Session"]["Photons"] for j in jobs)
    t0 = time.perf_counter()
    bs = B.BatchSimulation(jobs)
    setup_s = time.perf_counter() - t0
    trials = bs.trials()
    chosen = next(t for t in trials if t["chosen"])
    sync()
    t0 = time.perf_counter()
    batch_ms = bs.run()
    batch_wall = (time.perf_counter() - t0) * 1e3
    log(f"bench: batched {photons:.3e} photons in {batch_ms:.1f} ms (tuned: block {chosen['block']} x "
        f"{chosen['blocks']}, tile {chosen['tile_cells']} cells)")
    roi_b = np.array([B.roi_sum(bs.fluence(j), jobs[j]) for j in range(len(jobs))])
    bs.close()
    jobs2 = copy.deepcopy(jobs)
    for j in jobs2:
        j["Session"]["RNGSeed"] = int(j["Session"]["RNGSeed"]) + 1
    with B.BatchSimulation(jobs2, block=chosen["block"], blocks=chosen["blocks"],
                           tally_mode=1 if chosen["tile_cells"] else 0) as bs2:
        bs2.run()
        roi_2 = np.array([B.roi_sum(bs2.fluence(j), jobs2[j]) for j in range(len(jobs2))])
    rel = np.abs(roi_b - roi_2) / np.sqrt(2.0) / (0.5 * (roi_b + roi_2))
    log("bench: plain loop over the same jobs")
    res, loop_kernel_ms, loop_wall_ms = B.plain_loop(jobs)
    roi_p = np.array([B.roi_sum(f, c) for (f, _), c in zip(res, jobs)])
    matched = matched_noise_batch(jobs, chosen, sync)
    return {
        "workload": "256 uneven jobs, seed 7: edges U[32,128], 2-6 box labels, mua U(0.001,0.05), mus U(0.5,20), "
                    "g U(0.7,0.95), n U(1.33,1.45), Fresnel on, pencil at the z=0 face centre, photons log-U[1e5,1e7], "
                    "one 0-5 ns gate",
        "photons": photons, "batch_ms": round(batch_ms, 2), "batch_wall_ms": round(batch_wall, 2),
        "photons_per_ms": round(photons / batch_ms, 1),
        "plain_loop_kernel_ms": round(loop_kernel_ms, 2), "plain_loop_wall_ms": round(loop_wall_ms, 2),
        "speedup_vs_plain_kernel": round(loop_kernel_ms / batch_ms, 3),
        "speedup_vs_plain_wall": round(loop_wall_ms / batch_wall, 3),
        "setup_and_tuning_s": round(setup_s, 2),
        # the plain loop's wall time includes each job's session set-up and pilot; this ratio
        # charges the batch its set-up and autotuning too
        "speedup_vs_plain_wall_incl_setup": round(loop_wall_ms / (batch_wall + setup_s * 1e3), 3),
        "tuned": {"block": chosen["block"], "blocks": chosen["blocks"], "tile_cells": chosen["tile_cells"],
                  "trials": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in t.items()} for t in trials]},
        "reported_quantity": "per-job fluence summed over the 5x5x5 voxels under the source",
        "noise_rel_std_rms": float(np.sqrt(np.mean(rel ** 2))), "noise_rel_std_max_est": float(rel.max()),
        "batch_vs_plain_max_rel_diff": float(np.max(np.abs(roi_b - roi_p) / roi_p)),
        "at_1pct_noise": matched,
        "roofline": ceilings(counter_summary(BATCH_PROFILE), photons, batch_ms * 1e-3),
    }


def replica_noise(cfgs, shape, reps: int, seed_base: int):
    """Relative standard deviation of each job's reported quantity from `reps` independent
    replicas (seeds changed per replica), all run in one batch."""
    import copy

    import numpy as np

    from umcx_amd import batch as B
    rep_jobs = []
    for c in cfgs:
        for r in range(reps):
            c2 = copy.deepcopy(c)
            c2["Session"]["RNGSeed"] = int(c["Session"]["RNGSeed"]) + seed_base + 1000 * (r + 1)
            rep_jobs.append(c2)
    with B.BatchSimulation(rep_jobs, **shape) as bs:
        bs.run()
        roi = np.array([B.roi_sum(bs.fluence(j), rep_jobs[j]) for j in range(len(rep_jobs))]).reshape(len(cfgs), reps)
    return roi.std(axis=1, ddof=1) / roi.mean(axis=1)


def matched_noise_batch(jobs, chosen, sync, target: float = 0.01, reps: int = 16, aim: float = 0.005) -> dict:
    """The same comparison at photon counts where every job's reported quantity has at most a
    1 % relative standard deviation. A first estimate comes from 8 replicas of N/8 photons per job;
    each job's count is then set for a deviation of `aim` (N (s/aim)^2, at least 2000), below the
    target so that the spread of the estimate itself stays inside it. At those counts `reps`
    independent replicas of every job measure its deviation; a job above 0.8 % gets
    1.3 (s/aim)^2 times its count and is measured again with new replicas (at most three times).
    The reported deviations are each job's last measurement."""
    import copy

    import numpy as np

    from umcx_amd import batch as B
    shape = dict(block=chosen["block"], blocks=chosen["blocks"], tally_mode=1 if chosen["tile_cells"] else 0)
    first = []
    for c in jobs:
        c2 = copy.deepcopy(c)
        c2["Session"]["Photons"] = max(1, int(c["Session"]["Photons"]) // 8)
        first.append(c2)
    rel_full = replica_noise(first, shape, 8, 0) / np.sqrt(8)                  # at the full count N
    small = copy.deepcopy(jobs)
    for c, s_ in zip(small, rel_full):
        n = int(c["Session"]["Photons"])
        c["Session"]["Photons"] = int(max(2000, np.ceil(n * (s_ / aim) ** 2)))
    rel = replica_noise(small, shape, reps, 100000)
    raised, rounds = set(), 0
    for rounds in range(1, 4):
        over = np.flatnonzero(rel > 0.8 * target)
        log(f"bench: 1% noise check {rounds}: max {rel.max():.4f}, {len(over)} jobs above {0.8 * target:.3f}")
        if len(over) == 0:
            break
        sub = []
        for j in over:
            c = small[j]["Session"]
            c["Photons"] = int(np.ceil(int(c["Photons"]) * 1.3 * (rel[j] / aim) ** 2))
            raised.add(int(j))
            sub.append(small[j])
        rel[over] = replica_noise(sub, shape, reps, 100000 * (rounds + 1))
    log(f"bench: 1% noise: max {rel.max():.4f} over {len(rel)} jobs")
    photons = sum(int(c["Session"]["Photons"]) for c in small)
    t0 = time.perf_counter()
    with B.BatchSimulation(small, **shape) as bs:
        setup_ms = (time.perf_counter() - t0) * 1e3
        bs.run()
        sync()
        t0 = time.perf_counter()
        batch_ms = bs.run()
        batch_wall = (time.perf_counter() - t0) * 1e3
    log(f"bench: plain loop at the 1% noise counts ({photons:.3e} photons)")
    _, loop_kernel_ms, loop_wall_ms = B.plain_loop(small, keep=False)
    return {
        "photons": photons, "photons_min": min(int(c["Session"]["Photons"]) for c in small),
        "photons_max": max(int(c["Session"]["Photons"]) for c in small),
        "batch_ms": round(batch_ms, 2), "batch_wall_ms": round(batch_wall, 2),
        "plain_loop_kernel_ms": round(loop_kernel_ms, 2), "plain_loop_wall_ms": round(loop_wall_ms, 2),
        "speedup_vs_plain_kernel": round(loop_kernel_ms / batch_ms, 3),
        "speedup_vs_plain_wall": round(loop_wall_ms / batch_wall, 3),
        "setup_ms": round(setup_ms, 2),
        "speedup_vs_plain_wall_incl_setup": round(loop_wall_ms / (batch_wall + setup_ms), 3),
        "noise_replicas": reps, "noise_checks": rounds, "jobs_raised": sorted(raised),
        "noise_rel_std_max": float(rel.max()), "noise_rel_std_rms": float(np.sqrt(np.mean(rel ** 2))),
        "noise_rel_std_median": float(np.median(rel)),
        "noise_estimator": f"per job, sample std / mean of the reported quantity over {reps} independent replicas "
                           "at its final count (its last measurement)",
    }


def jacobian_leg(sync, photons: int = 10_000_000, fd_voxels: int = 3, target: float = 0.01) -> dict:
    """The Jacobian comparison at 1e7 photons per source, then again at the photon count per
    source that brings every pair's relative standard deviation of I_sd to the target (both
    sides scale with the photon count, so the ratio is expected to hold)."""
    import numpy as np

    res = jacobian_at(sync, photons, fd_voxels, warm_forward=True)
    worst = res["noise_rel_std_max"]
    n1 = int(np.ceil(photons * 1.1 * (worst / target) ** 2 / 1e6)) * 1_000_000 if worst > target else photons
    log(f"bench: jacobian at {n1:.3e} photons per source for <= {target:.0%} noise")
    at = jacobian_at(sync, n1, 1)
    res["at_1pct_noise"] = {k: at[k] for k in ("photons_per_source", "forward_ms", "replay_ms", "records",
                                              "entries_per_s", "fd_plain_launch_ms", "speedup_vs_plain_loop",
                                              "fd_vs_replay_max_rel_diff", "noise_rel_std_median",
                                              "noise_rel_std_max")}
    res["at_1pct_noise"]["jacobian_inclusion_noise"] = inclusion_jacobian_noise(n1)
    return res


def inclusion_jacobian_noise(photons: int, batches: int = 8) -> dict:
    """Noise of the Jacobian itself: d log I_sd / d mua summed over the inclusion's voxels, per
    source-detector pair, from `batches` independent seed batches of photons/batches per source
    (the sample std across batches / sqrt(batches) is the relative std at `photons` per source).
    The same batches give the empirical relative std of the detected intensities I_sd at
    `photons` per source, beside the single-run estimate sqrt(sum w^2) / sum w."""
    import copy

    import numpy as np

    from umcx_amd import jacobian as J
    cfg = J.slab_config(nphoton=photons)
    srcs = J.grid_sources(64, 64, 4, 16.5)
    lab = np.asarray(__import__("umcx_amd").domain(cfg) & 0x7FFFFFFF).reshape(-1)
    incl = np.flatnonzero(lab == 2)
    vals, ivals = [], []
    for b in range(batches):
        c = copy.deepcopy(cfg)
        c["Session"]["RNGSeed"] = 1648335518 + 7919 * (b + 1)
        m = J.MultiSourceModel(c, srcs, photons // batches)
        m.forward()
        ivals.append(np.asarray(m.I, dtype=np.float64).reshape(-1))
        Jt = m.jacobian()
        import torch
        idx = torch.as_tensor(incl, device=Jt.device)
        vals.append(Jt[:, idx, 0].double().sum(dim=1).cpu().numpy())
        m.close()
        del Jt
    v = np.array(vals)                                         # (batches, pairs)
    rel = v.std(axis=0, ddof=1) / np.abs(v.mean(axis=0)) / np.sqrt(batches)
    iv = np.array(ivals)                                       # (batches, pairs)
    irel = iv.std(axis=0, ddof=1) / iv.mean(axis=0) / np.sqrt(batches)
    log(f"bench: inclusion Jacobian noise at {photons:.3e} per source: median {np.median(rel):.4f}, max {rel.max():.4f}; "
        f"I_sd from the same batches: median {np.median(irel):.4f}, max {irel.max():.4f}")
    return {"quantity": "d log I_sd / d mua summed over the inclusion's 216 voxels, per pair",
            "batches": batches, "photons_per_source_per_batch": photons // batches,
            "noise_rel_std_median": float(np.median(rel)), "noise_rel_std_max": float(rel.max()),
            "pairs_at_or_below_1pct": int((rel <= 0.01).sum()), "pairs": int(rel.size),
            "photons_per_source_for_every_pair_at_1pct": int(np.ceil(photons * max(1.0, (rel.max() / 0.01) ** 2))),
            "intensity_noise_from_batches": {"quantity": "detected intensity I_sd per source-detector pair",
                                             "noise_rel_std_median": float(np.median(irel)),
                                             "noise_rel_std_max": float(irel.max())}}


def jacobian_at(sync, photons: int, fd_voxels: int, warm_forward: bool = False) -> dict:
    """d log I_sd / d(mua_v, mus_v) for 16 sources x 16 detectors on the 64x64x32 slab by one
    forward launch per source plus the replay kernel, against the loop of plain launches that
    central differences under common random numbers need (2 launches per (source, voxel,
    parameter) column, each giving the 16 detectors' entries). A mua perturbation leaves every
    trajectory unchanged under common random numbers, so both estimate the same quantity from
    the same photons (matched noise); a few columns are computed both ways as a check."""
    import numpy as np

    from umcx_amd import jacobian as J
    cfg = J.slab_config(nphoton=photons)
    srcs = J.grid_sources(64, 64, 4, 16.5)
    # the first use of PyTorch's index and reduction kernels loads their code objects: a small
    # model runs the same steps once before anything is timed
    w = J.MultiSourceModel(cfg, srcs, 2000)
    w.forward()
    w.jacobian()
    w.close()
    m = J.MultiSourceModel(cfg, srcs, photons)
    first_wall = None
    if warm_forward:
        # PyTorch loads the sort and scan kernels it picks for this many records on their first
        # use (about 80 ms, once per process), so the timed forward is the model's second; the
        # first one's wall time is reported beside it
        sync()
        t0 = time.perf_counter()
        m.forward()
        sync()
        first_wall = (time.perf_counter() - t0) * 1e3
    sync()
    t0 = time.perf_counter()
    m.forward()
    fwd_wall = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    Jt = m.jacobian()
    sync()
    jac_wall = (time.perf_counter() - t0) * 1e3
    S, D, nvox = len(srcs), m.ndet, m.nvox
    entries = S * D * nvox * 2
    noise = m.noise()
    recs = m.nrecords
    log(f"bench: jacobian forward {m.forward_ms:.1f} ms, replay {m.replay_ms:.1f} ms, {recs} records")
    # the inclusion's voxels carry the largest mua sensitivity; FD columns of source 0 there
    lab = np.asarray(__import__("umcx_amd").domain(cfg) & 0x7FFFFFFF).reshape(-1)
    media = cfg["Domain"]["Media"]
    mua0 = np.array([media[l]["mua"] for l in lab], dtype=np.float32)
    mus0 = np.array([media[l]["mus"] for l in lab], dtype=np.float32)
    J0 = Jt[:D, :, 0].float().cpu().numpy()
    cand = np.argsort(-np.abs(J0).sum(axis=0))[:fd_voxels]
    fd_ms, fd_wall, err = 0.0, 0.0, []
    m1 = J.MultiSourceModel(cfg, srcs[:1], photons)
    for v in cand:
        h = 1e-3 * float(mua0[v])
        lp = []
        for sgn in (1.0, -1.0):
            mua = mua0.copy()
            mua[v] = np.float32(mua0[v] + sgn * h)
            t0 = time.perf_counter()
            lp.append(m1.forward(mua.reshape(m.shape), mus0.reshape(m.shape))[0])
            fd_wall += (time.perf_counter() - t0) * 1e3
            fd_ms += m1.forward_ms
        fd = (lp[0] - lp[1]) / (2 * h)
        ok = np.abs(J0[:, v]) > 1e-3 * np.abs(J0[:, v]).max()
        err.append(float(np.max(np.abs(fd[ok] - J0[ok, v]) / np.abs(J0[ok, v]))))
    m1.close()
    m.close()
    per_launch_ms = fd_ms / (2 * len(cand))
    per_launch_wall = fd_wall / (2 * len(cand))
    fd_total_ms = 2 * 2 * nvox * S * per_launch_ms         # 2 launches x 2 parameters per column
    rep_ms = m.forward_ms + m.replay_ms
    return {
        "workload": "64x64x32 slab (mua 0.01, mus 1.0, g 0.9, n 1.37), 6x6x6 inclusion mua 0.035, 16 pencil "
                    f"sources x 16 disk detectors r=1.5 mm on 4x4 grids at z=0, {photons:.3g} photons per source",
        "photons_per_source": photons, "entries": entries, "records": recs,
        "forward_ms": round(m.forward_ms, 2), "replay_ms": round(m.replay_ms, 2),
        "forward_wall_ms": round(fwd_wall, 2), "jacobian_wall_ms": round(jac_wall, 2),
        "first_forward_wall_ms": None if first_wall is None else round(first_wall, 2),
        "entries_per_s": entries / (rep_ms * 1e-3),
        "fd_plain_launch_ms": round(per_launch_ms, 3), "fd_plain_launch_wall_ms": round(per_launch_wall, 3),
        "fd_loop_ms_full_jacobian": fd_total_ms,
        "speedup_vs_plain_loop": fd_total_ms / rep_ms,
        "speedup_vs_plain_loop_wall": (2 * 2 * nvox * S * per_launch_wall) / (fwd_wall + jac_wall),
        "fd_check_voxels": [int(v) for v in cand], "fd_vs_replay_max_rel_diff": max(err),
        "reported_quantity": "detected intensity I_sd per source-detector pair",
        "noise_rel_std_median": float(np.nanmedian(noise)), "noise_rel_std_max": float(np.nanmax(noise)),
    }


SKIN_CONFIG = {
    "Session": {"ID": "skinvessel200", "Photons": 100_000_000, "RNGSeed": 1648335518, "DoMismatch": 1,
                "DoSaveVolume": 1, "DoPartialPath": 0, "OutputType": "e"},
    "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
    "Domain": {"Dim": [200, 200, 200], "LengthUnit": 0.005,
               "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 1e-5, "mus": 0, "g": 1, "n": 1},
                         {"mua": 1.657, "mus": 37.59, "g": 0.9, "n": 1.37},
                         {"mua": 0.0458, "mus": 35.65, "g": 0.9, "n": 1.37},
                         {"mua": 23.05, "mus": 9.398, "g": 0.9, "n": 1.37}]},
    "Optode": {"Source": {"Type": "disk", "Pos": [100, 100, 0], "Dir": [0, 0, 1], "Param1": [60, 0, 0, 0]}},
    "Shapes": [{"Grid": {"Tag": 3, "Size": [200, 200, 200]}}, {"ZLayers": [[0, 20, 1], [20, 32, 2]]},
               {"Cylinder": {"Tag": 4, "C0": [0, 100, 100], "C1": [200, 100, 100], "R": 20}}]}


def other_configs_leg(sync) -> dict:
    """Photons/ms of the other single-GPU configurations: the small homogeneous cube (60^3,
    reflection off, 1e6 photons, no detectors), the reference's published cube60 benchmark
    (the same cube with its 4 detectors, 1e8 photons) and the layered skin model with a vessel
    (200^3 voxels of 5 um, Fresnel on, disk source r=60 voxels, 1e8 photons, energy output; its
    pilot counts about 450 global deposits per photon, so its session sums them through the
    deposit log instead of float atomics)."""
    import umcx_amd as U
    out = {}
    cube = U.config_from_argv(["--bench", "cube60", "-d", "0", "-n", "1000000"])
    # the reference's published benchmark: cube60 (4 detectors on) with 1e8 photons
    cube_pub = U.config_from_argv(["--bench", "cube60", "-n", "1e8"])
    for name, cfg, steps in (("cube60_1e6", cube, 5), ("cube60_published_1e8", cube_pub, 1),
                             ("skinvessel_200cube_1e8", SKIN_CONFIG, 1)):
        with U.Simulation(cfg) as sim:
            n = int(float(cfg["Session"]["Photons"]))
            sim.run()
            ms = []
            for _ in range(steps):
                sim.reset()
                ms.append(sim.run())
            kms = sum(ms) / len(ms)
            esc, tot, dep = sim.energies()
            out[name] = {"photons": n, "kernel_ms": round(kms, 3), "photons_per_ms": round(n / kms, 1),
                         "lds_label_bits": sim.info["lds_label_bits"], "tally_box": sim.tally_box(),
                         "deposit_cache": sim.deposit_cache(), "deposit_cache_pilot_hits": round(sim.deposit_cache_hits(), 4),
                         "deposit_log": sim.deposit_log(),
                         "global_deposits_per_photon": round(sim.global_deposits_per_photon(), 1),
                         "absorbed_percent": (tot - esc) / tot * 100.0}
            if name == "cube60_published_1e8":
                # webpage/umcx.html:451,455 of the reference: 175,075 photons/ms on an RTX 5090
                out[name]["published_rtx5090_photons_per_ms"] = 175075
                out[name]["vs_published_rtx5090"] = round(n / kms / 175075, 3)
        log(f"bench: {name}: {out[name]['photons_per_ms']} photons/ms")
    return out


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
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--no-jacobian", action="store_true", help="skip the Jacobian-vs-plain leg")
    ap.add_argument("--no-others", action="store_true", help="skip the other configurations")
    ap.add_argument("--batch-profile", nargs="?", const="batch", choices=["batch", "plain"],
                    help="only build the 256-job batch (autotuned) and run it once, or ('plain') run the loop of "
                         "plain launches over the same jobs once (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(photon ranges would shard across GPUs with one fluence all-reduce)", file=sys.stderr)
        sys.exit(2)

    import umcx_amd as U
    try:
        import torch
        sync = torch.cuda.synchronize
    except Exception:
        torch, sync = None, (lambda: None)
    if args.batch_profile:
        from umcx_amd import batch as B
        jobs = B.uneven_jobs(seed=7, njobs=256)
        photons = sum(int(j["Session"]["Photons"]) for j in jobs)
        if args.batch_profile == "plain":
            _, kms, wall = B.plain_loop(jobs, keep=False)
            print(json.dumps({"photons": photons, "plain_loop_kernel_ms": kms, "plain_loop_wall_ms": wall}))
            return
        with B.BatchSimulation(jobs) as bs:
            ms = bs.run()
        print(json.dumps({"batched_photons": photons, "batch_ms": ms}))
        return

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
    sim_cache = (sim.deposit_cache(), sim.deposit_cache_hits())
    sim.close()
    # the drop-in path end to end (outside the timed steps): a fresh session uploads the labels,
    # runs the pilot and the photons, and reads the fluence and the detected rows back to the host
    sync()
    t1 = time.perf_counter()
    with U.Simulation(cfg) as e2e:
        e2e.run()
        e2e_flu = e2e.fluence()
        e2e_rows = e2e.detp()[0]
    e2e_ms = (time.perf_counter() - t1) * 1e3
    log(f"bench: end to end {e2e_ms:.1f} ms ({e2e_flu.nbytes / 1e6:.1f} MB fluence, {len(e2e_rows)} rows read back)")
    ms_per_step = wall_ms / args.steps
    value = args.photons * args.steps / wall_ms
    kavg = sum(kernel_ms) / len(kernel_ms)
    achieved_tflops = OPS_PER_PHOTON * args.photons / (kavg * 1e-3) / 1e12
    # the kernel's per-photon counters come from its committed profile summary (build_match says
    # whether they were measured on the sources built now)
    pc = counter_summary(PHOTON_PROFILE)
    ceil = ceilings(pc, args.photons, kavg * 1e-3)
    roofline = {"bound": "valu" if ceil["binding"] == "valu_issue" else "atomics",
                "achieved": round(achieved_tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 5),
                "traffic": round(pc["hbm_bytes_per_photon"] * args.photons),
                "traffic_gbps": round(pc["hbm_bytes_per_photon"] * args.photons / (kavg * 1e-3) / 1e9, 1),
                "ops_per_photon": round(OPS_PER_PHOTON, 1), "kernel_ms": round(kavg, 3), **ceil}
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "photons/ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "heterogeneous reflective: 60^3 1 mm voxels, sphere r=15 (n=1.0) in n=1.37 medium, "
                               "Fresnel on, pencil (29.5,29.5,0)+z, 50 gates x 0.1 ns, 4 disk detectors r=1",
                   "photons_per_step": args.photons, "seed": 1648335518},
        "roofline": roofline,
        "end_to_end": {"wall_ms": round(e2e_ms, 1), "photons_per_ms": round(args.photons / e2e_ms, 1),
                       "includes": "session set-up, label upload, counting pilot, photon loop, read-back of the "
                                   "fluence grid and the detected rows over PCIe"},
        "tally_box": sim_box,
        "deposit_cache": {"on": sim_cache[0], "pilot_hits": round(sim_cache[1], 4)},
        "detected_per_photon": detected / args.photons,
        "absorbed_percent": (tot - esc) / tot * 100.0,
    }
    if not args.no_others:
        out["other_configs"] = other_configs_leg(sync)
    if not args.no_jacobian:
        log("bench: Jacobian leg")
        out["jacobian"] = jacobian_leg(sync)
    if not args.no_batch:
        log("bench: batched leg")
        out["batched"] = batched_leg(sync)
    if not args.no_cpu:
        log("bench: CPU baseline")
        try:
            cb = cpu_baseline()
            ratios = {"hetero_reflective_1e8": round(value / cb["value"], 1)}
            for name, v in cb.get("other_configs", {}).items():
                gpu = out.get("other_configs", {}).get(name)
                if gpu:
                    ratios[name] = round(gpu["photons_per_ms"] / v["value"], 1)
            cb["gpu_over_cpu"] = ratios
            out["cpu_baseline"] = cb
        except Exception as e:  # report, never fake
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
