#!/usr/bin/env python3
"""Write the golden vectors the parity tests compare against.

Runs the reference umcx binary, compiled from /root/reference/src by ``make -C oracle ref``
(output oracle/_ref/umcx), on a fixed list of small cases and stores what it produced:

* the printed summary line (energy, normalizer, detected count, absorbed %),
* every detected-photon row from ``<ID>_detp.jdb`` (sorted, bitwise),
* the fluence volume from ``<ID>.bnii`` -- in full for small grids, otherwise per-gate sums,
  the sum of squares and a fixed sample of voxels,
* the expanded JSON config (``--dumpjson``) and the label volume with detector bits
  (``--dumpmask``),
* the seed words of a few seeds (glibc srand/rand).

Run in the build container only (the reference is absent on GPU machines):
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile
import zlib
import base64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from umcx_amd import bjdata  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "umcx")
LINE = re.compile(r"simulated energy ([0-9.eE+-]+), speed ([0-9.eE+-]+|inf) photon/ms, duration ([0-9.eE+-]+) ms, "
                  r"normalizer ([0-9.eE+-]+|inf|nan), detected (\d+), absorbed ([0-9.eE+-]+|nan|-nan)%")

C2 = {"Session": {"ID": "c2", "Photons": 1000000, "RNGSeed": 1648335518, "DoMismatch": 1, "DoSaveVolume": 1,
                  "DoPartialPath": 1, "SaveDetFlag": 5, "OutputType": "x"},
      "Forward": {"T0": 0, "T1": 5e-9, "Dt": 1e-10},
      "Domain": {"Dim": [60, 60, 60], "LengthUnit": 1,
                 "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.005, "mus": 1.0, "g": 0.01, "n": 1.37},
                           {"mua": 0.002, "mus": 5.0, "g": 0.9, "n": 1.0}]},
      "Optode": {"Source": {"Type": "pencil", "Pos": [29.5, 29.5, 0], "Dir": [0, 0, 1]},
                 "Detector": [{"Pos": [20, 29, 0], "R": 1}, {"Pos": [40, 29, 0], "R": 1},
                              {"Pos": [29, 20, 0], "R": 1}, {"Pos": [29, 40, 0], "R": 1}]},
      "Shapes": [{"Grid": {"Tag": 1, "Size": [60, 60, 60]}}, {"Sphere": {"O": [30, 30, 30], "R": 15, "Tag": 2}}]}

C3 = {"Session": {"ID": "c3", "Photons": 1000000, "RNGSeed": 1648335518, "DoMismatch": 1, "DoSaveVolume": 1,
                  "DoPartialPath": 0, "OutputType": "e"},
      "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
      "Domain": {"Dim": [200, 200, 200], "LengthUnit": 0.005,
                 "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 1e-5, "mus": 0, "g": 1, "n": 1},
                           {"mua": 1.657, "mus": 37.59, "g": 0.9, "n": 1.37},
                           {"mua": 0.0458, "mus": 35.65, "g": 0.9, "n": 1.37},
                           {"mua": 23.05, "mus": 9.398, "g": 0.9, "n": 1.37}]},
      "Optode": {"Source": {"Type": "disk", "Pos": [100, 100, 0], "Dir": [0, 0, 1], "Param1": [60, 0, 0, 0]}},
      "Shapes": [{"Grid": {"Tag": 3, "Size": [200, 200, 200]}}, {"ZLayers": [[0, 20, 1], [20, 32, 2]]},
                 {"Cylinder": {"Tag": 4, "C0": [0, 100, 100], "C1": [200, 100, 100], "R": 20}}]}


def slab_volume(seed=11):
    """64x64x32 slab of label 1 with one 6x6x6 inclusion of label 2 at a seeded position."""
    rng = np.random.default_rng(seed)
    vol = np.ones((64, 64, 32), dtype=np.int32)
    x0, y0, z0 = (int(rng.integers(8, 64 - 14)), int(rng.integers(8, 64 - 14)), int(rng.integers(4, 32 - 10)))
    vol[x0:x0 + 6, y0:y0 + 6, z0:z0 + 6] = 2
    return vol, (x0, y0, z0)


def c4_single_source():
    vol, corner = slab_volume()
    cfg = {"Session": {"ID": "c4", "Photons": 1000000, "RNGSeed": 1648335518, "DoMismatch": 1,
                       "DoPartialPath": 1, "SaveDetFlag": 5, "OutputType": "x"},
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 5e-9},
           "Domain": {"Dim": [64, 64, 32], "LengthUnit": 1,
                      "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.01, "mus": 1.0, "g": 0.9, "n": 1.37},
                                {"mua": 0.035, "mus": 1.0, "g": 0.9, "n": 1.37}]},
           "Optode": {"Source": {"Type": "pencil", "Pos": [24.5, 24.5, 0], "Dir": [0, 0, 1]},
                      "Detector": [{"Pos": [20 + 8 * i, 20 + 8 * j, 0], "R": 1.5} for j in range(4) for i in range(4)]},
           "Shapes": {"_ArraySize_": [64, 64, 32], "_ArrayType_": "int32",
                      # JData column order: umcx copies this flat list straight into its [x][y][z] buffer
                      "_ArrayData_": vol.reshape(-1).tolist()}}
    return cfg


def zip_volume_case():
    vol = np.ones((40, 36, 30), dtype=np.uint8)
    vol[10:25, 8:20, 5:18] = 2
    raw = vol.reshape(-1).tobytes()
    cfg = {"Session": {"ID": "zipvol", "Photons": 1000, "RNGSeed": 7, "DoMismatch": 1, "SaveDetFlag": 53},
           "Forward": {"T0": 0, "T1": 5e-9, "Dt": 1e-9},
           "Domain": {"Dim": [40, 36, 30], "LengthUnit": 0.5,
                      "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 0.01, "mus": 2.0, "g": 0.8, "n": 1.4},
                                {"mua": 0.05, "mus": 8.0, "g": 0.95, "n": 1.33}]},
           "Optode": {"Source": {"Type": "pencil", "Pos": [20, 18, 0], "Dir": [0, 0, 1]},
                      "Detector": [{"Pos": [15, 18, 0], "R": 2}, {"Pos": [25, 18, 0], "R": 2}]},
           "Shapes": {"_ArraySize_": [40, 36, 30], "_ArrayType_": "uint8", "_ArrayZipType_": "zlib",
                      "_ArrayZipSize_": [1, int(vol.size)],
                      "_ArrayZipData_": base64.b64encode(zlib.compress(raw)).decode("ascii")}}
    return cfg


# the longest domain axis the reference's short voxel coordinates allow (32767 voxels): a 3x3 bar
# lit along x, nearly unscattering, so photons run its whole length, reflecting off its sides
LONGX = {"Session": {"ID": "longx", "Photons": 2000, "RNGSeed": 4242, "DoMismatch": 1, "DoPartialPath": 0},
         "Forward": {"T0": 0, "T1": 2e-7, "Dt": 2e-8},
         "Domain": {"Dim": [32767, 3, 3], "LengthUnit": 1,
                    "Media": [{"mua": 0, "mus": 0, "g": 1, "n": 1}, {"mua": 1e-5, "mus": 1e-4, "g": 0.9, "n": 1.37}]},
         "Optode": {"Source": {"Type": "pencil", "Pos": [0.5, 1.5, 1.5], "Dir": [1, 0, 0]}},
         "Shapes": [{"Grid": {"Tag": 1, "Size": [32767, 3, 3]}}]}

CASES = {
    # name: (argv after the binary, json file contents or None)
    "c1": (["--bench", "cube60", "-d", "0", "-n", "20000", "-U", "0"], None),
    "cube60": (["--bench", "cube60", "-n", "20000", "-w", "53", "-U", "0"], None),
    "cube60b": (["--bench", "cube60b", "-n", "20000", "-w", "53", "-U", "0"], None),
    "cubesph60b": (["--bench", "cubesph60b", "-n", "20000", "-w", "53"], None),
    "cube60planar": (["--bench", "cube60planar", "-n", "20000", "-w", "53", "-U", "0"], None),
    "sphshells": (["--bench", "sphshells", "-n", "10000", "-w", "53", "-U", "0", "-O", "f"], None),
    "spherebox": (["--bench", "spherebox", "-n", "20000", "-w", "53", "-U", "0"], None),
    "skinvessel": (["--bench", "skinvessel", "-n", "5000", "-U", "0"], None),
    "c2": (["-f", "INPUT.json", "-n", "20000", "-w", "53", "-U", "0"], C2),
    "c3": (["-f", "INPUT.json", "-n", "4000", "-U", "0"], C3),
    "c4src": (["-f", "INPUT.json", "-n", "20000", "-U", "0"], "c4"),
    "zipvol": (["-f", "INPUT.json", "-n", "20000", "-U", "0"], "zip"),
    "isotropic": (["--bench", "cube60", "--json", '{"Optode":{"Source":{"Type":"isotropic","Pos":[29,29,29]}}}',
                   "-d", "0", "-n", "20000", "-U", "0"], None),
    "cone": (["--bench", "cube60", "--json", '{"Domain":{"Media":[[0,0,1,1],[0.001,0.001,0,1]]},'
              '"Optode":{"Source":{"Type":"cone","Param1":[0.5,0,0,0]}}}', "-d", "0", "-n", "20000", "-U", "0"], None),
    "fourier": (["--bench", "cube60planar", "--json", '{"Optode":{"Source":{"Type":"fourier","Param1":[40,0,0,2.25],'
                 '"Param2":[0,40,0,1.5]}}}', "-n", "20000", "-w", "53", "-U", "0"], None),
    "save2pt0": (["--bench", "cube60b", "-S", "0", "-n", "10000", "-w", "5"], None),
    "overflow": (["--bench", "cube60b", "-n", "20000", "-H", "10", "-w", "5", "-U", "0"], None),
    "gated": (["--bench", "spherebox", "--json", '{"Forward":{"T0":0,"T1":1e-9,"Dt":1e-10}}',
               "-n", "20000", "-w", "53", "-U", "0", "-O", "e"], None),
    "normx": (["--bench", "cubesph60b", "-n", "5000", "-d", "0", "-u", "0.5"], None),
    "zerophotons": (["--bench", "cube60", "-n", "0"], None),
    "longx": (["-f", "INPUT.json", "-U", "0"], LONGX),
}
FULL_FLUENCE_MAX = 300_000   # store the whole volume when it has at most this many cells
SAMPLE = 4096


def run_ref(argv, cwd):
    p = subprocess.run([REF] + argv, cwd=cwd, capture_output=True, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="8"))
    if p.returncode != 0:
        raise RuntimeError(f"reference failed: {argv}\n{p.stdout}\n{p.stderr}")
    return p.stdout


def build_case(name, argv, cfg):
    with tempfile.TemporaryDirectory() as td:
        if cfg is not None:
            if cfg == "c4":
                cfg = c4_single_source()
            elif cfg == "zip":
                cfg = zip_volume_case()
            with open(os.path.join(td, "INPUT.json"), "w") as f:
                json.dump(cfg, f)
        argv = list(argv) + ["-s", name]
        dump = run_ref(argv + ["--dumpjson"], td)
        run_ref(argv + ["--dumpmask"], td)
        maskfile = os.path.join(td, f"{name}_vol.bnii")
        mask = bjdata.load_bnii(maskfile)[..., 0].astype(np.int32)
        with open(maskfile, "rb") as f:
            mask_head = np.frombuffer(f.read(128), dtype=np.uint8)
        mask_size = os.path.getsize(maskfile)
        out = run_ref(argv, td)
        m = LINE.search(out)
        if not m:
            raise RuntimeError(f"no summary line for {name}: {out}")
        rec = dict(name=name, argv=json.dumps(argv), dumpjson=dump,
                   energy=float(m.group(1)), normalizer=float(m.group(4)), detected=int(m.group(5)),
                   absorbed=float(m.group(6)) if "nan" not in m.group(6) else np.nan, line=out.strip())
        rec["mask"] = mask
        rec["mask_head"] = mask_head
        rec["mask_size"] = mask_size
        bnii = os.path.join(td, f"{name}.bnii")
        if os.path.exists(bnii):
            flu = bjdata.load_bnii(bnii)          # [x,y,z,t]
            flat = flu.transpose(3, 0, 1, 2).reshape(-1).astype(np.float32)   # memory order [t][x][y][z]
            rec["fluence_shape"] = np.array(flu.shape, dtype=np.int64)
            rec["gate_sums"] = flu.astype(np.float64).sum(axis=(0, 1, 2))
            rec["sumsq"] = float((flat.astype(np.float64) ** 2).sum())
            if flat.size <= FULL_FLUENCE_MAX:
                rec["fluence"] = flat
            else:
                nz = np.flatnonzero(flat)
                top = nz[np.argsort(flat[nz])[-SAMPLE:]]
                rng = np.random.default_rng(1234)
                rnd = rng.choice(nz, size=min(SAMPLE, nz.size), replace=False)
                idx = np.unique(np.concatenate([top, rnd]))
                rec["sample_idx"] = idx.astype(np.int64)
                rec["sample_val"] = flat[idx]
            with open(bnii, "rb") as f:
                head = f.read(128)
            rec["bnii_head"] = np.frombuffer(head, dtype=np.uint8)
            rec["bnii_size"] = os.path.getsize(bnii)
        detp = os.path.join(td, f"{name}_detp.jdb")
        if os.path.exists(detp):
            info, rows = bjdata.load_detp(detp)
            rows = rows[np.lexsort(rows.T[::-1])] if rows.size else rows
            rec["detp"] = rows.astype(np.float32)
            rec["detp_info"] = json.dumps(info)
            with open(detp, "rb") as f:
                raw = f.read()
            # header bytes up to the raw row data: pins the writer's layout
            rec["detp_head"] = np.frombuffer(raw[:raw.find(b"PhotonRawData") + 40], dtype=np.uint8)
            rec["detp_size"] = len(raw)
        return rec


def main():
    if not os.path.exists(REF):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    names = sys.argv[1:] or list(CASES)
    for name in names:
        argv, cfg = CASES[name]
        rec = build_case(name, argv, cfg)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **rec)
        print(f"{name}: {rec['line']}")
    seeds = {}
    for s in (1648335518, 7, 1, 123456, 2147483647):
        code = f"#include <stdio.h>\n#include <stdlib.h>\nint main(){{srand({s});int a=rand(),b=rand(),c=rand(),d=rand();printf(\"%d %d %d %d\",a,b,c,d);}}"
        with tempfile.TemporaryDirectory() as td:
            src = os.path.join(td, "s.c")
            with open(src, "w") as f:
                f.write(code)
            subprocess.run(["gcc", src, "-o", os.path.join(td, "s")], check=True)
            seeds[str(s)] = [int(v) for v in subprocess.run([os.path.join(td, "s")], capture_output=True,
                                                             text=True).stdout.split()]
    with open(os.path.join(HERE, "seed_words.json"), "w") as f:
        json.dump(seeds, f, indent=1)


if __name__ == "__main__":
    main()
