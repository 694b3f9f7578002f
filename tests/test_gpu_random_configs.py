"""Randomised GPU parity: 300 seeded random configurations (volume size, label layout with background
holes, 1-7 media with and without refractive-index steps, zero-scattering media, every source type
inside and outside the volume, reflection on/off, 1-5 time gates, detectors with every saved
column, all output types, length units) run on the GPU through the C ABI and on the oracle; the
per-photon final states must agree bitwise and the detected rows as multisets."""
import sys

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U

pytestmark = pytest.mark.gpu

SOURCES = ["pencil", "isotropic", "cone", "disk", "planar", "fourier"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sorted_rows(rows):
    return rows[np.lexsort(rows.T[::-1])] if rows.size else rows


def random_config(seed: int) -> dict:
    rng = np.random.default_rng(seed)
    dims = [int(v) for v in rng.integers(6, 41, size=3)]
    if seed % 5 == 4:
        dims = [int(v) for v in rng.integers(60, 90, size=3)]           # labels too large for LDS
    nmed = int(rng.integers(1, 8))
    media = [{"mua": 0.0, "mus": 0.0, "g": 1.0, "n": 1.0}]
    for m in range(nmed):
        mus = 0.0 if rng.random() < 0.1 else float(rng.uniform(0.05, 8.0))
        g = 0.0 if rng.random() < 0.1 else float(rng.uniform(-0.9, 0.95))
        n = 1.37 if rng.random() < 0.3 else float(rng.uniform(1.0, 1.6))
        media.append({"mua": float(rng.uniform(0.0, 0.1)), "mus": mus, "g": g, "n": n})
    vol = np.ones(dims, dtype=np.int32)
    for _ in range(int(rng.integers(0, 5))):
        lo = [int(rng.integers(0, d)) for d in dims]
        hi = [int(rng.integers(l + 1, d + 1)) for l, d in zip(lo, dims)]
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = int(rng.integers(0, nmed + 1))   # 0: background hole
    st = SOURCES[int(rng.integers(0, len(SOURCES)))]
    z0 = float(rng.choice([0.0, -2.5, 0.5]))
    pos = [float(rng.uniform(0.2, dims[0] - 0.2)), float(rng.uniform(0.2, dims[1] - 0.2)), z0]
    src = {"Type": st, "Pos": pos, "Dir": [0.0, 0.0, 1.0]}
    if st == "isotropic":
        src["Pos"] = [dims[0] / 2, dims[1] / 2, dims[2] / 2]
    elif st == "cone":
        src["Param1"] = [float(rng.uniform(0.05, 1.2)), 0, 0, 0]
    elif st == "disk":
        src["Param1"] = [float(rng.uniform(0.5, 4.0)), float(rng.uniform(0.0, 0.5)), 0, 0]
    elif st in ("planar", "fourier"):
        src["Pos"] = [1.0, 1.0, z0]
        src["Param1"] = [float(dims[0] - 2), 0.0, 0.0, float(rng.integers(1, 4)) + (0.5 if st == "fourier" else 0)]
        src["Param2"] = [0.0, float(dims[1] - 2), 0.0, float(rng.uniform(0.0, 0.5)) if st == "fourier" else 0]
    dets = [{"Pos": [float(rng.uniform(0, dims[0])), float(rng.uniform(0, dims[1])), 0.0],
             "R": float(rng.uniform(1.0, 4.0))} for _ in range(int(rng.integers(0, 4)))]
    t1 = float(rng.choice([1e-9, 2e-9, 5e-9]))
    gates = int(rng.integers(1, 6))
    cfg = {
        "Session": {"ID": f"rand{seed}", "Photons": 1500, "RNGSeed": int(rng.integers(1, 2 ** 31 - 1)),
                    "DoMismatch": int(rng.random() < 0.6), "DoPartialPath": int(bool(dets)),
                    "SaveDetFlag": int(rng.choice([1, 5, 17, 33, 53])), "OutputType": str(rng.choice(["x", "f", "e"])),
                    "DoNormalize": 0},
        "Forward": {"T0": 0.0, "T1": t1, "Dt": t1 / gates},
        "Domain": {"Dim": dims, "LengthUnit": float(rng.choice([0.5, 1.0, 2.0])), "Media": media},
        "Optode": {"Source": src, "Detector": dets},
        "Shapes": {"_ArraySize_": dims, "_ArrayType_": "int32",
                   "_ArrayData_": vol.reshape(-1).tolist()},
    }
    return cfg


# seed 782: overlapping detectors save more rows (1,625) than there are photons (1,500)
@pytest.mark.parametrize("seed", list(range(300)) + [782])
def test_random_config_matches_oracle(seed):
    cfg = random_config(seed)
    src = cfg["Optode"]["Source"]
    say(f"random config {seed}: dims {cfg['Domain']['Dim']}, {len(cfg['Domain']['Media']) - 1} media, "
        f"{src['Type']} at z={src['Pos'][2]}, reflect {cfg['Session']['DoMismatch']}, "
        f"{len(cfg['Optode']['Detector'])} detectors")
    prob = O.problem_from_config(cfg, U.domain(cfg))
    n = int(cfg["Session"]["Photons"])
    o = O.run(prob, nphoton=n, traces=True, nthreads=8)
    with U.Simulation(cfg, trace=True, label_cache=-1 if seed % 3 == 2 else 0) as sim:
        sim.run(0, n)
        t = sim.trace(n)
        rows = sorted_rows(sim.detp()[0]) if prob["issavedet"] and len(prob["detpos"]) else None
        e = sim.energies()
    ot = o["traces"]
    exact = lambda a: np.ascontiguousarray(a).view(np.uint32)
    for key in ("pos", "vec", "len"):
        assert np.array_equal(exact(t[key]), exact(ot[key])), key
    assert np.array_equal(t["mediaid"], ot["mediaid"])
    assert np.array_equal(t["lastvoxelidx"], ot["lastvoxelidx"])
    assert np.array_equal(exact(t["w0"]), exact(ot["w0"]))
    if rows is not None:
        assert np.array_equal(rows.view(np.uint32), sorted_rows(o["detp"]).view(np.uint32))
    assert e[1] == o["energies"][1]
    assert e[0] == pytest.approx(o["energies"][0], rel=1e-10, abs=1e-12)
