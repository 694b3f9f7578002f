#!/usr/bin/env python3
"""Write the fixtures of the BJData input path (MCX_run_simulation with argn < 0: argv[0] is a
BJData config buffer, src/umcx.cpp:475-478).

Every fixture holds bytes and the value the reference's JSON library reads from them: the bytes are
either configs encoded by that library (json::to_bjdata, three size/type settings) or hand-built
buffers covering the rest of the BJData grammar (no-ops, half floats, high-precision numbers,
chars, uncounted and typed containers, N-D arrays of every element type, 1 x n and empty
dimension vectors, malformed input); the expected value is what json::from_bjdata returns, as
JSON text, or "" where it rejects the buffer. Both come from oracle/_ref/bjenc, built from the
reference's vendored src/nlohmann/json.hpp by ``make -C oracle ref``.

Run in the build container only (the reference is absent on GPU machines):
    python tests/golden/make_bjdata.py
"""
from __future__ import annotations

import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BJENC = os.path.join(ROOT, "oracle", "_ref", "bjenc")


def golden_cfg(name: str) -> dict:
    d = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    return json.loads(str(d["dumpjson"]))


def encode(cfg: dict, use_size: int, use_type: int) -> bytes:
    with tempfile.TemporaryDirectory() as t:
        src, dst = os.path.join(t, "in.json"), os.path.join(t, "out.bjd")
        with open(src, "w") as f:
            json.dump(cfg, f)
        subprocess.run([BJENC, src, dst, str(use_size), str(use_type)], check=True)
        return open(dst, "rb").read()


def decode(buf: bytes) -> str:
    with tempfile.TemporaryDirectory() as t:
        src, dst = os.path.join(t, "in.bjd"), os.path.join(t, "out.json")
        open(src, "wb").write(buf)
        r = subprocess.run([BJENC, "-d", src, dst], capture_output=True)
        return open(dst).read() if r.returncode == 0 else ""


def key(s: str) -> bytes:
    return b"U" + bytes([len(s)]) + s.encode()


def grammar_cases() -> dict[str, bytes]:
    c = {}
    c["noops_and_scalars"] = (b"{" + key("a") + b"N" + b"i\xfe" + key("b") + b"U\xc8" + key("c") + b"I" + struct.pack("<h", -300)
                              + key("d") + b"u" + struct.pack("<H", 60000) + key("e") + b"l" + struct.pack("<i", -70000)
                              + key("f") + b"m" + struct.pack("<I", 3000000000) + key("g") + b"L" + struct.pack("<q", -2**40)
                              + key("h") + b"M" + struct.pack("<Q", 2**63 + 5) + key("k") + b"T" + key("l") + b"F"
                              + key("m") + b"Z" + key("n") + b"d" + struct.pack("<f", 0.1) + key("o") + b"D" + struct.pack("<d", -2.5e-300)
                              + b"}")
    c["half_floats"] = b"[" + b"".join(b"h" + struct.pack("<e", v) for v in (1.5, -0.000061035156, 65504.0, 0.0)) + b"]"
    c["high_precision"] = b"[" + b"HU\x0512.75" + b"HU\x031e3" + b"]"
    c["char_and_string"] = b"[" + b"Ca" + b"SU\x05hello" + b"Sl" + struct.pack("<i", 3) + b"xyz" + b"]"
    c["counted_untyped"] = b"[#U\x04" + b"i\x01" + b"SU\x01q" + b"[]" + b"{#U\x01" + key("z") + b"T"
    c["typed_array"] = b"[$d#U\x03" + struct.pack("<3f", 1.0, -2.25, 3.5)
    c["typed_object"] = b"{$l#U\x02" + key("p") + struct.pack("<i", 7) + key("q") + struct.pack("<i", -8)
    c["nd_uint8_2x3"] = b"[$U#[$U#U\x02\x02\x03" + bytes([1, 2, 3, 4, 5, 6])
    c["nd_int32_2x2x2"] = b"[$l#[$i#U\x03\x02\x02\x02" + struct.pack("<8i", *range(-4, 4))
    c["nd_single_3x2"] = b"[$d#[$U#U\x02\x03\x02" + struct.pack("<6f", *[0.5 * i for i in range(6)])
    c["nd_double_untyped_dims"] = b"[$D#[U\x02U\x02]" + struct.pack("<4d", 1.0, 2.0, 3.0, 4.0)
    c["nd_uint16_int16_int64"] = (b"[#U\x03" + b"[$u#[$U#U\x02\x01\x02" + struct.pack("<2H", 1, 65535)
                                  + b"[$I#[$U#U\x02\x02\x01" + struct.pack("<2h", -1, 2)
                                  + b"[$L#[$U#U\x02\x02\x02" + struct.pack("<4q", 1, -2, 3, -4))
    c["nd_char_and_uint32"] = (b"[#U\x02" + b"[$C#[$U#U\x02\x02\x02" + b"abcd"
                               + b"[$m#[$U#U\x02\x02\x01" + struct.pack("<2I", 7, 4000000000))
    c["nd_one_by_n"] = b"[$U#[$U#U\x02\x01\x04" + bytes([9, 8, 7, 6])
    c["nd_one_dim"] = b"[$i#[$U#U\x01\x03" + bytes([1, 2, 255])
    c["nd_zero_dim"] = b"[$U#[$U#U\x02\x00\x05"
    c["nested_uncounted"] = b"{" + key("a") + b"[" + b"{" + key("b") + b"[$U#U\x02\x05\x06" + b"}" + b"]" + b"}"
    # malformed: must be rejected
    c["bad_truncated_int"] = b"l\x01\x02"
    c["bad_truncated_array"] = b"[$d#U\x03" + struct.pack("<2f", 1.0, 2.0)
    c["bad_unknown_marker"] = b"[Q]"
    c["bad_negative_count"] = b"[#i\xff]"
    c["bad_unclosed_object"] = b"{" + key("a") + b"T"
    c["bad_nd_without_type"] = b"[#[$U#U\x02\x02\x02" + b"U\x01U\x02U\x03U\x04"
    c["bad_string_type"] = b"[$S#U\x01U\x01a"
    c["bad_key_marker"] = b"{" + b"T" + b"T" + b"}"
    return c


def main() -> None:
    if not os.path.exists(BJENC):
        sys.exit("oracle/_ref/bjenc is missing: run make -C oracle ref first")
    out = {}
    names = []
    # configs as the reference's JSON library encodes them
    zb = golden_cfg("zipvol")
    vol = np.arange(12 * 10 * 8, dtype=np.int64).reshape(12, 10, 8) % 3
    cube = golden_cfg("cube60b")
    jvol = json.loads(json.dumps(cube))
    jvol["Domain"]["Dim"] = [12, 10, 8]
    jvol["Shapes"] = {"_ArrayType_": "uint8", "_ArraySize_": [12, 10, 8], "_ArrayData_": vol.reshape(-1).tolist()}
    jvol["Optode"]["Source"]["Pos"] = [6, 5, 0]
    jvol["Optode"]["Detector"] = [{"Pos": [3, 5, 0], "R": 1.5}]
    jvol["Session"]["ID"] = "bjvol"
    import base64
    import zlib
    zbytes = golden_cfg("zipvol")
    zarr = json.loads(json.dumps(zbytes))
    zarr["Shapes"]["_ArrayZipData_"] = list(base64.b64decode(zarr["Shapes"]["_ArrayZipData_"]))
    zarr["Session"]["ID"] = "zipbytes"
    configs = {"c2": golden_cfg("c2"), "zipvol": zb, "cube60b": cube, "jdatavol": jvol, "zipbytes": zarr}
    for name, cfg in configs.items():
        for us, ut in ((1, 1), (1, 0), (0, 0)):
            tag = f"cfg_{name}_{us}{ut}"
            buf = encode(cfg, us, ut)
            out[f"bjd_{tag}"] = np.frombuffer(buf, dtype=np.uint8)
            out[f"json_{tag}"] = np.array(decode(buf))
            out[f"src_{tag}"] = np.array(json.dumps(cfg))
            names.append(tag)
    for name, buf in grammar_cases().items():
        tag = f"gram_{name}"
        out[f"bjd_{tag}"] = np.frombuffer(buf, dtype=np.uint8)
        out[f"json_{tag}"] = np.array(decode(buf))
        names.append(tag)
    out["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "bjdata_inputs.npz"), **out)
    for n in names:
        print(n, len(out[f"bjd_{n}"]), "bytes", "rejected" if not str(out[f"json_{n}"]) else "ok")


if __name__ == "__main__":
    main()
