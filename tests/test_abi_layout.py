"""The ctypes structures of umcx_amd/_native.py have the layout of include/umcx.h: a C program
built against the header prints the size of each struct and the offset of every field, and each
must equal the ctypes one (a field added to the header but not to the binding, or the other way
round, shifts every field after it for Python callers)."""
import ctypes as C
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from umcx_amd._native import UmcxBatchTrial, UmcxDeviceProblem, UmcxInfo, UmcxLaunch  # noqa: E402

STRUCTS = {"umcx_launch": UmcxLaunch, "umcx_info": UmcxInfo, "umcx_batch_trial": UmcxBatchTrial,
           "umcx_device_problem": UmcxDeviceProblem}


def c_layout(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "umcx.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        name, field, value = line.split()
        layout[(name, field)] = int(value)
    return layout


def test_ctypes_structs_match_the_header(tmp_path):
    layout = c_layout(tmp_path)
    for cname, cls in STRUCTS.items():
        # equal sizes: no header field lacks a binding at the end; equal offsets: none in between
        assert C.sizeof(cls) == layout[(cname, "sizeof")], cname
        for field, _ in cls._fields_:
            assert getattr(cls, field).offset == layout[(cname, field)], f"{cname}.{field}"
