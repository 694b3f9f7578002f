"""Build the native parts in-tree: libumcx.so + the umcx CLI (umcx_amd/csrc/Makefile) and the
CPU oracle (oracle/Makefile; plus the reference binary when /root/reference is present).

    python -m umcx_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def build(oracle: bool = True, jobs: int = 8) -> None:
    # always recompile: the build check must not trust stale objects
    subprocess.run(["make", "-C", os.path.join(PKG, "csrc"), "clean"], check=True, capture_output=True)
    subprocess.run(["make", "-C", os.path.join(PKG, "csrc"), f"-j{jobs}"], check=True)
    if oracle:
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-B"], check=True)
        if os.path.isdir("/root/reference/src"):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)


if __name__ == "__main__":
    build(oracle="--no-oracle" not in sys.argv)
