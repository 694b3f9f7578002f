"""The reference's own command-line test windows (test/testumcx.sh:13-131) run against this
build's `umcx` binary on the GPU: the same command lines, and the printed absorbed fraction or
detected count must fall in the same windows."""
import os
import re
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UMCX = os.path.join(ROOT, "umcx_amd", "bin", "umcx")

# (description, arguments, regex that the output must match) -- test/testumcx.sh:59-131
CASES = [
    ("homogeneous cube60", ["--bench", "cube60", "-S", "0"], r"absorbed.*17\.[0-9]+%"),
    ("boundary reflection cube60b", ["--bench", "cube60b", "-S", "0"], r"absorbed.*27\.[0-9]+%"),
    ("reflection flag -b 1", ["--bench", "cube60", "-b", "1", "-S", "0"], r"absorbed.*27\.[0-9]+%"),
    ("photon detection cube60b", ["--bench", "cube60b"], r"detected.*4[0-9]+"),
    ("planar source cube60planar", ["--bench", "cube60planar"], r"absorbed.*25\.[0-9]+%"),
    ("isotropic source", ["--bench", "cube60", "--json",
                          '{"Optode":{"Source":{"Type":"isotropic","Pos":[29,29,29]}}}', "-d", "0", "-S", "0"],
     r"absorbed.*88\.[0-9]+%"),
    ("cone beam source", ["--bench", "cube60", "--json",
                          '{"Domain":{"Media":[[0,0,1,1],[0.001,0.001,0,1]]},"Optode":{"Source":{"Type":"cone",'
                          '"Param1":[0.5,0,0,0]}}}', "-d", "0", "-S", "0"], r"absorbed.*6\.[0-9]+%"),
    ("unitinmm skinvessel", ["--bench", "skinvessel", "-S", "0", "-n", "1e5"], r"absorbed.*39\.[0-9]+%"),
    ("heterogeneous spherebox", ["--bench", "spherebox", "-S", "0"], r"absorbed.*1[01]\.[0-9]+%"),
]


@pytest.mark.parametrize("desc,args,pattern", CASES, ids=[c[0] for c in CASES])
def test_reference_cli_window(desc, args, pattern, tmp_path):
    print(f"reference CLI window: {desc}", flush=True)
    sys.stdout.flush()
    out = subprocess.run([UMCX] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert re.search(pattern, out.stdout), out.stdout
