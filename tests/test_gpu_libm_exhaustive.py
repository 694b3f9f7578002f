"""The GPU replicas of the host C library's logf, sinf, cosf, acosf and expf, compared bit for bit
with the host library itself (glibc, called through the oracle's helper) on EVERY float of the
ranges the photon loop can reach, and beyond:

* logf: every positive normal float (the loop takes logf(u + FLT_EPSILON), u in [0, 1),
  src/umcx.cpp:213);
* expf: every float from -inf to +inf (the loop takes expf(-mua d) <= 0, :326);
* acosf: every float of [-1, 1] (Henyey-Greenstein cosines for any g, 2u - 1, cone angles:
  :232, :372, :374);
* sinf and cosf: every float (2 pi u and scattering angles, :437-438, stay below 120, where the
  replicas reduce in double; user-set cone and fourier angles, :231, :252, may reach any value,
  where the reduction uses the bits of 2/pi in integer arithmetic).

Results compare as bit patterns; two NaNs count as equal (the two processors' default NaNs differ
in sign). Inputs are streamed in chunks of 2^27 floats.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
import umcx_amd as U

pytestmark = pytest.mark.gpu

CHUNK = 1 << 27


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mismatches(fn: int, lo: int, hi: int) -> tuple[int, int, list]:
    """Compare device replica and host library on every bit pattern in [lo, hi]."""
    bad, tested, examples = 0, 0, []
    for start in range(lo, hi + 1, CHUNK):
        stop = min(hi + 1, start + CHUNK)
        x = np.arange(start, stop, dtype=np.uint64).astype(np.uint32).view(np.float32)
        dev = U.libm_eval(fn, x, device=0)
        host = O.host_libm(fn, x)
        diff = dev.view(np.uint32) != host.view(np.uint32)
        if diff.any():
            diff &= ~(np.isnan(dev) & np.isnan(host))
            n = int(diff.sum())
            if n:
                bad += n
                idx = np.flatnonzero(diff)[:3]
                examples += [(float(x[i]), float(dev[i]), float(host[i])) for i in idx]
        tested += stop - start
    return bad, tested, examples


def run_ranges(name, fn, ranges):
    total_bad, total = 0, 0
    for lo, hi in ranges:
        say(f"libm exhaustive {name}: bits {lo:#010x}..{hi:#010x}")
        bad, tested, ex = mismatches(fn, lo, hi)
        total_bad += bad
        total += tested
        assert bad == 0, (name, hex(lo), hex(hi), bad, ex)
    say(f"libm exhaustive {name}: {total} floats, {total_bad} mismatches")
    return total


def test_logf_every_positive_normal_float():
    assert run_ranges("logf", 0, [(0x00800000, 0x7F7FFFFF)]) == 0x7F7FFFFF - 0x00800000 + 1


def test_expf_every_float():
    # -0 .. -inf and +0 .. +inf (NaN inputs excluded)
    run_ranges("expf", 4, [(0x80000000, 0xFF800000), (0x00000000, 0x7F800000)])


def test_acosf_every_float_of_the_unit_interval():
    run_ranges("acosf", 3, [(0x00000000, 0x3F800000), (0x80000000, 0xBF800000)])


@pytest.mark.parametrize("fn,name", [(1, "sinf"), (2, "cosf")])
def test_sin_cos_every_float(fn, name):
    # every finite float and the infinities (NaN inputs excluded)
    run_ranges(name, fn, [(0x00000000, 0x7F800000), (0x80000000, 0xFF800000)])
