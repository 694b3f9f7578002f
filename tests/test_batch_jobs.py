"""The batched workload generator (no GPU): deterministic from seed 7, parameters in the stated
ranges, and the zlib+base64 JData volumes decode to the generated labels through the library's
own input parser."""
import base64
import zlib

import numpy as np

import umcx_amd as U
from umcx_amd import batch as B


def test_uneven_jobs_ranges_and_determinism():
    a = B.uneven_jobs(seed=7, njobs=256)
    b = B.uneven_jobs(seed=7, njobs=256)
    assert a == b and len(a) == 256
    photons = np.array([j["Session"]["Photons"] for j in a])
    assert photons.min() >= 1e5 and photons.max() <= 1e7
    assert 1.5e6 < photons.mean() < 3.0e6            # log-uniform on [1e5, 1e7] has mean ~2.15e6
    for j in a:
        assert all(32 <= d <= 128 for d in j["Domain"]["Dim"])
        media = j["Domain"]["Media"][1:]
        assert 2 <= len(media) <= 6
        for m in media:
            assert 0.001 <= m["mua"] <= 0.05 and 0.5 <= m["mus"] <= 20 and 0.7 <= m["g"] <= 0.95
            assert 1.33 <= m["n"] <= 1.45
        assert j["Session"]["DoMismatch"] == 1 and j["Optode"]["Source"]["Type"] == "pencil"


def test_job_volumes_decode_through_the_library():
    for cfg in B.uneven_jobs(seed=7, njobs=4):
        dims = cfg["Domain"]["Dim"]
        raw = zlib.decompress(base64.b64decode(cfg["Shapes"]["_ArrayZipData_"]))
        vol = np.frombuffer(raw, dtype=np.uint8).reshape(dims)
        lab = U.domain(cfg)
        assert np.array_equal(lab, vol.astype(np.int32))
