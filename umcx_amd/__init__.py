"""MI355X-native photon-propagation hot path of umcx (voxel Monte Carlo light transport).

The product is libumcx.so (C ABI in include/umcx.h): hand-written gfx950 HIP kernels behind
umcx's JSON/JData contract and its MCX_run_cmd / MCX_run_json exports. This package is its
Python front end.
"""
from .api import (Result, Simulation, UmcxError, config_from_argv, config_from_bjdata, device_count,  # noqa: F401
                  domain, last_error, libm_eval, run_bjdata, run_cmd, run_json, seed_words, simulate)

__version__ = "0.1.0"
