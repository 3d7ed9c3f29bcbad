"""EDUS propagation engine for MI355X: Python harness over the C ABI.

The product is the C ABI library (include/edus_gpu.h, lib/libedus_gpu.so);
this package only marshals arrays and parameters for tests and benchmarks.
"""
from . import config  # noqa: F401
from .engine import (PREC_FP64, PREC_MIXED, LAYOUT_FULL, LAYOUT_HERMITIAN, EdusError, Engine, Batch, Mlp, Simulation, Tuning, lib,  # noqa: F401
                     LIB_PATH, HEADER_PATH, EXPORTS, kgrad_stencil, lindhard, lindhard_device,
                     lindhard_device_noalloc)
