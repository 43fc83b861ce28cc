"""MI355X-native Kobayashi (1993) phase-field solver.

The numerics run in hand-written gfx950 HIP kernels behind the C ABI of
include/kobayashi_gpu.h (libkobayashi_gpu.so, built in-tree by
`python -m crystalgrowth_amd.build`).  This package is the Python host side:
`Solver` (general handle) and `Kobayashi` (the reference class's surface).
"""
from ._lib import KGError, Params, Stats  # noqa: F401
from .kobayashi import (Batch, Kobayashi, Solver, default_params, device_count,  # noqa: F401
                        philox4x32_10, validate)

__all__ = ["Batch", "Kobayashi", "Solver", "Params", "Stats", "KGError", "default_params",
           "device_count", "philox4x32_10", "validate"]
