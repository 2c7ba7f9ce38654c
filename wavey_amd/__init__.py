"""MI355X-native FDTD wave solver for Wavey (core/wavey-solver FDTD path).

The compute path is the HIP library ``libwavey_fdtd.so`` (C ABI in
``include/wavey_fdtd.h``); this package is its Python view plus seeded
workload generators. Nothing here computes a solve on the CPU.
"""
from . import configs, fdtd  # noqa: F401
from .fdtd import FdtdError, FdtdGrid, Gpu, Plan  # noqa: F401

__all__ = ["configs", "fdtd", "FdtdGrid", "Gpu", "Plan", "FdtdError"]
