"""MI355X coupled IBM-DEM time step: Python host side.

The compute path is the HIP library ``lib/libibmdem.so`` built from
``csrc/`` (see ``build.py``). ``gpu_library()`` opens exactly that file and
raises if it is missing: there is no CPU fallback in the product.
"""
from __future__ import annotations

import os

from .abi import (BC_PERIODIC, BC_WALL, EXPORTED, FP64, MIXED, Batch,
                  BodyArrays, Case, IbmdemError, MarkerArrays, Simulation,
                  SolveInfo, load_library)

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_LIB = os.path.join(PKG_ROOT, "lib", "libibmdem.so")

_lib = None


def gpu_library():
    """The HIP library (loaded once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(GPU_LIB):
            raise RuntimeError(
                f"HIP library {GPU_LIB} is missing: run __graft_entry__.build() "
                "(or python ibm-dem_amd/build.py) first; there is no CPU fallback")
        _lib = load_library(GPU_LIB)
    return _lib


def simulation(case: Case) -> Simulation:
    """A context on the GPU library."""
    return Simulation(gpu_library(), case)


def batch(cases) -> Batch:
    """A batch of independent cases on the GPU library."""
    return Batch(gpu_library(), cases)


__all__ = ["BC_PERIODIC", "BC_WALL", "FP64", "MIXED", "Batch", "BodyArrays", "Case", "batch",
           "IbmdemError", "MarkerArrays", "Simulation", "SolveInfo",
           "EXPORTED", "gpu_library", "simulation", "GPU_LIB"]
