"""gcdt_amd: MI355X (gfx950) constrained Delaunay triangulation.

The compute path is libgcdt.so (HIP kernels behind the C ABI of
include/gcdt.h).  This package only loads it and moves host arrays in and out;
if the library is missing every call raises -- there is no CPU fallback.
"""
from .lib import GcdtError, lib, load_error  # noqa: F401
from .api import (DeviceBuffer, get_tunable, pack_instances, set_tunable, solve, solve_batch,  # noqa: F401
                  solve_batch_device, solve_carved, solve_device)

__all__ = ["GcdtError", "lib", "load_error", "DeviceBuffer", "solve", "solve_device", "solve_batch",
           "solve_batch_device", "pack_instances", "set_tunable", "get_tunable", "solve_carved"]
