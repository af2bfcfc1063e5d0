"""storm_amd: MI355X-native Implicit Monte Carlo transport (gray, Cartesian).

The compute path is libstorm_amd.so (HIP kernels for gfx950 behind the C ABI
in include/storm_gpu.h).  This package holds the Python mirror of the
problem description and a thin handle over the C ABI.
"""
from . import capi, problems  # noqa: F401
from .engine import GpuEngine, StormError, device_info  # noqa: F401

__all__ = ["capi", "problems", "GpuEngine", "StormError", "device_info"]
