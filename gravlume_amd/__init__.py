"""MI355X (gfx950) Kerr-Newman null-geodesic trace.

The trace runs in hand-written HIP kernels (gravlume_amd/csrc) behind the C
ABI of include/gravlume.h; this package is the host-side mirror of the
reference's trace seams. Importing it loads libgravlume_hip.so and fails if
the library has not been built: there is no CPU fallback.
"""
# PyTorch first: its HIP runtime is the process's runtime, and the library
# binds to the libamdhip64 already loaded. Loaded the other way round, the
# library's runtime can find no device on a host where PyTorch does.
import torch  # noqa: F401,E402

from . import _capi  # noqa: E402
from ._capi import GravlumeError  # noqa: E402

_capi.load()

from .scene import Scene, escape_map, isco_prograde, outer_horizon  # noqa: E402
from .trace import (FrameGraph, FrequencyRatioImage, GeodesicBatch,  # noqa: E402
                    GpuTraceInputError, TileRegion, TracePipeline, TraceTarget, VjpTrace,
                    flop_constants)

__all__ = ["FrameGraph", "FrequencyRatioImage", "GeodesicBatch", "VjpTrace", "GpuTraceInputError", "GravlumeError", "Scene",
           "TileRegion", "TracePipeline", "TraceTarget", "escape_map", "flop_constants",
           "isco_prograde", "outer_horizon"]
