"""toaster_amd -- MI355X-native path-tracing hot path of the Toaster renderer.

The compute path is libtoaster_hip.so (hand-written HIP for gfx950 behind the C ABI in
include/toaster_hip.h); this package is its Python host mirror of toaster-gpu's
renderer interface.  Importing the package does not touch the GPU.
"""
from . import _abi  # noqa: F401
from .scene import EnvironmentMap, PackedScene, Scene, pack  # noqa: F401
from .renderer import (GpuContext, GpuScene, RenderOptions, device_count, display_rgba8,  # noqa: F401
                       render)

__all__ = ["Scene", "EnvironmentMap", "PackedScene", "pack", "GpuContext", "GpuScene",
           "RenderOptions", "device_count", "display_rgba8", "render"]
