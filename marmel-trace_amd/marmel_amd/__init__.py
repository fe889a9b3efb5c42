"""marmel_amd -- MI355X (gfx950) path-tracing hot path of marmel-trace.

The GPU library (libmarmel_hip) and the C11 host (libmarmel_host) are native;
this package only binds them for tests, benchmarks and Python callers.
"""
from . import abi, native
from .scene import (Batch, GpuScene, HostScene, MarmelError, device_info, encode_bmp, flare_apply,
                    last_sample_stats, make_camera, write_bmp)

__all__ = ["abi", "native", "Batch", "GpuScene", "HostScene", "MarmelError", "device_info",
           "encode_bmp", "flare_apply", "last_sample_stats", "make_camera", "write_bmp"]
