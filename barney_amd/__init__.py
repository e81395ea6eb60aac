"""barney-amd: MI355X-native (gfx950) ray-tracing core for barney.

The compute path is libbnamd.so (hand-written HIP kernels behind the C ABI
declared in include/bnamd.h); this package is the Python host layer over
it. Importing the package does not touch the GPU; api.init() does.
"""
from . import capi  # noqa: F401

__all__ = ["capi", "api", "scenes"]
