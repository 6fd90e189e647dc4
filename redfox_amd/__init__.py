"""redfox_amd - MI355X-native relational-algebra operators for Red Fox plans.

The compute lives in librfx_hip.so (HIP kernels for gfx950 behind the C-ABI
in include/rfx_ra.h).  This package loads it, manages device buffers and
contexts, and mirrors the reference's operator interface for tests and
benchmarks.  Importing it fails if the library is not built.
"""
from ._lib import LIB_PATH, RfxError
from .rfx import Context, DeviceBuffer, Graph, Stream, Timer, device_count, lib, read_u64, u64_cell

# Load eagerly so a missing or broken library fails at import time.
lib()

__all__ = [
    "LIB_PATH", "RfxError", "Context", "DeviceBuffer", "Graph", "Stream", "Timer", "device_count", "lib",
    "read_u64", "u64_cell",
]
