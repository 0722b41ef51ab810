"""openvdb_amd: MI355X-native (gfx950 HIP) read-only query path for NanoVDB grids.

The kernels live in libnvq.so (C ABI, include/nvq.h); this package is the thin
host side: grid upload (``DeviceGrid``) and batched queries (``ops``).
"""
from ._lib import LIB_PATH, NvqError, lib  # noqa: F401
from .grid import DeviceGrid, aligned_buffer  # noqa: F401
from . import ops, tools  # noqa: F401

__all__ = ["DeviceGrid", "NvqError", "aligned_buffer", "ops", "tools", "lib", "LIB_PATH"]
