"""MI355X (gfx950) dense linear-solve engine for the scirs2-linalg solve path.

The compute path is libscirs_gpu.so (HIP kernels behind the C ABI of
include/scirs_gpu.h); ``scirs_amd.linalg`` is its numpy/torch front end.
"""
from . import _native  # noqa: F401

__all__ = ["_native"]
