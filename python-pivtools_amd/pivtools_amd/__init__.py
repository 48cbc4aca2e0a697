"""MI355X-native PIVtools correlation hot path.

The compute lives in hand-written HIP kernels for gfx950 behind a C ABI
(``include/pivtools_amd.h``): ``lib/libbulkxcorr2d.so`` and
``lib/libfusedwarp.so`` are drop-ins for the reference's ctypes libraries,
``lib/libpivtools_amd.so`` is the device-pointer layer this package drives.
There is no CPU fallback: every entry point raises if the libraries are not
built or no device is present.
"""

from . import _native  # noqa: F401
from .window import (BUILT_FFT_SIZES, build_pass_grids, compute_window_centers,  # noqa: F401
                     window_weight)

__all__ = ["BUILT_FFT_SIZES", "build_pass_grids", "compute_window_centers", "window_weight"]
