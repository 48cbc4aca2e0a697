"""The C ABI: every function include/pivtools_amd.h declares is exported by
the library that owns it, and the libraries load without a device (no
compute call is made here)."""

import ctypes
import os
import re

import pytest

from pivtools_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pivtools_amd.h")

WARP_NAMES = {"fused_warp_set_impl", "fused_warp_get_impl", "fused_symmetric_warp",
              "fused_symmetric_warp_batch"}
KSPACE_NAMES = {"kspace_lm_fit_batch", "kspace_lm_fit_max_threads", "kspace_debug_prep",
                "kspace_fft2_centered"}


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"extern", "if", "defined"}
    return sorted({n for n in names if n not in skip})


def test_header_declares_reference_symbols():
    d = set(_declared())
    for name in ("bulkxcorr2d", "bulkxcorr2d_accumulate", "bulkxcorr2d_accumulate_triple",
                 "lsqpeaklocate_lm", "lsqpeaklocate_lm_batch", "peak_detectability",
                 "pivtools_cpu_supported", "fused_symmetric_warp_batch", "convolve", "xcorr",
                 "kspace_lm_fit_batch", "kspace_fft2_centered", "kspace_debug_prep"):
        assert name in d


@pytest.mark.parametrize("name", _declared())
def test_symbol_exported(name):
    if name.startswith("pivamd_"):
        lib = ctypes.CDLL(_native.lib_path("libpivtools_amd.so"))
    elif name in WARP_NAMES:
        lib = ctypes.CDLL(_native.lib_path("libfusedwarp.so"))
    elif name in KSPACE_NAMES:
        lib = ctypes.CDLL(_native.lib_path("libkspacefit.so"))
    else:
        lib = ctypes.CDLL(_native.lib_path("libbulkxcorr2d.so"))
    assert hasattr(lib, name), name


def test_reference_style_bindings_load():
    L = _native.bulkxcorr2d_lib()
    assert L.bulkxcorr2d.restype is ctypes.c_ubyte
    assert _native.fusedwarp_lib().fused_symmetric_warp_batch.restype is ctypes.c_int
    assert _native.kspacefit_lib().kspace_lm_fit_batch.restype is ctypes.c_int
    assert _native.core().pivamd_size_supported(32, 32) == 1
    assert _native.core().pivamd_size_supported(20, 20) == 0
    assert _native.core().pivamd_size_supported(32, 64) == 1        # rectangular, both axes built
    assert _native.core().pivamd_size_supported(128, 128) == 1
    assert _native.core().pivamd_size_supported(128, 96) == 1
    assert _native.core().pivamd_size_supported(48, 20) == 0


def test_every_device_entry_point_has_argtypes():
    """ctypes passes untyped Python ints as 32-bit C int, which truncates
    device pointers: every pivamd_* entry point must be bound with argtypes
    as soon as the library is loaded, whichever module loads it first."""
    L = _native.core()
    missing = [n for n in _declared() if n.startswith("pivamd_") and getattr(L, n).argtypes is None]
    assert not missing, missing


def test_batch_lane_width_within_reference_bound():
    """peak_locate_lm_batch.h:26-28: PEAKFIT_MAX_LANES = 16 bounds the lane
    width callers size their output arrays by, and the reference's batch gate
    requires peakfit_batch_lanes() == codelet_lanes()
    (PIV_2d_cross_correlate.c:231-239).  Neither call touches the device."""
    L = _native.bulkxcorr2d_lib()
    assert 1 <= L.peakfit_batch_lanes() <= 16
    assert L.peakfit_batch_lanes() == L.codelet_lanes()
    assert L.peakfit_batch_available() == 1


def test_ragged_job_layout_matches_the_library():
    """The ctypes mirror of PivamdXcorrJob has the library's size."""
    assert ctypes.sizeof(_native.XcorrJob) == _native.core().pivamd_xcorr_job_size()
