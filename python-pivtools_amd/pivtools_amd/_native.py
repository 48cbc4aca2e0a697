"""ctypes bindings of the three in-tree libraries.

* ``libpivtools_amd.so``: the device-pointer API (``pivamd_*``), driven by
  the host layer of this package with device buffers it owns.
* ``libbulkxcorr2d.so`` / ``libfusedwarp.so`` / ``libkspacefit.so``:
  host-pointer drop-ins with the reference's exported names and arguments;
  bound here with the same ``argtypes`` the reference's callers declare
  (``pivtools_cli/piv/piv_backend/cpu_instantaneous.py:68-165``,
  ``kspace_lm_fitting.py:230-262``).

Loading never falls back: a missing library raises ``FileNotFoundError``.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i = ctypes.c_int

_CORE = None
_XC = None
_FW = None
_KS = None


def lib_path(name: str) -> str:
    path = os.path.join(LIB_DIR, name)
    if not os.path.isfile(path):
        raise FileNotFoundError(
            f"{path} is not built; run python python-pivtools_amd/build.py "
            "(or __graft_entry__.build()) first")
    return path


class XcorrJob(ctypes.Structure):
    """PivamdXcorrJob of include/pivtools_amd.h: one correlation call of a
    ragged launch (device pointers)."""

    _fields_ = [("imgA", ctypes.c_void_p), ("imgB", ctypes.c_void_p), ("u16", ctypes.c_int),
                ("H", ctypes.c_int), ("W", ctypes.c_int), ("n_images", ctypes.c_int),
                ("mask", ctypes.c_void_p), ("ctrX", ctypes.c_void_p), ("ctrY", ctypes.c_void_p),
                ("nWy", ctypes.c_int), ("nWx", ctypes.c_int),
                ("wA", ctypes.c_void_p), ("wB", ctypes.c_void_p), ("planeWeight", ctypes.c_void_p),
                ("pkX", ctypes.c_void_p), ("pkY", ctypes.c_void_p), ("pkH", ctypes.c_void_p),
                ("sx", ctypes.c_void_p), ("sy", ctypes.c_void_p), ("sxy", ctypes.c_void_p),
                ("ratio", ctypes.c_void_p), ("planes", ctypes.c_void_p), ("work", ctypes.c_void_p),
                ("work_bytes", ctypes.c_ulong)]


def core() -> ctypes.CDLL:
    """libpivtools_amd.so with argtypes of the device-pointer API."""
    global _CORE
    if _CORE is None:
        L = ctypes.CDLL(lib_path("libpivtools_amd.so"))
        L.pivamd_size_supported.argtypes = [_i, _i]
        L.pivamd_size_supported.restype = _i
        L.pivamd_set_dft_mode.argtypes = [_i]
        L.pivamd_set_dft_mode.restype = _i
        L.pivamd_xcorr_variants.argtypes = [_i, _vp, _i]
        L.pivamd_xcorr_variants.restype = _i
        L.pivamd_xcorr_default_variant.argtypes = [_i]
        L.pivamd_xcorr_default_variant.restype = _i
        L.pivamd_xcorr_set_variant.argtypes = [_i, _i]
        L.pivamd_xcorr_set_variant.restype = _i
        L.pivamd_xcorr_get_variant.argtypes = [_i]
        L.pivamd_xcorr_get_variant.restype = _i
        L.pivamd_plane_weight.argtypes = [_vp, _i, _i, _vp, _vp]
        L.pivamd_plane_weight.restype = _i
        L.pivamd_convolve.argtypes = [_vp, _vp, _i, _i, _vp, _vp]
        L.pivamd_convolve.restype = _i
        L.pivamd_xcorr_workspace.argtypes = [_i, _i, _i, _i, _i, _i, _i]
        L.pivamd_xcorr_workspace.restype = ctypes.c_size_t
        L.pivamd_xcorr.argtypes = [
            _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i,
            _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp,
        ]
        L.pivamd_xcorr.restype = _i
        L.pivamd_xcorr_u16.argtypes = L.pivamd_xcorr.argtypes
        L.pivamd_infill_knn_workspace.argtypes = [_i]
        L.pivamd_infill_knn_workspace.restype = ctypes.c_ulong
        L.pivamd_infill_knn.argtypes = [_vp, _vp, _i, _i, _i, ctypes.c_float, ctypes.c_float, _i, _i, _vp, _vp,
                                        ctypes.c_ulong, _vp, _vp]
        L.pivamd_infill_knn.restype = _i
        L.pivamd_infill_griddata_workspace.argtypes = [_i, _i, _i]
        L.pivamd_infill_griddata_workspace.restype = ctypes.c_ulong
        L.pivamd_infill_griddata.argtypes = [_vp, _vp, _i, _i, _i, ctypes.c_float, ctypes.c_float, _vp, _vp,
                                             ctypes.c_ulong, _vp, _vp]
        L.pivamd_infill_griddata.restype = _i
        L.pivamd_infill_rbf_workspace.argtypes = [_i, _i, _i]
        L.pivamd_infill_rbf_workspace.restype = ctypes.c_ulong
        L.pivamd_infill_rbf.argtypes = [_vp, _vp, _i, _i, _i, ctypes.c_float, ctypes.c_float, _i, _i,
                                        ctypes.c_double, ctypes.c_double, _vp, _vp, ctypes.c_ulong, _vp, _vp]
        L.pivamd_infill_rbf.restype = _i
        L.pivamd_xcorr_ragged.argtypes = [ctypes.POINTER(XcorrJob), _i, _i, _i, _i, _i, _i, _vp]
        L.pivamd_xcorr_ragged.restype = _i
        L.pivamd_xcorr_job_size.argtypes = []
        L.pivamd_xcorr_job_size.restype = _i
        L.pivamd_xcorr_noalloc.argtypes = L.pivamd_xcorr.argtypes[:27] + [_vp]
        L.pivamd_xcorr_noalloc.restype = _i
        L.pivamd_xcorr_noalloc_max_bytes.argtypes = []
        L.pivamd_xcorr_noalloc_max_bytes.restype = ctypes.c_size_t
        L.pivamd_xcorr_u16.restype = _i
        L.pivamd_warp.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _i, _i, _i, _vp]
        L.pivamd_warp.restype = _i
        L.pivamd_warp_u16.argtypes = L.pivamd_warp.argtypes
        L.pivamd_warp_u16.restype = _i
        L.pivamd_davis_payload_bytes.argtypes = [_i, _i, _i]
        L.pivamd_davis_payload_bytes.restype = ctypes.c_ulong
        L.pivamd_davis_tokens.argtypes = [_vp, _sz, ctypes.POINTER(_sz), _vp, _sz, ctypes.POINTER(ctypes.c_longlong),
                                          ctypes.POINTER(_sz)]
        L.pivamd_davis_tokens.restype = _i
        L.pivamd_lz4_block.argtypes = [_vp, _sz, _vp, _sz]
        L.pivamd_lz4_block.restype = ctypes.c_longlong
        for f in ("pivamd_tiff_lzw", "pivamd_packbits"):
            getattr(L, f).argtypes = [_vp, _sz, _vp, _sz]
            getattr(L, f).restype = ctypes.c_longlong
        L.pivamd_png_unfilter.argtypes = [_vp, _sz, _sz, _i, _vp]
        L.pivamd_png_unfilter.restype = _i
        L.pivamd_jpeg_info.argtypes = [_vp, _sz, _vp]
        L.pivamd_jpeg_info.restype = _i
        L.pivamd_jpeg_decode.argtypes = [_vp, _sz, _vp, _sz]
        L.pivamd_jpeg_decode.restype = _i
        L.pivamd_warp_set_variant.argtypes = [_i]
        L.pivamd_warp_set_variant.restype = _i
        L.pivamd_warp_get_variant.argtypes = []
        L.pivamd_warp_get_variant.restype = _i
        L.pivamd_peakfit.argtypes = [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]
        L.pivamd_peakfit.restype = _i
        L.pivamd_infill_ls_workspace.argtypes = [_i, _i, _i]
        L.pivamd_infill_ls_workspace.restype = ctypes.c_ulong
        L.pivamd_outlier_methods_workspace.argtypes = [_i, _i, _i]
        L.pivamd_outlier_methods_workspace.restype = ctypes.c_ulong
        for name, argt in GLUE_SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = argt
            f.restype = _i
        L.pivamd_kspace_workspace.argtypes = [_i, _i, _i]
        L.pivamd_kspace_workspace.restype = ctypes.c_size_t
        L.pivamd_kspace_fit.argtypes = [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i] + [_vp] * 11 + [
            _i, _vp, ctypes.c_size_t, _vp]
        L.pivamd_sum_images.argtypes = [_vp, _i, ctypes.c_ulong, _vp, _vp]
        L.pivamd_mean_image.argtypes = [_vp, ctypes.c_double, ctypes.c_ulong, _vp, _vp]
        L.pivamd_ensemble_finalize.argtypes = [_vp] * 8 + [_i, _i, _i, ctypes.c_float, ctypes.c_float, _i,
                                                           _vp, _vp, _vp, _vp]
        L.pivamd_ensemble_extract.argtypes = [_vp, _vp, _vp, _i, _i, _i, ctypes.c_double] + [_vp] * 8
        L.pivamd_pad_frames.argtypes = [_vp, _i, _i, _i, _i, _i, _i, _i, _i, _vp, _vp]
        L.pivamd_ensemble_workspace.argtypes = [_i, _i, _i, _i]
        L.pivamd_ensemble_workspace.restype = ctypes.c_size_t
        L.pivamd_accumulate.argtypes = [_vp, _vp, _vp, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _i, _i,
                                        _i, _i, _i, _i, _i, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]
        for f in ("pivamd_sum_images", "pivamd_mean_image", "pivamd_ensemble_finalize", "pivamd_ensemble_extract",
                  "pivamd_accumulate", "pivamd_pad_frames"):
            getattr(L, f).restype = _i
        L.pivamd_kspace_fit.restype = _i
        L.pivamd_kspace_fft2_centered.argtypes = [_vp, _i, _i, _vp, _vp, _vp, ctypes.c_size_t, _vp]
        L.pivamd_kspace_fft2_centered.restype = _i
        L.pivamd_kspace_debug_prep.argtypes = [_vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, ctypes.c_size_t, _vp]
        L.pivamd_kspace_debug_prep.restype = _i
        if hasattr(L, "pivamd_accumulate_triple"):
            L.pivamd_accumulate_triple.argtypes = [
                _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i,
                _i, _i, _i, _vp, _vp, _vp, _vp,
            ]
            L.pivamd_accumulate_triple.restype = _i
        _CORE = L
    return _CORE


# device-pointer glue entry points (multipass loop on the device); bound
# with the library itself so that no caller can reach one without argtypes
_sz = ctypes.c_size_t
_f = ctypes.c_float
GLUE_SIGNATURES: dict = {
    "pivamd_u16_to_f32": [_vp, _vp, _sz, _vp],
    "pivamd_copy_f32": [_vp, _vp, _sz, _vp],
    "pivamd_gate": [_vp, _vp, _vp, _vp, _vp, _i, _i, _f, _f, _vp, _vp, _vp, _vp, _vp],
    "pivamd_median": [_vp, _vp, _vp, _i, _i, _i, _f, _f, _i, _vp, _vp, _vp, _vp, _vp],
    "pivamd_infill_nearest": [_vp, _vp, _i, _i, _i, _f, _f, _vp, _vp, _vp],
    "pivamd_outlier_stage": [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _i, _i, _f, _vp, _vp, _vp, _vp, _vp, _vp,
                             _vp],
    "pivamd_outlier_detect": [_vp, _vp, _vp, _i, _i, _i, _i, _f, _f, _i, _i, _f, _vp, _vp, _vp, _vp, _vp],
    "pivamd_gate_peaks": [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                          _vp],
    "pivamd_peak_select": [_vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pivamd_peak_final": [_vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "pivamd_pad_edge": [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp],
    "pivamd_smooth_axis": [_vp, _i, _i, _i, _i, _vp, _i, _vp, _vp],
    "pivamd_remap_cubic": [_vp, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp],
    "pivamd_fill_masked_nearest": [_vp, _vp, _i, _i, _i, _f, _f, _vp, _vp, _vp],
    "pivamd_infill_ls": [_vp, _vp, _i, _i, _i, _f, _f, _i, _vp, _vp, ctypes.c_ulong, _vp, _vp],
    "pivamd_outlier_sigma": [_vp, _vp, _i, _i, _i, _f, _i, _vp, _vp, ctypes.c_ulong, _vp],
    "pivamd_outlier_div_vort": [_vp, _vp, _i, _i, _i, _f, _f, _vp, _vp, ctypes.c_ulong, _vp],
    "pivamd_fill_linear": [_vp, _i, ctypes.c_long, _i, _vp, _i, _vp, _vp, _vp],
    "pivamd_unpack_mono12p": [_vp, _sz, _i, _i, _i, _vp, _vp, _vp],
    "pivamd_bayer_rg8_superpixel": [_vp, _sz, _i, _i, _i, _vp, _vp, _vp],
    "pivamd_davis_decode": [_vp, _sz, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp],
    "pivamd_gray_from_samples": [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp],
}


def bulkxcorr2d_lib() -> ctypes.CDLL:
    """libbulkxcorr2d.so bound exactly as the reference's correlators bind it."""
    global _XC
    if _XC is None:
        L = ctypes.CDLL(lib_path("libbulkxcorr2d.so"))
        L.bulkxcorr2d.restype = ctypes.c_ubyte
        L.bulkxcorr2d.argtypes = [
            _f32p, _f32p, _f32p, _i32p, _i, _f32p, _f32p, _i32p, _f32p, ctypes.c_bool, _f32p, _i32p,
            _i, _i, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _vp,
        ]
        if hasattr(L, "bulkxcorr2d_accumulate"):
            L.bulkxcorr2d_accumulate.restype = ctypes.c_ubyte
            L.bulkxcorr2d_accumulate.argtypes = [
                _f32p, _f32p, _f32p, _i32p, _i, _f32p, _f32p, _i32p, _f32p, _f32p, _i32p, _vp, _f32p,
            ]
            L.bulkxcorr2d_accumulate_triple.restype = ctypes.c_ubyte
            L.bulkxcorr2d_accumulate_triple.argtypes = [
                _f32p, _f32p, _f32p, _i32p, _i, _f32p, _f32p, _i32p, _f32p, _f32p, _f32p, _f32p, _i32p,
                _vp, _i, _i, _f32p, _f32p, _f32p,
            ]
        L.lsqpeaklocate_lm.restype = None
        L.lsqpeaklocate_lm.argtypes = [_f32p, _i32p, _f32p, _i, _i, _f32p, _vp]
        L.lsqpeaklocate_lm_batch.restype = None
        L.lsqpeaklocate_lm_batch.argtypes = [_f32p, _i, _i32p, _i, _f32p, _f32p, _vp]
        L.peak_detectability.restype = ctypes.c_float
        L.peak_detectability.argtypes = [_f32p, _i32p]
        L.bulkxcorr2d_set_peakfit_impl.argtypes = [_i]
        L.bulkxcorr2d_set_peakfit_impl.restype = _i
        L.bulkxcorr2d_get_peakfit_impl.restype = _i
        L.peakfit_batch_lanes.restype = _i
        L.pivtools_cpu_supported.restype = _i
        L.convolve.restype = ctypes.c_uint
        L.convolve.argtypes = [_f32p, _f32p, _f32p, _i32p]
        L.xcorr.restype = ctypes.c_uint
        L.xcorr.argtypes = [_f32p, _f32p, _f32p, _i32p]
        L.fminvec.restype = ctypes.c_float
        L.fminvec.argtypes = [_f32p, _i]
        L.fmaxvec.restype = ctypes.c_float
        L.fmaxvec.argtypes = [_f32p, _i]
        L.bulkxcorr2d_get_timing.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.bulkxcorr2d_get_timing.restype = None
        _XC = L
    return _XC


def fusedwarp_lib() -> ctypes.CDLL:
    """libfusedwarp.so bound as cpu_instantaneous.py:81-104 binds the reference."""
    global _FW
    if _FW is None:
        L = ctypes.CDLL(lib_path("libfusedwarp.so"))
        L.fused_symmetric_warp_batch.restype = ctypes.c_int
        L.fused_symmetric_warp_batch.argtypes = [
            _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i, _i, _i, _i, _i, _f32p, _f32p, _i, _i, _i,
        ]
        L.fused_symmetric_warp.restype = ctypes.c_int
        L.fused_symmetric_warp.argtypes = [
            _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _i, _i, _i, _i, _f32p, _f32p, _i, _i,
        ]
        L.fused_warp_set_impl.argtypes = [_i]
        L.fused_warp_get_impl.restype = _i
        _FW = L
    return _FW


def kspacefit_lib() -> ctypes.CDLL:
    """libkspacefit.so bound as kspace_lm_fitting.py:230-262 binds the reference."""
    global _KS
    if _KS is None:
        L = ctypes.CDLL(lib_path("libkspacefit.so"))
        f64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
        L.kspace_lm_fit_batch.restype = _i
        L.kspace_lm_fit_batch.argtypes = [f64, f64, f64, u8, _vp, _i, _i, _i, _i, _i, _i, f64, _i32p, f64,
                                          f64, f64, f64, f64, f64, _i32p, _i32p, u8]
        L.kspace_lm_fit_max_threads.restype = _i
        L.kspace_fft2_centered.restype = _i
        L.kspace_fft2_centered.argtypes = [f64, _i, _i, f64, f64]
        L.kspace_debug_prep.restype = _i
        L.kspace_debug_prep.argtypes = [f64, f64, f64, _vp, _i, _i, f64, f64]
        _KS = L
    return _KS


def check(rc: int, what: str) -> None:
    if rc != 0:
        names = {1: "bad argument", 2: "unsupported window size", 3: "kernel launch failed",
                 4: "out of device memory", 5: "runtime error"}
        raise RuntimeError(f"{what} failed: {names.get(rc, rc)}")
