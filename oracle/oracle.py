"""CPU oracle for the PIVtools correlation hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker used by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``.  The shipped library never imports
it.  Two flavours live here:

* ``liboracle.so`` (``piv_oracle.c``): a C restatement of the reference C
  path, same arguments and outputs as the reference's ctypes bindings
  (``pivtools_cli/piv/piv_backend/cpu_instantaneous.py:68-165``).
* FP64 NumPy functions (``*_fp64``): the same mathematics in double precision,
  used to state the error of every result against an exact-ish answer.

Pins (which check fixes each function) are listed in ``piv_oracle.c``'s
header and in the tests that import this module.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LIB_CR = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(libm: str = "host") -> ctypes.CDLL:
    """The C restatement.  libm="host": expf/logf from the host libm, as the
    reference's C; libm="cr": correctly rounded expf/logf (the device
    fitter's definition, liboracle_cr.so)."""
    global _LIB, _LIB_CR
    if libm == "cr":
        if _LIB_CR is None:
            path = os.path.join(HERE, "liboracle_cr.so")
            if not os.path.isfile(path):
                build()
            _LIB_CR = _bind(ctypes.CDLL(path))
        return _LIB_CR
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.isfile(path):
            build()
        _LIB = _bind(ctypes.CDLL(path))
    return _LIB


def _bind(L: ctypes.CDLL) -> ctypes.CDLL:
    L.orc_bulkxcorr2d.restype = ctypes.c_ubyte
    L.orc_bulkxcorr2d.argtypes = [
        _f32p, _f32p, _f32p, _i32p, ctypes.c_int, _f32p, _f32p, _i32p, _f32p,
        ctypes.c_int, _f32p, _i32p, ctypes.c_int, ctypes.c_int,
        _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_void_p,
    ]
    L.orc_lsqpeaklocate_lm.restype = None
    L.orc_lsqpeaklocate_lm.argtypes = [
        _f32p, _i32p, _f32p, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_void_p,
    ]
    L.orc_peak_detectability.restype = ctypes.c_float
    L.orc_peak_detectability.argtypes = [_f32p, _i32p]
    L.orc_convolve.restype = ctypes.c_uint
    L.orc_convolve.argtypes = [_f32p, _f32p, _f32p, _i32p]
    L.orc_fused_symmetric_warp_batch.restype = ctypes.c_int
    L.orc_fused_symmetric_warp_batch.argtypes = [
        _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p, ctypes.c_int,
        ctypes.c_int, ctypes.c_int,
    ]
    L.orc_bulkxcorr2d_accumulate_triple.restype = ctypes.c_ubyte
    L.orc_bulkxcorr2d_accumulate_triple.argtypes = [
        _f32p, _f32p, _f32p, _i32p, ctypes.c_int, _f32p, _f32p, _i32p,
        _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_int, _f32p, _f32p, _f32p,
    ]
    L.orc_bulkxcorr2d_accumulate.restype = ctypes.c_ubyte
    L.orc_bulkxcorr2d_accumulate.argtypes = [
        _f32p, _f32p, _f32p, _i32p, ctypes.c_int, _f32p, _f32p, _i32p,
        _f32p, _f32p, _i32p, ctypes.c_void_p, _f32p,
    ]
    L.orc_max_threads.restype = ctypes.c_int
    if not hasattr(L, "orc_kspace_lm_fit_batch"):     # liboracle_cr.so: peak path only
        return L
    _f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    _u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
    L.orc_kspace_lm_fit_batch.restype = ctypes.c_int
    L.orc_kspace_lm_fit_batch.argtypes = [
        _f64p, _f64p, _f64p, _u8p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, _i32p, _f64p, _f64p, _f64p, _f64p,
        _f64p, _f64p, _i32p, _i32p, _u8p,
    ]
    L.orc_kspace_fft2_centered.restype = ctypes.c_int
    L.orc_kspace_fft2_centered.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, _f64p, _f64p]
    L.orc_kspace_debug_prep.restype = ctypes.c_int
    L.orc_kspace_debug_prep.argtypes = [_f64p, _f64p, _f64p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_int, _f64p, _f64p]
    return L


# ---------------------------------------------------------------------------
# C restatement wrappers (same argument meaning as the reference bindings)
# ---------------------------------------------------------------------------


def bulkxcorr2d(img_a, img_b, mask, ctrs_x, ctrs_y, win, weight_a, weight_b,
                n_peaks=1, peak_finder=3, ensemble=False, want_planes=False):
    """Run the oracle ``bulkxcorr2d``; returns a dict of the output arrays."""
    img_a = np.ascontiguousarray(img_a, np.float32)
    img_b = np.ascontiguousarray(img_b, np.float32)
    if img_a.ndim == 2:
        img_a, img_b = img_a[None], img_b[None]
    N, H, W = img_a.shape
    ny, nx = len(ctrs_y), len(ctrs_x)
    out_shape = (N, n_peaks, ny, nx)
    o = {k: np.zeros(out_shape, np.float32) for k in ("pk_x", "pk_y", "pk_h", "sx", "sy", "sxy")}
    o["ratio"] = np.zeros((N, ny, nx), np.float32)
    planes = np.zeros(N * ny * nx * win[0] * win[1], np.float32) if want_planes else None
    err = lib().orc_bulkxcorr2d(
        img_a, img_b, np.ascontiguousarray(mask, np.float32), np.array([H, W], np.int32), N,
        np.ascontiguousarray(ctrs_x, np.float32), np.ascontiguousarray(ctrs_y, np.float32),
        np.array([ny, nx], np.int32), np.ascontiguousarray(weight_a, np.float32), int(ensemble),
        np.ascontiguousarray(weight_b, np.float32), np.array(win, np.int32), n_peaks, peak_finder,
        o["pk_x"], o["pk_y"], o["pk_h"], o["sx"], o["sy"], o["sxy"], o["ratio"],
        planes.ctypes.data if planes is not None else None,
    )
    o["err"] = int(err)
    if planes is not None:
        o["planes"] = planes.reshape(N, ny, nx, win[0], win[1])
    return o


def lsqpeaklocate_lm(plane, n_peaks=1, fit_type=3, weight=None, libm="host"):
    plane = np.ascontiguousarray(plane, np.float32)
    N = np.array(plane.shape, np.int32)
    loc = np.zeros(3 * n_peaks, np.float32)
    std = np.zeros(3 * n_peaks, np.float32)
    w = None if weight is None else np.ascontiguousarray(weight, np.float32)
    lib(libm).orc_lsqpeaklocate_lm(plane, N, loc, n_peaks, fit_type, std,
                                   None if w is None else w.ctypes.data)
    return loc, std


def peak_detectability(plane):
    plane = np.ascontiguousarray(plane, np.float32)
    return float(lib().orc_peak_detectability(plane, np.array(plane.shape, np.int32)))


def plane_weight(weight_b):
    """numel / (wB (*) wB) on the central crop (xcorr.c:22-89)."""
    wb = np.ascontiguousarray(weight_b, np.float32)
    conv = np.zeros_like(wb)
    err = lib().orc_convolve(wb, wb, conv, np.array(wb.shape, np.int32))
    if err:
        raise ValueError(f"convolve error {err}")
    return (np.float32(wb.size) / conv).astype(np.float32)


def fused_warp_batch(imgs_a, imgs_b, pred_dy, pred_dx, ctrs_y, ctrs_x, interp_mode=0,
                     shared_predictor=False, round_shifts=False):
    imgs_a = np.ascontiguousarray(imgs_a, np.float32)
    imgs_b = np.ascontiguousarray(imgs_b, np.float32)
    N, H, W = imgs_a.shape
    pred_dy = np.ascontiguousarray(pred_dy, np.float32)
    pred_dx = np.ascontiguousarray(pred_dx, np.float32)
    nPY, nPX = pred_dy.shape[-2:]
    oa = np.zeros_like(imgs_a)
    ob = np.zeros_like(imgs_b)
    r = lib().orc_fused_symmetric_warp_batch(
        imgs_a, imgs_b, oa, ob, pred_dy, pred_dx, N, H, W, nPY, nPX,
        np.ascontiguousarray(ctrs_y, np.float32), np.ascontiguousarray(ctrs_x, np.float32),
        interp_mode, int(shared_predictor), int(round_shifts))
    if r != 0:
        raise RuntimeError(f"oracle warp failed ({r})")
    return oa, ob


def accumulate_triple(imgs_a, imgs_b, mask, ctrs_x, ctrs_y, win, w_a_ab, w_b_ab, w_auto_a,
                      w_auto_b, fit_win=None, mean_subtract=False, per_pair_norm=False):
    imgs_a = np.ascontiguousarray(imgs_a, np.float32)
    imgs_b = np.ascontiguousarray(imgs_b, np.float32)
    N, H, W = imgs_a.shape
    ny, nx = len(ctrs_y), len(ctrs_x)
    oh, ow = fit_win if fit_win is not None else win
    outs = [np.zeros(ny * nx * oh * ow, np.float32) for _ in range(3)]
    fw = None if fit_win is None else np.array(fit_win, np.int32)
    err = lib().orc_bulkxcorr2d_accumulate_triple(
        imgs_a, imgs_b, np.ascontiguousarray(mask, np.float32), np.array([H, W], np.int32), N,
        np.ascontiguousarray(ctrs_x, np.float32), np.ascontiguousarray(ctrs_y, np.float32),
        np.array([ny, nx], np.int32),
        *[np.ascontiguousarray(x, np.float32) for x in (w_a_ab, w_b_ab, w_auto_a, w_auto_b)],
        np.array(win, np.int32), None if fw is None else fw.ctypes.data,
        int(mean_subtract), int(per_pair_norm), *outs)
    if err:
        raise RuntimeError(f"oracle accumulate_triple error {err}")
    return [o.reshape(ny, nx, oh, ow) for o in outs]


# ---------------------------------------------------------------------------
# FP64 NumPy oracle
# ---------------------------------------------------------------------------


def window_origin(ctr, n):
    """row_min/col_min = floor(ctr - (n-1)/2 + 0.5) (PIV_2d_cross_correlate.c:299-300)."""
    return int(np.floor(np.float64(np.float32(ctr)) - (n - 1.0) / 2.0 + 0.5))


def xcorr_plane_fp64(win_a, win_b):
    """shift(ifft2(fft2(B) * conj(fft2(A)))) in FP64: zero lag at [H/2][W/2]."""
    fa = np.fft.fft2(np.asarray(win_a, np.float64))
    fb = np.fft.fft2(np.asarray(win_b, np.float64))
    return np.fft.fftshift(np.real(np.fft.ifft2(fb * np.conj(fa))))


def xcorr_plane_bruteforce(win_a, win_b):
    """Direct circular correlation ref[m,n] = sum B[(i+m)%H,(j+n)%W] A[i,j], shifted.

    The brute-force FP64 definition the reference's own gate uses
    (pivtools_cli/lib/test_codelet_gate.c:35-51).  O(N^4): small windows only.
    """
    a = np.asarray(win_a, np.float64)
    b = np.asarray(win_b, np.float64)
    H, W = a.shape
    out = np.zeros((H, W))
    for m in range(H):
        for n in range(W):
            out[m, n] = np.sum(np.roll(np.roll(b, -m, 0), -n, 1) * a)
    return np.fft.fftshift(out)


def windows_fp64(img_a, img_b, ctrs_x, ctrs_y, win, wa, wb, ensemble=False):
    """Extract, weight and mean-subtract every window in FP64 (instantaneous rule)."""
    H, W = img_a.shape
    h, w = win
    out_a, out_b, ok = [], [], []
    for jj, cy in enumerate(ctrs_y):
        r0 = window_origin(cy, h)
        for ii, cx in enumerate(ctrs_x):
            c0 = window_origin(cx, w)
            valid = r0 >= 0 and c0 >= 0 and r0 + h <= H and c0 + w <= W
            ok.append(valid)
            if not valid:
                out_a.append(np.zeros((h, w)))
                out_b.append(np.zeros((h, w)))
                continue
            a = img_a[r0:r0 + h, c0:c0 + w].astype(np.float64) * wa
            b = img_b[r0:r0 + h, c0:c0 + w].astype(np.float64) * wb
            if not ensemble:
                a = a - a.mean()
                b = b - b.mean()
            out_a.append(a)
            out_b.append(b)
    return np.array(out_a), np.array(out_b), np.array(ok)


def planes_fp64(img_a, img_b, ctrs_x, ctrs_y, win, wa, wb, ensemble=False):
    """All window planes of one pair in FP64, shape (ny*nx, h, w)."""
    a, b, ok = windows_fp64(img_a, img_b, ctrs_x, ctrs_y, win, wa, wb, ensemble)
    fa = np.fft.fft2(a)
    fb = np.fft.fft2(b)
    planes = np.fft.fftshift(np.real(np.fft.ifft2(fb * np.conj(fa))), axes=(-2, -1))
    return planes, ok


# ---------------------------------------------------------------------------
# k-space transfer-function fit (kspace_oracle.c)
# ---------------------------------------------------------------------------
KSPACE_DIAG = ("gain", "N0", "b4x", "b4y", "cost_per_pt", "n_valid", "iter", "conv")


def kspace_alloc(n):
    """Output arrays of kspace_lm_fit_batch, initialised as the reference
    caller does (kspace_lm_fitting.py:862-875)."""
    return {
        "gauss": np.zeros((n, 16)), "status": np.full(n, -1, np.int32), "init": np.zeros((n, 16)),
        "gain": np.full(n, np.nan), "N0": np.full(n, np.nan), "b4x": np.full(n, np.nan),
        "b4y": np.full(n, np.nan), "cost_per_pt": np.full(n, np.nan),
        "n_valid": np.zeros(n, np.int32), "iter": np.zeros(n, np.int32), "conv": np.zeros(n, np.uint8),
    }


def kspace_fit(r_aa, r_bb, r_ab, mask, corr, p_win=None, kx4=False, ky4=False, threads=0):
    """Oracle kspace_lm_fit_batch on (n, h, w) FP64 planes; returns (ret, outputs)."""
    h, w = corr
    aa = np.ascontiguousarray(r_aa, np.float64).reshape(-1)
    bb = np.ascontiguousarray(r_bb, np.float64).reshape(-1)
    ab = np.ascontiguousarray(r_ab, np.float64).reshape(-1)
    n = aa.size // (h * w)
    m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
    pw = None if p_win is None else np.ascontiguousarray(p_win, np.float64).reshape(-1)
    o = kspace_alloc(n)
    ret = lib().orc_kspace_lm_fit_batch(
        aa, bb, ab, m, None if pw is None else pw.ctypes.data, n, h, w, int(kx4), int(ky4), threads,
        o["gauss"], o["status"], o["init"], o["gain"], o["N0"], o["b4x"], o["b4y"], o["cost_per_pt"],
        o["n_valid"], o["iter"], o["conv"])
    return ret, o


def kspace_fft2_centered(plane):
    plane = np.ascontiguousarray(plane, np.float64)
    h, w = plane.shape
    re = np.empty((h, w))
    im = np.empty((h, w))
    ret = lib().orc_kspace_fft2_centered(plane, h, w, re, im)
    return ret, re + 1j * im


def kspace_debug_prep(r_aa, r_bb, r_ab, p_centred=None):
    h, w = r_aa.shape
    dv = np.empty(h * w)
    n0 = np.zeros(1)
    pc = None if p_centred is None else np.ascontiguousarray(p_centred, np.float64)
    ret = lib().orc_kspace_debug_prep(np.ascontiguousarray(r_aa, np.float64),
                                      np.ascontiguousarray(r_bb, np.float64),
                                      np.ascontiguousarray(r_ab, np.float64),
                                      None if pc is None else pc.ctypes.data, h, w, dv, n0)
    return ret, dv.reshape(h, w), float(n0[0])
