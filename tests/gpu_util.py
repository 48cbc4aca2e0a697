"""Helpers shared by the GPU tests: drop-in calls with reference argtypes."""

import numpy as np

from pivtools_amd import _native


def dropin_bulkxcorr2d(img_a, img_b, mask, ctrs_x, ctrs_y, win, wa, wb, n_peaks=1, finder=3,
                       ensemble=False, want_planes=False, prefill=0.0):
    L = _native.bulkxcorr2d_lib()
    img_a = np.ascontiguousarray(img_a, np.float32)
    img_b = np.ascontiguousarray(img_b, np.float32)
    if img_a.ndim == 2:
        img_a, img_b = img_a[None], img_b[None]
    N, H, W = img_a.shape
    ny, nx = len(ctrs_y), len(ctrs_x)
    shape = (N, n_peaks, ny, nx)
    o = {k: np.full(shape, prefill, np.float32) for k in ("pk_x", "pk_y", "pk_h", "sx", "sy", "sxy")}
    o["ratio"] = np.full((N, ny, nx), prefill, np.float32)
    planes = np.zeros(N * ny * nx * win[0] * win[1], np.float32) if want_planes else None
    err = L.bulkxcorr2d(img_a, img_b, np.ascontiguousarray(mask, np.float32), np.array([H, W], np.int32), N,
                        np.ascontiguousarray(ctrs_x, np.float32), np.ascontiguousarray(ctrs_y, np.float32),
                        np.array([ny, nx], np.int32), np.ascontiguousarray(wa, np.float32), bool(ensemble),
                        np.ascontiguousarray(wb, np.float32), np.array(win, np.int32), n_peaks, finder,
                        o["pk_x"], o["pk_y"], o["pk_h"], o["sx"], o["sy"], o["sxy"], o["ratio"],
                        planes.ctypes.data if planes is not None else None)
    o["err"] = int(err)
    if planes is not None:
        o["planes"] = planes.reshape(N, ny, nx, win[0], win[1])
    return o


def device_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
