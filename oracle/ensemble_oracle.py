"""CPU oracle of one ensemble pass -- TEST INFRASTRUCTURE ONLY.

Chains the C restatements in liboracle.so with the finalize arithmetic of
pivtools_cli/piv/piv_backend/single_pass_accumulator.py (restated here in
NumPy, float32 where the reference is float32):

* sums of every pair's AB, AA, BB planes (orc_bulkxcorr2d_accumulate_triple,
  weights wA/wB for AB and wB for the autos, cpu_ensemble.py:996-1010);
* mean images and their background correlation (:352-506), R = sum/N - bg
  (:555-640), optional plane-mean removal (:691-697, 105-118), optional
  envelope divide (:700-755), division by sqrt(max(AA(0) BB(0), 1e-12))
  (:757-795);
* the k-space fit (orc_kspace_lm_fit_batch) and the displacement / stress
  extraction (:1067-1290, without the optional outlier methods).

Pins: each stage is pinned separately (accumulate against FP64 NumPy and the
reference identities, the fitter against closed-form transfer functions);
this chain only composes them.
"""

from __future__ import annotations

import numpy as np

import oracle as O


def ensemble_pass(A, B, ctrs_x, ctrs_y, win, weight, bg_method="correlation", env=None, p_centred=None,
                  mask=None, kx4=False, ky4=False):
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    n = A.shape[0]
    ny, nx = len(ctrs_y), len(ctrs_x)
    nwin = ny * nx
    h, w = win
    mask = np.zeros(nwin, bool) if mask is None else np.asarray(mask, bool).reshape(-1)
    maskf = mask.astype(np.float32)
    mean_sub = bg_method in ("window_mean", "correlation+window_mean")
    ab, aa, bb = O.accumulate_triple(A, B, maskf, ctrs_x, ctrs_y, win, weight, weight, weight, weight,
                                     mean_subtract=mean_sub)
    nf32 = np.float32(n)
    R = [x / nf32 for x in (aa, bb, ab)]
    if bg_method.startswith("correlation"):
        ma = (A.astype(np.float64).sum(0) / n).astype(np.float32)
        mb = (B.astype(np.float64).sum(0) / n).astype(np.float32)
        bab, baa, bbb = O.accumulate_triple(ma[None], mb[None], maskf, ctrs_x, ctrs_y, win, weight, weight,
                                            weight, weight, mean_subtract=mean_sub)
        R = [R[0] - baa, R[1] - bbb, R[2] - bab]
    R = [r.reshape(nwin, h, w) for r in R]
    if bg_method.endswith("dc_zero"):
        R = [r - r.mean(axis=(1, 2), keepdims=True, dtype=np.float32) for r in R]
    if env is not None:
        e = np.asarray(env, np.float32)[None]
        R = [r / e for r in R]
    cy, cx = h // 2, w // 2
    nfac = np.sqrt(np.maximum(R[0][:, cy, cx] * R[1][:, cy, cx], np.float32(1e-12)))[:, None, None]
    planes = [(r / nfac).astype(np.float64) for r in R]
    pw = None if p_centred is None else np.tile(np.asarray(p_centred, np.float64).ravel(), nwin)
    ret, o = O.kspace_fit(planes[0], planes[1], planes[2], mask, win, p_win=pw, kx4=kx4, ky4=ky4)
    if ret != 0:
        raise RuntimeError(f"oracle kspace fit failed ({ret})")
    g = o["gauss"]
    ux = g[:, 14].astype(np.float32) - np.float32(w / 2.0 + 1)
    uy = g[:, 15].astype(np.float32) - np.float32(h / 2.0 + 1)
    st = o["status"].copy()
    bad = ~np.isfinite(ux) | ~np.isfinite(uy) | (np.abs(ux) > 0.75 * w) | (np.abs(uy) > 0.75 * h)
    ux[bad] = np.nan
    uy[bad] = np.nan
    st[bad] = 6

    def safe(v, lim):
        r = np.clip(v, -lim, lim)
        return np.where(np.isfinite(r), r, 0.0).astype(np.float32)

    amp = [safe(g[:, i], 1e10) for i in range(3)]
    peak = np.clip(amp[2] / np.sqrt(np.maximum(amp[0] * amp[1], 1e-12)), 0.0, 1.0).astype(np.float32)
    sa = [safe(g[:, i], 1e6) for i in (6, 7, 8)]
    sab = [safe(g[:, i], 1e6) for i in (9, 10, 11)]
    uu = np.maximum(sab[0] - sa[0], 0.0)
    vv = np.maximum(sab[1] - sa[1], 0.0)
    uv = sab[2] - sa[2]
    reason = st.astype(np.int32)
    for arr in (ux, uy, uu, vv, uv, peak):
        arr[mask] = 0.0
    reason[mask] = -1
    out_mask = (st != 0) & ~mask
    for arr in (ux, uy, uu, vv, uv, peak):
        arr[out_mask] = np.nan
    shp = (ny, nx)
    return {"ux": ux.reshape(shp), "uy": uy.reshape(shp), "UU": uu.reshape(shp), "VV": vv.reshape(shp),
            "UV": uv.reshape(shp), "peakheight": peak.reshape(shp), "nan_reason": reason.reshape(shp),
            "planes": planes, "fit": o}
