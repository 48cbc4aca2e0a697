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

The 'single' run type is the same chain with the frames zero-padded
(``pad_single``, window_utils.py:447-520), wA != wB for AB and wB for both
autos at the sum size, a central crop to the fitting window, AB divided by
its own envelope and scaled by ``ab_scale`` (:700-795), and the peak height
scaled by ``peak_scale`` in FP64 before the clip (:1191-1210).

Pins: each stage is pinned separately (accumulate against FP64 NumPy and the
reference identities, the fitter against closed-form transfer functions);
this chain only composes them.
"""

from __future__ import annotations

import numpy as np

import oracle as O


def pad_single(frames, padding):
    """Zero border (top, bottom, left, right) around (n, H, W) frames."""
    pt, pb, pl, pr = padding
    return np.pad(np.asarray(frames, np.float32), ((0, 0), (pt, pb), (pl, pr)), mode="constant")


def ensemble_pass(A, B, ctrs_x, ctrs_y, win, weight, bg_method="correlation", env=None, p_centred=None,
                  mask=None, kx4=False, ky4=False, weight_a=None, fit_win=None, env_ab=None, ab_scale=1.0,
                  peak_scale=1.0):
    """One pass; ``win`` is the correlation (FFT) size and ``weight`` the B /
    auto weight; ``weight_a`` (default ``weight``) weights A in AB and
    ``fit_win`` (default ``win``) is the stored plane size."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    n = A.shape[0]
    ny, nx = len(ctrs_y), len(ctrs_x)
    nwin = ny * nx
    wa = weight if weight_a is None else weight_a
    h, w = win if fit_win is None else fit_win
    mask = np.zeros(nwin, bool) if mask is None else np.asarray(mask, bool).reshape(-1)
    maskf = mask.astype(np.float32)
    mean_sub = bg_method in ("window_mean", "correlation+window_mean")
    ab, aa, bb = O.accumulate_triple(A, B, maskf, ctrs_x, ctrs_y, win, wa, weight, weight, weight,
                                     fit_win=fit_win, mean_subtract=mean_sub)
    nf32 = np.float32(n)
    R = [x / nf32 for x in (aa, bb, ab)]
    if bg_method.startswith("correlation"):
        ma = (A.astype(np.float64).sum(0) / n).astype(np.float32)
        mb = (B.astype(np.float64).sum(0) / n).astype(np.float32)
        bab, baa, bbb = O.accumulate_triple(ma[None], mb[None], maskf, ctrs_x, ctrs_y, win, wa, weight,
                                            weight, weight, fit_win=fit_win, mean_subtract=mean_sub)
        R = [R[0] - baa, R[1] - bbb, R[2] - bab]
    R = [r.reshape(nwin, h, w) for r in R]
    if bg_method.endswith("dc_zero"):
        R = [r - r.mean(axis=(1, 2), keepdims=True, dtype=np.float32) for r in R]
    if env is not None:
        e = np.asarray(env, np.float32)[None]
        eab = e if env_ab is None else np.asarray(env_ab, np.float32)[None]
        R = [R[0] / e, R[1] / e, R[2] / eab]
    cy, cx = h // 2, w // 2
    nfac = np.sqrt(np.maximum(R[0][:, cy, cx] * R[1][:, cy, cx], np.float32(1e-12)))[:, None, None]
    if ab_scale != 1.0:
        R[2] = R[2] * np.float32(ab_scale)
    planes = [(r / nfac).astype(np.float64) for r in R]
    if p_centred is None:
        pw = None
    elif np.ndim(p_centred) == 2 and np.shape(p_centred)[0] == nwin and np.shape(p_centred)[1] == h * w:
        pw = np.asarray(p_centred, np.float64).ravel()          # one floor shape per window
    else:
        pw = np.tile(np.asarray(p_centred, np.float64).ravel(), nwin)
    ret, o = O.kspace_fit(planes[0], planes[1], planes[2], mask, (h, w), p_win=pw, kx4=kx4, ky4=ky4)
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
    peak = (amp[2] / np.sqrt(np.maximum(amp[0] * amp[1], 1e-12))).astype(np.float32)
    if peak_scale != 1.0:
        peak = (peak.astype(np.float64) * float(peak_scale)).astype(np.float32)
    peak = np.clip(peak, 0.0, 1.0).astype(np.float32)
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


def ensemble_multipass(A, B, grids, weight_type="square", bg_method="correlation", floor_fn=None,
                       rounding=False, smoothing=False, mid_infill="nearest", final_infill="biharmonic",
                       interp_mode=0, run_types=None, sum_window=(32, 32), fit_window=None):
    """Multi-pass ensemble (cpu_ensemble.py:912-960,1445-1660,
    single_pass_accumulator.py:820-890,1146-1400): pass p > 0 warps every
    pair with the previous pass's infilled field as one shared predictor
    (edge-padded onto the previous padded centres, optionally smoothed and
    rounded to even integers), correlates the warped pairs on its own grid,
    and adds the cubic remap of the predictor at its window centres back to
    the fitted displacements.  floor_fn(pred_win or None, win, run_type)
    gives the coloured floor shape (None: flat).  A 'single' pass pads the
    warped pairs and correlates at the sum size (its grid centres are in
    original coordinates; the padding is added back for the correlation)."""
    import multipass as M
    from pivtools_amd.window import window_weight

    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    out, field = [], None
    for p, g in enumerate(grids):
        Ap, Bp, win = A, B, None
        if p > 0:
            gp = grids[p - 1]
            pad = M.pad_edge(field, gp.n_pre, gp.n_post)
            if smoothing and g.sd > 0:
                pad = M.smooth(pad, g.sd, g.ksize)
            if rounding:
                pad = (2.0 * np.round(pad / 2.0)).astype(np.float32)
            win = M.remap_cubic(pad, g.pred_map_x, g.pred_map_y)[:, 0]
            if rounding:
                win = (2.0 * np.round(win / 2.0)).astype(np.float32)
            Ap, Bp = O.fused_warp_batch(A, B, pad[0, 0], pad[1, 0], np.asarray(gp.ctrs_y_all, np.float32),
                                        np.asarray(gp.ctrs_x_all, np.float32), interp_mode, True, rounding)
        rt = "standard" if run_types is None else run_types[p]
        pc = None if floor_fn is None else floor_fn(win, g.win, rt)
        if rt == "single":
            from pivtools_amd.window import compute_padding_for_single_mode
            pad_s = compute_padding_for_single_mode(g.win, sum_window)
            wa = window_weight(g.win, "singlepix", sum_window)
            wb = window_weight(sum_window, "bsingle", sum_window)
            area = sum_window[0] * sum_window[1] / (g.win[0] * g.win[1])
            r = ensemble_pass(pad_single(Ap, pad_s), pad_single(Bp, pad_s), g.ctrs_x + np.float32(pad_s[2]),
                              g.ctrs_y + np.float32(pad_s[0]), tuple(sum_window), wb, bg_method=bg_method,
                              p_centred=pc, weight_a=wa, fit_win=fit_window, ab_scale=area,
                              peak_scale=np.sqrt(area))
        else:
            w = window_weight(g.win, weight_type)
            r = ensemble_pass(Ap, Bp, g.ctrs_x, g.ctrs_y, g.win, w, bg_method=bg_method, p_centred=pc)
        if win is not None:
            r["ux"] = (r["ux"] + win[1]).astype(np.float32)
            r["uy"] = (r["uy"] + win[0]).astype(np.float32)
        r["predictor"] = win
        u = np.stack([r["uy"], r["ux"]])[:, None].astype(np.float32)
        method = final_infill if p == len(grids) - 1 else mid_infill
        if method != "none" and np.isnan(u).any():
            if method == "nearest":
                u = M.infill_nearest(u, g.spacing).astype(np.float32)
            else:
                u = M.infill_ls(u, g.spacing, 1 if method == "harmonic" else 2).astype(np.float32)
            r["uy"], r["ux"] = u[0, 0], u[1, 0]
        field = u
        out.append(r)
    return out
