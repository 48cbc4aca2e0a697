"""Whole ensemble pass on the device (accumulate -> finalize -> k-space fit
-> extraction) against the oracle chain on the same seeded pairs.

Parity is stated in two parts, as for the LM peak fit:
* the finalized FP64 planes agree with the oracle's to 1e-5 relative L2
  (the device keeps FP64 spectral sums where the reference sums float32
  planes, so the planes differ at the float32 rounding level);
* the device fit of the device planes equals the oracle fit of those same
  planes to the reference's own C-vs-oracle gate (identical statuses,
  |dSigma|, |dmu| < 1e-8);
and end to end the statuses and the stresses agree within the tolerance the
plane differences allow.
"""

import numpy as np
import pytest
import torch

import ensemble_oracle as EO
import oracle as O
from pivtools_amd import synthetic as S
from pivtools_amd.ensemble import EnsemblePIV
from pivtools_amd.window import compute_window_centers, window_weight

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _pairs(n, shape, seed0=500, log_scale=1.5):
    A, B = [], []
    for k in range(n):
        a, b = S.ensemble_pair(seed0 + k, shape=shape, log_scale=log_scale)
        A.append(a.numpy())
        B.append(b.numpy())
    return np.stack(A), np.stack(B)


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


@pytest.mark.parametrize("bg,floor", [("correlation", "coloured"), ("correlation", "flat"),
                                      ("window_mean", "coloured"), ("correlation+dc_zero", "flat")])
def test_ensemble_pass_matches_oracle(bg, floor):
    shape, win = (192, 192), (32, 32)
    A, B = _pairs(12, shape)
    print(f"ensemble pass bg={bg} floor={floor}", flush=True)
    ens = EnsemblePIV(shape, win, 75, "square", bg_method=bg, kspace_floor=floor)
    for i in range(0, 12, 5):        # ragged batches: 5, 5, 2
        ens.accumulate(torch.as_tensor(A[i:i + 5], device=dev), torch.as_tensor(B[i:i + 5], device=dev))
    res = ens.finalize()
    torch.cuda.synchronize()
    wc = compute_window_centers(shape, win, 75)
    w = window_weight(win, "square")
    pc = ens.floor_shape() if floor == "coloured" else None
    ref = EO.ensemble_pass(A, B, wc.win_ctrs_x, wc.win_ctrs_y, win, w, bg_method=bg, p_centred=pc)
    dplanes = [p.cpu().numpy() for p in res.planes]
    for dp, rp in zip(dplanes, ref["planes"]):
        assert _rel(dp, rp) < 1e-5
    # the fitter on identical inputs
    pw = None if pc is None else np.tile(pc.ravel(), ens.nwin)
    ret, o = O.kspace_fit(dplanes[0], dplanes[1], dplanes[2], np.zeros(ens.nwin, bool), win, p_win=pw)
    st = res.status.cpu().numpy()
    g = res.gauss.cpu().numpy()
    np.testing.assert_array_equal(st, o["status"])
    ok = st == 0
    assert ok.sum() > 0.5 * ens.nwin
    assert np.abs(g[ok, 9:12] - o["gauss"][ok, 9:12]).max() < 1e-8
    assert np.abs(g[ok, 14:16] - o["gauss"][ok, 14:16]).max() < 1e-8
    # end to end
    same = (st == 0) & (ref["fit"]["status"] == 0)
    assert np.mean(st != ref["fit"]["status"]) < 0.02
    for k in ("ux", "uy", "UU", "VV", "UV"):
        d = res.__dict__[k].cpu().numpy().ravel()[same] - ref[k].ravel()[same]
        assert np.abs(d).max() < 1e-3, (k, np.abs(d).max())
    agree = st == ref["fit"]["status"]
    np.testing.assert_array_equal(np.isnan(res.ux.cpu().numpy().ravel()[agree]), np.isnan(ref["ux"].ravel()[agree]))
    np.testing.assert_array_equal(res.nan_reason.cpu().numpy().ravel()[agree], ref["nan_reason"].ravel()[agree])


def test_ensemble_uint16_frames_and_mask():
    shape, win = (160, 160), (32, 32)
    A, B = _pairs(6, shape, seed0=900)
    Au = np.clip(np.rint(A * 3000), 0, 4095).astype(np.int16)
    Bu = np.clip(np.rint(B * 3000), 0, 4095).astype(np.int16)
    wc = compute_window_centers(shape, win, 75)
    mask = np.zeros(wc.n_win_y * wc.n_win_x, bool)
    mask[::5] = True
    ens = EnsemblePIV(shape, win, 75, kspace_floor="flat", mask=mask)
    ens.accumulate(torch.as_tensor(Au, device=dev), torch.as_tensor(Bu, device=dev))
    res = ens.finalize()
    ref = EO.ensemble_pass(Au.astype(np.float32), Bu.astype(np.float32), wc.win_ctrs_x, wc.win_ctrs_y, win,
                           window_weight(win, "square"), mask=mask)
    reason = res.nan_reason.cpu().numpy().ravel()
    assert np.all(reason[mask] == -1)
    assert np.all(res.ux.cpu().numpy().ravel()[mask] == 0.0)
    for dp, rp in zip(res.planes, ref["planes"]):
        dp = dp.cpu().numpy()[~mask]
        assert _rel(dp, rp[~mask]) < 1e-5
    np.testing.assert_array_equal(reason[mask], ref["nan_reason"].ravel()[mask])


def test_stress_recovery_uniform_shear():
    """Physical check: the Reynolds stresses of a flow with a known
    displacement PDF.  Every pair translates by a random constant (normal,
    std 0.5 px in x, 0.3 px in y), so the ensemble transfer ratio is the
    characteristic function of that PDF: UU = 0.25, VV = 0.09, UV = 0."""
    shape, win = (160, 160), (32, 32)
    rng = np.random.default_rng(77)
    A, B, dxs, dys = [], [], [], []
    for k in range(48):
        r = np.random.default_rng(3000 + k)
        x, y, inten = S._positions(r, shape, 0.02)
        dx, dy = rng.normal(1.0, 0.5), rng.normal(-0.5, 0.3)
        dxs.append(dx)
        dys.append(dy)
        A.append(S.render(shape, x, y, inten).numpy())
        B.append(S.render(shape, x + dx, y + dy, inten).numpy())
    ens = EnsemblePIV(shape, win, 75, kspace_floor="flat")
    ens.accumulate(torch.as_tensor(np.stack(A), device=dev), torch.as_tensor(np.stack(B), device=dev))
    res = ens.finalize()
    ok = res.nan_reason.cpu().numpy() == 0
    assert ok.mean() > 0.8
    uu = np.median(res.UU.cpu().numpy()[ok])
    vv = np.median(res.VV.cpu().numpy()[ok])
    ux = np.median(res.ux.cpu().numpy()[ok])
    # the sample moments of the 48 drawn shifts are what the fit sees
    assert abs(ux - np.mean(dxs)) < 0.05, (ux, np.mean(dxs))
    assert abs(uu - np.var(dxs)) < 0.05, (uu, np.var(dxs))
    assert abs(vv - np.var(dys)) < 0.05, (vv, np.var(dys))


def _floor_fn(mp_ens):
    """The floor shape the device used, for the oracle chain (pinned
    separately against the brute-force quadratic form)."""
    def fn(win_pred, win, run_type="standard"):
        e = EnsemblePIV((mp_ens.H, mp_ens.W), win, 50, "square", device=dev, run_type=run_type,
                        sum_window=mp_ens.sum_window)
        if win_pred is not None:
            e.set_predictor(torch.as_tensor(win_pred, device=dev))
        return e.floor_shape()
    return fn


@pytest.mark.parametrize("rounding", [False, True])
def test_multipass_ensemble_matches_oracle(rounding):
    """Two ensemble passes, 32 x 32 then 16 x 16 at 50 %, the second warped
    with the first pass's infilled field as one shared predictor (optionally
    rounded to even integers): pass planes within 1e-5 relative L2 of the
    oracle chain's, identical fit statuses on pass 0, and displacements with
    the predictor added back within 2e-3 px where both fits succeed."""
    from pivtools_amd.ensemble import EnsembleMultipass
    import ensemble_oracle as EO2

    shape = (160, 192)
    A, B = _pairs(10, shape, seed0=800, log_scale=2.0)
    mp = EnsembleMultipass(shape, [(32, 32), (16, 16)], [50, 50], predictor_rounding=rounding,
                           mid_infill="nearest", final_infill="none")
    print(f"multipass ensemble rounding={rounding}", flush=True)
    res = mp.run(torch.as_tensor(A, device=dev), torch.as_tensor(B, device=dev), batch=4)
    torch.cuda.synchronize()
    ref = EO2.ensemble_multipass(A, B, mp.grids, floor_fn=_floor_fn(mp), rounding=rounding,
                                 mid_infill="nearest", final_infill="none")
    for p, (r, o) in enumerate(zip(res, ref)):
        st, ost = r.status.cpu().numpy(), o["fit"]["status"]
        assert np.mean(st != ost) < 0.02, p
        ux, uy = r.ux.cpu().numpy(), r.uy.cpu().numpy()
        both = np.isfinite(ux) & np.isfinite(o["ux"]) & (st.reshape(ux.shape) == ost.reshape(ux.shape))
        assert both.mean() > 0.8
        d = np.hypot(ux[both] - o["ux"][both], uy[both] - o["uy"][both])
        if p == 0:
            for dp, rp in zip([x.cpu().numpy() for x in r.planes], o["planes"]):
                assert _rel(dp, rp) < 1e-5
            assert d.max() < 1e-3
        else:
            # a fit that flips between the chains changes the infilled
            # predictor around it, so the next pass warps differently there
            assert np.median(d) < 1e-4 and np.mean(d < 2e-3) > 0.95, (np.median(d), np.mean(d < 2e-3))
    # the second pass measured the residual of warped pairs: the added-back
    # field follows the flow, not the residual
    assert np.nanmedian(np.abs(res[1].ux.cpu().numpy())) > 0.5
    if rounding:
        w = ref[1]["predictor"]
        assert np.all(np.abs(w / 2.0 - np.round(w / 2.0)) == 0)


# ---------------------------------------------------------------- single run type
def _single_ref(A, B, ens, bg, floor, envelope=False):
    from pivtools_amd.window import compute_window_centers_single_mode
    m, S_ = ens.win, ens.comp
    wp = compute_window_centers_single_mode((ens.H, ens.W), m, S_, ens.overlap)
    assert wp.padding == ens.padding
    wa = window_weight(m, "singlepix", S_)
    wb = window_weight(S_, "bsingle", S_)
    env = env_ab = None
    if envelope:
        from pivtools_amd import floor_psd
        env = floor_psd.pair_envelope(wb, wb, ens.corr).astype(np.float32)
        env_ab = floor_psd.pair_envelope(wa, wb, ens.corr).astype(np.float32)
    pc = ens.floor_shape() if floor == "coloured" else None
    area = S_[0] * S_[1] / (m[0] * m[1])
    return EO.ensemble_pass(EO.pad_single(A, wp.padding), EO.pad_single(B, wp.padding), wp.win_ctrs_x,
                            wp.win_ctrs_y, S_, wb, bg_method=bg, env=env, env_ab=env_ab, p_centred=pc, weight_a=wa,
                            fit_win=None if ens.corr == S_ else ens.corr, ab_scale=area, peak_scale=np.sqrt(area))


@pytest.mark.parametrize("m,S_,fit,bg,floor,envelope", [
    ((8, 8), (32, 32), None, "correlation", "coloured", False),
    ((6, 6), (32, 32), (24, 24), "correlation", "flat", False),
    ((4, 6), (24, 24), None, "window_mean", "coloured", False),
    ((8, 8), (48, 48), (32, 32), "correlation+dc_zero", "coloured", True),
])
def test_single_run_type_matches_oracle(m, S_, fit, bg, floor, envelope):
    """The single run type on the device (padding kernel, singlepix / bsingle
    weights at the sum size, fitting-window crop, AB envelope and area scale,
    scaled peak height) against the oracle chain on the padded frames."""
    shape = (96, 112)
    A, B = _pairs(8, shape, seed0=1500, log_scale=1.0)
    print(f"single run type m={m} sum={S_} fit={fit} bg={bg} floor={floor} env={envelope}", flush=True)
    ens = EnsemblePIV(shape, m, 50, bg_method=bg, kspace_floor=floor, envelope_divide=envelope,
                      run_type="single", sum_window=S_, fit_window=fit)
    for i in range(0, 8, 3):           # ragged batches: 3, 3, 2
        ens.accumulate(torch.as_tensor(A[i:i + 3], device=dev), torch.as_tensor(B[i:i + 3], device=dev))
    res = ens.finalize()
    torch.cuda.synchronize()
    ref = _single_ref(A, B, ens, bg, floor, envelope)
    assert res.ux.shape == ref["ux"].shape == (ens.ny, ens.nx)
    dplanes = [p.cpu().numpy() for p in res.planes]
    for dp, rp in zip(dplanes, ref["planes"]):
        assert _rel(dp, rp) < 1e-5
    # the fitter on identical inputs (the device planes)
    pc = ens.floor_shape() if floor == "coloured" else None
    pw = None if pc is None else np.tile(pc.ravel(), ens.nwin)
    ret, o = O.kspace_fit(dplanes[0], dplanes[1], dplanes[2], np.zeros(ens.nwin, bool), ens.corr, p_win=pw)
    st, g = res.status.cpu().numpy(), res.gauss.cpu().numpy()
    np.testing.assert_array_equal(st, o["status"])
    ok = st == 0
    assert np.abs(g[ok, 9:12] - o["gauss"][ok, 9:12]).max() < 1e-8
    assert np.abs(g[ok, 14:16] - o["gauss"][ok, 14:16]).max() < 1e-8
    # end to end: the plane rounding moves a few ill-conditioned fits a little
    ost = ref["fit"]["status"]
    assert np.mean(st != ost) < 0.02
    same = (st == 0) & (ost == 0)
    assert same.mean() > 0.5
    for k in ("ux", "uy", "UU", "VV", "UV", "peakheight"):
        d = np.abs(res.__dict__[k].cpu().numpy().ravel()[same] - ref[k].ravel()[same])
        assert np.quantile(d, 0.99) < 1e-4 and d.max() < 1e-2, (k, np.quantile(d, 0.99), d.max())


def test_single_run_type_uint16_frames_and_shift_recovery():
    """16-bit counts padded and converted in one device pass give the same
    planes as float frames of the same counts, and a uniform shift is
    recovered on the small-window grid."""
    shape, m, S_ = (128, 128), (8, 8), (32, 32)
    rng = np.random.default_rng(4)
    A, B = [], []
    for k in range(24):
        x, y, inten = S._positions(rng, shape, 0.03)
        A.append(S.render(shape, x, y, inten).numpy())
        B.append(S.render(shape, x + 1.3, y - 0.6, inten).numpy())
    Au = np.clip(np.rint(np.stack(A) * 2000), 0, 4095).astype(np.int16)
    Bu = np.clip(np.rint(np.stack(B) * 2000), 0, 4095).astype(np.int16)
    out = []
    for a, b in ((Au, Bu), (Au.astype(np.float32), Bu.astype(np.float32))):
        ens = EnsemblePIV(shape, m, 0, kspace_floor="flat", run_type="single", sum_window=S_)
        ens.accumulate(torch.as_tensor(a, device=dev), torch.as_tensor(b, device=dev))
        out.append(ens.finalize())
    torch.cuda.synchronize()
    for p16, p32 in zip(out[0].planes, out[1].planes):
        assert torch.equal(p16, p32)
    res = out[0]
    assert (ens.ny, ens.nx) == (16, 16)
    ok = res.nan_reason.cpu().numpy() == 0
    assert ok.mean() > 0.8
    assert np.median(res.ux.cpu().numpy()[ok]) == pytest.approx(1.3, abs=0.05)
    assert np.median(res.uy.cpu().numpy()[ok]) == pytest.approx(-0.6, abs=0.05)
    ph = res.peakheight.cpu().numpy()[ok]
    assert np.all((ph >= 0) & (ph <= 1)) and np.median(ph) > 0.3


def test_pad_frames_kernel_matches_numpy_pad():
    from pivtools_amd import _native
    L = _native.core()
    rng = np.random.default_rng(9)
    a = rng.integers(0, 65535, (3, 7, 9)).astype(np.uint16).view(np.int16)
    for src in (torch.as_tensor(a, device=dev), torch.as_tensor(rng.normal(size=(3, 7, 9)).astype(np.float32),
                                                                device=dev)):
        out = torch.full((3, 7 + 5, 9 + 3), -1.0, device=dev)
        rc = L.pivamd_pad_frames(src.data_ptr(), int(src.dtype == torch.int16), 3, 7, 9, 3, 1, 12, 12,
                                 out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert rc == 0
        s = src.cpu().numpy()
        want = np.pad(s.view(np.uint16).astype(np.float32) if s.dtype == np.int16 else s,
                      ((0, 0), (3, 2), (1, 2)))
        np.testing.assert_array_equal(out.cpu().numpy(), want)
    assert L.pivamd_pad_frames(src.data_ptr(), 0, 3, 7, 9, 3, 1, 9, 12, out.data_ptr(), 0) != 0


def test_multipass_with_a_single_pass_matches_oracle():
    """A standard 32 x 32 pass, then a single-mode pass (8 x 8 A windows in
    a 32 x 32 sum window at 50 %) on pairs warped with the first pass's
    field: each pass against the oracle chain."""
    from pivtools_amd.ensemble import EnsembleMultipass

    shape = (128, 144)
    A, B = _pairs(8, shape, seed0=1700, log_scale=1.5)
    mp = EnsembleMultipass(shape, [(32, 32), (8, 8)], [50, 50], run_types=["standard", "single"],
                           sum_window=(32, 32), mid_infill="nearest", final_infill="none")
    print("multipass standard -> single", flush=True)
    res = mp.run(torch.as_tensor(A, device=dev), torch.as_tensor(B, device=dev), batch=4)
    torch.cuda.synchronize()
    ref = EO.ensemble_multipass(A, B, mp.grids, floor_fn=_floor_fn(mp), mid_infill="nearest", final_infill="none",
                                run_types=["standard", "single"], sum_window=(32, 32))
    assert res[1].ux.shape == ref[1]["ux"].shape == mp.grids[1].n_win
    for p, (r, o) in enumerate(zip(res, ref)):
        st, ost = r.status.cpu().numpy(), o["fit"]["status"]
        assert np.mean(st != ost) < 0.02, p
        ux, uy = r.ux.cpu().numpy(), r.uy.cpu().numpy()
        both = np.isfinite(ux) & np.isfinite(o["ux"]) & (st.reshape(ux.shape) == ost.reshape(ux.shape))
        assert both.mean() > 0.8
        d = np.hypot(ux[both] - o["ux"][both], uy[both] - o["uy"][both])
        if p == 0:
            for dp, rp in zip([x.cpu().numpy() for x in r.planes], o["planes"]):
                assert _rel(dp, rp) < 1e-5
            assert d.max() < 1e-3
        else:
            assert np.median(d) < 1e-4 and np.mean(d < 2e-3) > 0.95, (np.median(d), np.mean(d < 2e-3))
    assert np.nanmedian(np.abs(res[1].ux.cpu().numpy())) > 0.5


def test_single_run_type_with_mask_and_two_single_passes():
    """Masked windows of a single-mode pass are zero with reason -1 and the
    rest equal the oracle chain with the same mask; two single passes
    (16 then 8 in one 32 sum window) chain like standard ones."""
    from pivtools_amd.ensemble import EnsembleMultipass
    from pivtools_amd.window import compute_window_centers_single_mode

    shape, m, S_ = (96, 112), (8, 8), (32, 32)
    A, B = _pairs(6, shape, seed0=1900, log_scale=1.0)
    wp = compute_window_centers_single_mode(shape, m, S_, 50)
    mask = np.zeros(wp.n_win_y * wp.n_win_x, bool)
    mask[::7] = True
    print("single run type with a mask", flush=True)
    ens = EnsemblePIV(shape, m, 50, kspace_floor="flat", run_type="single", sum_window=S_, mask=mask)
    ens.accumulate(torch.as_tensor(A, device=dev), torch.as_tensor(B, device=dev))
    res = ens.finalize()
    torch.cuda.synchronize()
    wa = window_weight(m, "singlepix", S_)
    wb = window_weight(S_, "bsingle", S_)
    area = S_[0] * S_[1] / (m[0] * m[1])
    ref = EO.ensemble_pass(EO.pad_single(A, wp.padding), EO.pad_single(B, wp.padding), wp.win_ctrs_x,
                           wp.win_ctrs_y, S_, wb, mask=mask, weight_a=wa, ab_scale=area, peak_scale=np.sqrt(area))
    reason = res.nan_reason.cpu().numpy().ravel()
    assert np.all(reason[mask] == -1) and np.all(res.ux.cpu().numpy().ravel()[mask] == 0.0)
    for dp, rp in zip(res.planes, ref["planes"]):
        assert _rel(dp.cpu().numpy()[~mask], rp[~mask]) < 1e-5
    np.testing.assert_array_equal(reason[mask], ref["nan_reason"].ravel()[mask])

    print("multipass single -> single", flush=True)
    mp = EnsembleMultipass(shape, [(16, 16), (8, 8)], [50, 50], run_types=["single", "single"], sum_window=S_,
                           mid_infill="nearest", final_infill="none")
    out = mp.run(torch.as_tensor(A, device=dev), torch.as_tensor(B, device=dev), batch=3)
    torch.cuda.synchronize()
    oref = EO.ensemble_multipass(A, B, mp.grids, floor_fn=_floor_fn(mp), mid_infill="nearest",
                                 final_infill="none", run_types=["single", "single"], sum_window=S_)
    for p, (r, o) in enumerate(zip(out, oref)):
        st, ost = r.status.cpu().numpy(), o["fit"]["status"]
        assert np.mean(st != ost) < 0.02, p
        ux, uy = r.ux.cpu().numpy(), r.uy.cpu().numpy()
        both = np.isfinite(ux) & np.isfinite(o["ux"]) & (st.reshape(ux.shape) == ost.reshape(ux.shape))
        assert both.mean() > 0.8
        d = np.hypot(ux[both] - o["ux"][both], uy[both] - o["uy"][both])
        assert np.median(d) < 1e-4 and np.mean(d < 2e-3) > 0.95, (p, np.median(d), np.mean(d < 2e-3))
