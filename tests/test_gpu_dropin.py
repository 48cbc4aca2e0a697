"""Device parity of the drop-in libraries against the oracle and the
reference's own C outputs (golden fixtures).

Bar (per the reference's correctness gate): identical integer peak
indices and valid (NaN) masks, correlation planes within 1e-5 relative L2,
displacements within 0.005 px.  The warp and the peak fitter are held to
bit-identity / fixture agreement.
"""

import os

import numpy as np
import pytest

import oracle as O
from gpu_util import dropin_bulkxcorr2d
from pivtools_amd import _native
from pivtools_amd import synthetic as S
from pivtools_amd.window import compute_window_centers, window_weight

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
SIZES = [8, 12, 16, 24, 32, 48, 64, 96, 128]


def _pair(seed, shape, shift=(2.3, -1.4), ppp=0.04):
    rng = np.random.default_rng(seed)
    x, y, inten = S._positions(rng, shape, ppp)
    a = S.render(shape, x, y, inten).numpy()
    b = S.render(shape, x + shift[0], y + shift[1], inten).numpy()
    a += rng.normal(0, 0.01, shape).astype(np.float32)
    b += rng.normal(0, 0.01, shape).astype(np.float32)
    return a.astype(np.float32), b.astype(np.float32)


def _check_outputs(got, ref, planes=True):
    assert got["err"] == 0 and ref["err"] == 0
    gx, rx = got["pk_x"], ref["pk_x"]
    nan_g, nan_r = np.isnan(gx), np.isnan(rx)
    assert np.array_equal(nan_g, nan_r), f"NaN masks differ at {np.argwhere(nan_g != nan_r)[:5]}"
    ok = ~nan_r
    dx = np.abs(gx[ok] - rx[ok])
    dy = np.abs(got["pk_y"][ok] - ref["pk_y"][ok])
    assert dx.max(initial=0) < 0.005 and dy.max(initial=0) < 0.005, (dx.max(initial=0), dy.max(initial=0))
    np.testing.assert_allclose(got["pk_h"], ref["pk_h"], rtol=2e-5, atol=1e-7)
    np.testing.assert_allclose(got["ratio"], ref["ratio"], rtol=2e-5, atol=1e-7)
    if planes and "planes" in got:
        g = got["planes"].reshape(-1, *got["planes"].shape[-2:]).astype(np.float64)
        r = ref["planes"].reshape(-1, *ref["planes"].shape[-2:]).astype(np.float64)
        num = np.linalg.norm((g - r).reshape(len(g), -1), axis=1)
        den = np.linalg.norm(r.reshape(len(r), -1), axis=1)
        live = den > 0
        assert (num[live] / den[live]).max(initial=0) < 1e-5
        gi = np.argmax(g.reshape(len(g), -1), axis=1)
        ri = np.argmax(r.reshape(len(r), -1), axis=1)
        assert np.array_equal(gi[live], ri[live])


@pytest.mark.parametrize("n", SIZES)
def test_bulkxcorr2d_gauss3_all_sizes(n):
    shape = (3 * n + 7, 4 * n + 5)
    a, b = _pair(n, shape)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    mask[0, 1] = 1.0
    got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    _check_outputs(got, ref)


def _refit_on_planes(planes, w, finder, n_peaks=1, libm="cr"):
    """Oracle fit of the device's own planes: separates the fitter from the
    last-ulp FFT differences that feed it.  libm="cr" is the oracle with
    correctly rounded expf/logf, the device fitter's definition (exact
    agreement expected); libm="host" is the reference's host-libm flavour."""
    pw = O.plane_weight(w)
    flat = planes.reshape(-1, planes.shape[-2], planes.shape[-1])
    loc = np.array([O.lsqpeaklocate_lm(p, n_peaks, finder, pw, libm=libm)[0] for p in flat])
    return loc


@pytest.mark.parametrize("finder", [4, 5, 6])
@pytest.mark.parametrize("impl", [0, 1])
@pytest.mark.parametrize("n", [16, 32, 64])
def test_bulkxcorr2d_lm_fits(finder, impl, n):
    """LM trust decisions are chaotic on marginal windows (the reference's
    own scalar-vs-batch gate tolerates flips, unit-tests/test_peakfit_batch.py:226-289),
    so parity is split: (1) the device fit of the device plane equals the
    oracle fit of that same plane -- identical NaN masks, positions within
    1e-4 px; (2) end to end against the oracle, displacements agree within
    0.005 px where both are valid and the NaN-flip rate stays small."""
    L = _native.bulkxcorr2d_lib()
    assert L.bulkxcorr2d_set_peakfit_impl(impl) == 0
    try:
        shape = (5 * n, 5 * n)
        a, b = _pair(100 + n, shape, shift=(1.7, 0.6))
        wc = compute_window_centers(shape, (n, n), 50)
        w = window_weight((n, n), "gaussian")
        mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, finder,
                                 want_planes=True)
        ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, finder, want_planes=True)
    finally:
        L.bulkxcorr2d_set_peakfit_impl(0)
    gx = got["pk_x"].ravel()
    if impl == 0:
        loc = _refit_on_planes(got["planes"], w, finder)
        rx = loc[:, 1] - n / 2.0
        assert np.array_equal(np.isnan(gx), np.isnan(rx))
        ok = ~np.isnan(rx)
        assert np.abs(gx[ok] - rx[ok]).max(initial=0) < 1e-4
        # host libm: its expf/logf are not correctly rounded for every input,
        # so a window whose trust test sits on the threshold may flip (glibc
        # vs correctly rounded differ in 2 of the 768 reference fixture fits)
        hx = _refit_on_planes(got["planes"], w, finder, libm="host")[:, 1] - n / 2.0
        assert np.sum(np.isnan(gx) != np.isnan(hx)) <= 1
        both = ~np.isnan(gx) & ~np.isnan(hx)
        assert np.median(np.abs(gx[both] - hx[both])) < 1e-6
    ex = ref["pk_x"].ravel()
    both = ~np.isnan(ex) & ~np.isnan(gx)
    tol = 0.005 if impl == 0 else 1e-3 * 5
    assert np.abs(gx[both] - ex[both]).max(initial=0) < tol
    assert np.mean(np.isnan(ex) != np.isnan(gx)) < 0.06


def test_bulkxcorr2d_untouched_masked_and_oob():
    n = 32
    shape = (128, 160)
    a, b = _pair(7, shape)
    ctrs_x = np.array([-20.0, 15.5, 47.5, 79.5, 111.5, 143.5, 200.0], np.float32)
    ctrs_y = np.array([15.5, 47.5, 79.5, 111.5, 140.0], np.float32)
    mask = np.zeros((len(ctrs_y), len(ctrs_x)), np.float32)
    mask[2, 3] = 1.0
    w = window_weight((n, n), "gaussian")
    got = dropin_bulkxcorr2d(a, b, mask, ctrs_x, ctrs_y, (n, n), w, w, 1, 3, prefill=123.0)
    ref = O.bulkxcorr2d(a, b, mask, ctrs_x, ctrs_y, (n, n), w, w, 1, 3)
    skipped = np.zeros_like(mask, bool)
    skipped[:, 0] = skipped[:, -1] = True
    skipped[-1, :] = True
    skipped[2, 3] = True
    for k in ("pk_x", "pk_y", "pk_h", "sx", "sy", "sxy"):
        assert np.all(got[k][0, 0][skipped] == 123.0), k
    assert np.all(got["ratio"][0][skipped] == 123.0)
    live = ~skipped
    np.testing.assert_allclose(got["pk_x"][0, 0][live], ref["pk_x"][0, 0][live], atol=0.005)


def test_bulkxcorr2d_multi_peak():
    n = 32
    shape = (160, 160)
    a, b = _pair(11, shape)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 3, 6, want_planes=True)
    ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 3, 6)
    # the device's planes refitted by the oracle give the device's answers
    loc = _refit_on_planes(got["planes"], w, 6, n_peaks=3)
    gx = got["pk_x"][0].reshape(3, -1).T          # (windows, peaks)
    rx = loc[:, 3:6] - n / 2.0                     # col rows of the (3, nPeaks) layout
    assert np.array_equal(np.isnan(gx), np.isnan(rx))
    ok = ~np.isnan(rx)
    assert np.abs(gx[ok] - rx[ok]).max(initial=0) < 1e-4
    # gauss3 later peaks agree end to end wherever the first peak agrees
    e = ref["pk_x"][0].reshape(3, -1).T
    same = np.isnan(e[:, 0]) == np.isnan(gx[:, 0])
    assert np.mean(same) > 0.94


def test_bulkxcorr2d_batch_of_images_and_ensemble_planes():
    n = 16
    shape = (96, 112)
    pairs = [_pair(20 + k, shape) for k in range(3)]
    A = np.stack([p[0] for p in pairs])
    B = np.stack([p[1] for p in pairs])
    wc = compute_window_centers(shape, (n, n), 75)
    w = window_weight((n, n), "square")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    got = dropin_bulkxcorr2d(A, B, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    ref = O.bulkxcorr2d(A, B, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    _check_outputs(got, ref)
    gE = dropin_bulkxcorr2d(A, B, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, ensemble=True,
                            want_planes=True)
    rE = O.bulkxcorr2d(A, B, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, ensemble=True,
                       want_planes=True)
    g = gE["planes"].astype(np.float64)
    r = rE["planes"].astype(np.float64)
    assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-5


def test_bulkxcorr2d_single_pass_1024_vs_fp64():
    """The 1024 x 1024 single-pass workload (32 x 32 at 50 %, gaussian
    weight, gauss3): oracle parity plus the error against FP64."""
    a, b = S.single_pass_pair(seed=0)
    a, b = a.numpy(), b.numpy()
    wc = compute_window_centers(a.shape, (32, 32), 50)
    assert (wc.n_win_y, wc.n_win_x) == (63, 63)
    w = window_weight((32, 32), "gaussian")
    mask = np.zeros((63, 63), np.float32)
    got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (32, 32), w, w, 1, 3, want_planes=True)
    ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (32, 32), w, w, 1, 3, want_planes=True)
    _check_outputs(got, ref)
    p64, _ = O.planes_fp64(a, b, wc.win_ctrs_x, wc.win_ctrs_y, (32, 32), w.astype(np.float64), w.astype(np.float64))
    g = got["planes"].reshape(-1, 32, 32).astype(np.float64)
    rel = np.linalg.norm(g - p64) / np.linalg.norm(p64)
    assert rel < 1e-6


def test_unsupported_window_size_returns_4():
    a, b = _pair(1, (64, 64))
    w = np.ones((20, 20), np.float32)
    got = dropin_bulkxcorr2d(a, b, np.zeros((1, 1), np.float32), np.array([31.5], np.float32),
                             np.array([31.5], np.float32), (20, 20), w, w, 1, 3)
    assert got["err"] == 4


@pytest.mark.parametrize("ft", [3, 4, 5, 6])
@pytest.mark.parametrize("n", [16, 32, 64])
def test_lsqpeaklocate_lm_matches_reference_fixtures(ft, n):
    g = np.load(os.path.join(GOLD, "peakfit_ref.npz"))
    L = _native.bulkxcorr2d_lib()
    planes, w = g[f"n{n}_planes"], g[f"n{n}_weight"]
    ref_loc, ref_std = g[f"n{n}_ft{ft}_loc"], g[f"n{n}_ft{ft}_std"]
    N = np.array([n, n], np.int32)
    nan_flip = 0
    for k, pl in enumerate(planes):
        loc = np.zeros(3, np.float32)
        std = np.zeros(3, np.float32)
        L.lsqpeaklocate_lm(np.ascontiguousarray(pl), N, loc, 1, ft, std, w.ctypes.data)
        if np.isnan(loc[0]) != np.isnan(ref_loc[k][0]):
            nan_flip += 1
            continue
        if not np.isnan(loc[0]):
            assert abs(loc[0] - ref_loc[k][0]) < 1e-4 and abs(loc[1] - ref_loc[k][1]) < 1e-4, (k, loc, ref_loc[k])
            assert loc[2] == ref_loc[k][2]
    assert nan_flip == 0
    r = np.array([L.peak_detectability(np.ascontiguousarray(p), N) for p in planes], np.float32)
    np.testing.assert_array_equal(r, g[f"n{n}_ratio"])


@pytest.mark.parametrize("ft", [4, 5, 6])
@pytest.mark.parametrize("n", [16, 32, 64])
def test_lsqpeaklocate_lm_batch_matches_reference_batch(ft, n):
    """The batch flavour (the reference's config default) at the scalar bar:
    identical NaN masks and positions within 1e-4 px of the reference C's
    own batch fitter (peak_locate_lm_batch.c, fixtures from oracle/_ref)."""
    g = np.load(os.path.join(GOLD, "peakfit_ref.npz"))
    L = _native.bulkxcorr2d_lib()
    W = L.peakfit_batch_lanes()
    assert W == 16
    planes, w = g[f"n{n}_planes"], g[f"n{n}_weight"]
    ref = g[f"n{n}_ft{ft}_batch_loc"]
    got = np.zeros((len(planes), 3), np.float32)
    for s in range(0, len(planes), W):
        chunk = np.ascontiguousarray(planes[s:s + W])
        loc = np.zeros(3 * W, np.float32)
        std = np.zeros(3 * W, np.float32)
        L.lsqpeaklocate_lm_batch(chunk.reshape(-1), len(chunk), np.array([n, n], np.int32), ft, loc, std,
                                 w.ctypes.data)
        got[s:s + len(chunk)] = loc.reshape(3, W)[:, :len(chunk)].T
    np.testing.assert_array_equal(np.isnan(got[:, 0]), np.isnan(ref[:, 0]))
    ok = ~np.isnan(ref[:, 0])
    assert np.abs(got[ok, :2] - ref[ok, :2]).max(initial=0) < 1e-4
    np.testing.assert_array_equal(got[ok, 2], ref[ok, 2])


def test_lsqpeaklocate_lm_batch_writes_exactly_three_lane_widths():
    """A caller that sizes its outputs by PEAKFIT_MAX_LANES (3 x 16 floats,
    PIV_2d_cross_correlate.c:149) must find the guard words behind them
    untouched, also when L_real exceeds the lane width."""
    g = np.load(os.path.join(GOLD, "peakfit_ref.npz"))
    L = _native.bulkxcorr2d_lib()
    W = L.peakfit_batch_lanes()
    assert W <= 16 and L.codelet_lanes() == W
    n = 32
    planes = np.ascontiguousarray(g[f"n{n}_planes"][:W + 4])
    canary = np.float32(-12345.5)
    for l_real in (3, W, W + 4):
        loc = np.full(3 * 16 + 8, canary, np.float32)
        std = np.full(3 * 16 + 8, canary, np.float32)
        L.lsqpeaklocate_lm_batch(planes.reshape(-1), l_real, np.array([n, n], np.int32), 5, loc, std,
                                 g[f"n{n}_weight"].ctypes.data)
        assert (loc[3 * W:] == canary).all() and (std[3 * W:] == canary).all(), l_real
        assert not (loc[:3 * W] == canary).any()
        lanes = loc[:3 * W].reshape(3, W)
        assert np.isnan(lanes[0, min(l_real, W):]).all()
        assert (lanes[2, min(l_real, W):] == 0).all()


@pytest.mark.parametrize("ft", [3, 4, 5, 6])
@pytest.mark.parametrize("shape", [(17, 17), (33, 33), (65, 65), (33, 65)])
def test_lsqpeaklocate_lm_closed_form(ft, shape):
    """unit-tests/test_instantaneous_peaks.py:173-234 through the drop-in."""
    L = _native.bulkxcorr2d_lib()
    h, w = shape
    ii, jj = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    for i0, j0 in ((0.0, 0.0), (0.35, 0.27), (0.1, -0.4)):
        di, dj = ii - (h - 1) / 2 - i0, jj - (w - 1) / 2 - j0
        if ft == 6:
            pl = 1000 * np.exp(-0.5 * (di * di / 2.25 + dj * dj / 2.25))
        else:
            pl = 1000 * np.exp(-(di * di + dj * dj) / 2.25)
        loc = np.zeros(3, np.float32)
        std = np.zeros(3, np.float32)
        L.lsqpeaklocate_lm(pl.astype(np.float32), np.array(shape, np.int32), loc, 1, ft, std, None)
        assert abs(loc[0] - (h - 1) / 2 - i0) < 0.01 and abs(loc[1] - (w - 1) / 2 - j0) < 0.01


def test_fused_warp_matches_reference_fixtures_bitwise():
    g = np.load(os.path.join(GOLD, "warp_ref.npz"))
    L = _native.fusedwarp_lib()
    A, B = g["A"], g["B"]
    N, H, W = A.shape
    for mode in (0, 1):
        for shared in (0, 1):
            for rnd in (0, 1):
                dy = np.ascontiguousarray(g["dy"][0] if shared else g["dy"])
                dx = np.ascontiguousarray(g["dx"][0] if shared else g["dx"])
                oa = np.zeros_like(A)
                ob = np.zeros_like(B)
                r = L.fused_symmetric_warp_batch(A, B, oa, ob, dy, dx, N, H, W, len(g["cy"]), len(g["cx"]),
                                                 g["cy"], g["cx"], mode, shared, rnd)
                assert r == 0
                assert np.array_equal(oa.view(np.int32), g[f"m{mode}_s{shared}_r{rnd}_a"].view(np.int32))
                assert np.array_equal(ob.view(np.int32), g[f"m{mode}_s{shared}_r{rnd}_b"].view(np.int32))


def test_fused_warp_large_vs_oracle_bitwise():
    rng = np.random.default_rng(9)
    N, H, W = 2, 300, 517
    A = rng.uniform(0, 4095, (N, H, W)).astype(np.float32)
    B = rng.uniform(0, 4095, (N, H, W)).astype(np.float32)
    wc = compute_window_centers((H, W), (32, 32), 50)
    cy = np.concatenate([[0.0], wc.win_ctrs_y, [H - 1.0]]).astype(np.float32)
    cx = np.concatenate([[0.0], wc.win_ctrs_x, [W - 1.0]]).astype(np.float32)
    dy = rng.uniform(-6, 6, (N, len(cy), len(cx))).astype(np.float32)
    dx = rng.uniform(-6, 6, (N, len(cy), len(cx))).astype(np.float32)
    L = _native.fusedwarp_lib()
    for mode in (0, 1):
        oa = np.zeros_like(A)
        ob = np.zeros_like(B)
        assert L.fused_symmetric_warp_batch(A, B, oa, ob, dy, dx, N, H, W, len(cy), len(cx), cy, cx, mode, 0, 0) == 0
        ra, rb = O.fused_warp_batch(A, B, dy, dx, cy, cx, mode, False, False)
        assert np.array_equal(oa.view(np.int32), ra.view(np.int32))
        assert np.array_equal(ob.view(np.int32), rb.view(np.int32))


def test_convolve_and_xcorr_dropins():
    L = _native.bulkxcorr2d_lib()
    rng = np.random.default_rng(4)
    for n in (8, 32, 96):
        w = window_weight((n, n), "gaussian")
        conv = np.zeros_like(w)
        assert L.convolve(w, w, conv, np.array([n, n], np.int32)) == 0
        ref = np.zeros_like(w)
        O.lib().orc_convolve(w, w, ref, np.array([n, n], np.int32))
        np.testing.assert_allclose(conv, ref, rtol=1e-6)
        a = rng.normal(size=(n, n)).astype(np.float32)
        b = rng.normal(size=(n, n)).astype(np.float32)
        c = np.zeros_like(a)
        assert L.xcorr(a, b, c, np.array([n, n], np.int32)) == 0
        # xcorr(a, b) = shift(IFFT(FFT(a) conj(FFT(b)))) / numel
        fa, fb = np.fft.fft2(a.astype(np.float64)), np.fft.fft2(b.astype(np.float64))
        r64 = np.fft.fftshift(np.real(np.fft.ifft2(fa * np.conj(fb))))
        assert np.linalg.norm(c - r64) / np.linalg.norm(r64) < 1e-5
    assert L.xcorr(np.zeros((20, 20), np.float32), np.zeros((20, 20), np.float32),
                   np.zeros((20, 20), np.float32), np.array([20, 20], np.int32)) == 4


def test_plane_weight_cache_follows_the_weight():
    """bulkxcorr2d recomputes the plane weight only when wB changes: three
    calls of one size with weights gaussian, top-hat, gaussian each match
    the oracle for their own weight."""
    n = 32
    shape = (160, 192)
    a, b = _pair(41, shape)
    wc = compute_window_centers(shape, (n, n), 50)
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    for w in (window_weight((n, n), "gaussian"), np.ones((n, n), np.float32),
              window_weight((n, n), "gaussian")):
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3)
        ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3)
        _check_outputs(got, ref)
        ok = ~np.isnan(ref["pk_x"])
        # the widths see the plane weight's curvature across the patch
        np.testing.assert_allclose(got["sx"][ok], ref["sx"][ok], rtol=1e-4)
        np.testing.assert_allclose(got["sy"][ok], ref["sy"][ok], rtol=1e-4)


RECT = [(32, 64), (64, 32), (16, 48), (128, 96), (24, 8), (96, 128)]


@pytest.mark.parametrize("win", RECT)
def test_bulkxcorr2d_rectangular_windows(win):
    """Rectangular windows (row length W and column length H bound
    independently, codelet_fft.c:52-70): identical NaN masks and integer
    peaks, planes within 1e-5 relative L2 of the oracle and 5e-6 of FP64,
    displacements within 0.005 px; a masked window is left untouched."""
    shape = (3 * win[0] + 11, 3 * win[1] + 7)
    a, b = _pair(sum(win), shape, shift=(1.3, -0.7))
    wc = compute_window_centers(shape, win, 50)
    w = window_weight(win, "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    mask[0, 1] = 1.0
    got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, win, w, w, 1, 3, want_planes=True,
                             prefill=-3.0)
    ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, win, w, w, 1, 3, want_planes=True)
    assert got["err"] == 0
    assert (got["pk_x"][0, 0, 0, 1] == -3.0) and (got["planes"][0, 0, 1] == 0).all()
    got["pk_x"][0, 0, 0, 1] = ref["pk_x"][0, 0, 0, 1]
    for k in ("pk_y", "pk_h", "sx", "sy", "sxy"):
        got[k][0, 0, 0, 1] = ref[k][0, 0, 0, 1]
    got["ratio"][0, 0, 1] = ref["ratio"][0, 0, 1]
    _check_outputs(got, ref)
    gi = np.argmax(got["planes"].reshape(-1, win[0] * win[1]), axis=1)
    ri = np.argmax(ref["planes"].reshape(-1, win[0] * win[1]), axis=1)
    live = mask.reshape(-1) == 0
    assert np.array_equal(gi[live], ri[live])
    p64, _ = O.planes_fp64(a, b, wc.win_ctrs_x, wc.win_ctrs_y, win, w.astype(np.float64), w.astype(np.float64))
    g = got["planes"].reshape(-1, *win).astype(np.float64)[live]
    assert np.linalg.norm(g - p64[live]) / np.linalg.norm(p64[live]) < 5e-6


@pytest.mark.parametrize("finder,n_peaks", [(6, 1), (5, 1), (3, 3)])
def test_bulkxcorr2d_rectangular_fits_and_peaks(finder, n_peaks):
    """LM fits and several peaks on rectangular windows: the device fit of
    the device plane equals the oracle fit of that plane (identical NaN
    masks, 1e-4 px); end to end within 0.005 px of the oracle."""
    win = (32, 64)
    shape = (140, 260)
    a, b = _pair(77, shape, shift=(0.6, 2.2))
    wc = compute_window_centers(shape, win, 50)
    w = window_weight(win, "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, win, w, w, n_peaks, finder, want_planes=True)
    ref = O.bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, win, w, w, n_peaks, finder, want_planes=True)
    assert got["err"] == 0
    loc = _refit_on_planes(got["planes"], w, finder, n_peaks)
    gx = got["pk_x"][:, 0].ravel()
    rx = loc[:, n_peaks] - win[1] / 2.0     # column of the first peak
    assert np.array_equal(np.isnan(gx), np.isnan(rx))
    ok = ~np.isnan(rx)
    assert np.abs(gx[ok] - rx[ok]).max(initial=0) < 1e-4
    ex = ref["pk_x"][:, 0].ravel()
    both = ~np.isnan(ex) & ~np.isnan(gx)
    assert np.abs(gx[both] - ex[both]).max(initial=0) < 0.005
    assert np.mean(np.isnan(ex) != np.isnan(gx)) < 0.06
