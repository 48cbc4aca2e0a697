"""Pins of the CPU oracle (no GPU).

Each oracle function is checked against something that does not restate
its own formula: the reference's compiled C (golden fixtures written by
tests/golden/make_golden.py from oracle/_ref), brute-force FP64 definitions,
closed forms, and known-answer planes of the reference's own tests
(unit-tests/test_instantaneous_peaks.py:173-560).
"""

import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def peak_gold():
    return np.load(os.path.join(GOLD, "peakfit_ref.npz"))


@pytest.fixture(scope="module")
def warp_gold():
    return np.load(os.path.join(GOLD, "warp_ref.npz"))


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.mark.parametrize("n", [16, 32, 64])
@pytest.mark.parametrize("ft", [3, 4, 5, 6])
def test_oracle_peakfit_matches_reference_c_bitwise(peak_gold, n, ft):
    planes = peak_gold[f"n{n}_planes"]
    w = peak_gold[f"n{n}_weight"]
    ref_loc = peak_gold[f"n{n}_ft{ft}_loc"]
    ref_std = peak_gold[f"n{n}_ft{ft}_std"]
    for k, pl in enumerate(planes):
        loc, std = O.lsqpeaklocate_lm(pl, 1, ft, w)
        assert np.array_equal(_bits(loc), _bits(ref_loc[k])), (k, loc, ref_loc[k])
        assert np.array_equal(_bits(std), _bits(ref_std[k])), (k, std, ref_std[k])


def test_cr_libm_flavour_against_reference_c(peak_gold):
    """The oracle with correctly rounded expf/logf (the device fitter's
    definition) on the reference C fixtures: identical NaN masks, positions
    within 1e-5 px, and bitwise equal except where the host libm rounds an
    exponential or logarithm differently (2 of the 768 fits)."""
    n_diff = 0
    for n in (16, 32, 64):
        planes, w = peak_gold[f"n{n}_planes"], peak_gold[f"n{n}_weight"]
        for ft in (3, 4, 5, 6):
            ref_loc = peak_gold[f"n{n}_ft{ft}_loc"]
            for k, pl in enumerate(planes):
                loc, _ = O.lsqpeaklocate_lm(pl, 1, ft, w, libm="cr")
                assert np.array_equal(np.isnan(loc), np.isnan(ref_loc[k]))
                ok = ~np.isnan(loc)
                assert np.all(np.abs(loc[ok] - ref_loc[k][ok]) < 1e-5)
                n_diff += not np.array_equal(_bits(loc), _bits(ref_loc[k]))
    assert n_diff <= 4


@pytest.mark.parametrize("n", [16, 32, 64])
def test_oracle_detectability_matches_reference_c(peak_gold, n):
    planes = peak_gold[f"n{n}_planes"]
    got = np.array([O.peak_detectability(p) for p in planes], np.float32)
    assert np.array_equal(_bits(got), _bits(peak_gold[f"n{n}_ratio"]))


def test_oracle_warp_matches_reference_c_bitwise(warp_gold):
    g = warp_gold
    for mode in (0, 1):
        for shared in (0, 1):
            for rnd in (0, 1):
                dy = g["dy"][0] if shared else g["dy"]
                dx = g["dx"][0] if shared else g["dx"]
                oa, ob = O.fused_warp_batch(g["A"], g["B"], dy, dx, g["cy"], g["cx"], mode, shared, rnd)
                assert np.array_equal(_bits(oa), _bits(g[f"m{mode}_s{shared}_r{rnd}_a"]))
                assert np.array_equal(_bits(ob), _bits(g[f"m{mode}_s{shared}_r{rnd}_b"]))


def test_warp_integer_shift_is_exact_roll():
    """round_shifts with an even-integer uniform predictor is a pure pixel roll
    (unit-tests/test_ensemble_options.py:219-290)."""
    rng = np.random.default_rng(3)
    A = rng.uniform(0, 1, (1, 40, 48)).astype(np.float32)
    cy = np.array([0, 20, 39], np.float32)
    cx = np.array([0, 24, 47], np.float32)
    dy = np.full((3, 3), 4.0, np.float32)
    dx = np.full((3, 3), -2.0, np.float32)
    oa, ob = O.fused_warp_batch(A, A, dy, dx, cy, cx, 0, True, True)
    # A sampled at (i - 2, j + 1), B at (i + 2, j - 1)
    assert np.array_equal(oa[0, 2:, :-1], A[0, :-2, 1:])
    assert np.array_equal(ob[0, :-2, 1:], A[0, 2:, :-1])


@pytest.mark.parametrize("n", [8, 12, 16])
def test_plane_fp64_equals_bruteforce(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n))
    b = rng.normal(size=(n, n))
    np.testing.assert_allclose(O.xcorr_plane_fp64(a, b), O.xcorr_plane_bruteforce(a, b), atol=1e-10)


@pytest.mark.parametrize("n", [8, 12, 16, 24, 32, 48, 64, 96, 128])
def test_oracle_plane_vs_fp64(n):
    rng = np.random.default_rng(100 + n)
    H = W = 2 * n
    A = rng.uniform(0, 1, (H, W)).astype(np.float32)
    B = (np.roll(A, (1, 2), (0, 1)) + rng.normal(0, 0.05, (H, W))).astype(np.float32)
    wa = np.ones((n, n), np.float32)
    c = np.array([n - 0.5], np.float32)
    o = O.bulkxcorr2d(A, B, np.zeros((1, 1), np.float32), c, c, (n, n), wa, wa, 1, 3, want_planes=True)
    p64, _ = O.planes_fp64(A, B, c, c, (n, n), np.ones((n, n)), np.ones((n, n)))
    pl = o["planes"][0, 0, 0].astype(np.float64)
    assert np.linalg.norm(pl - p64[0]) / np.linalg.norm(p64[0]) < 1e-6
    assert np.unravel_index(np.argmax(pl), pl.shape) == (n // 2 + 1, n // 2 + 2)


@pytest.mark.parametrize("n", [8, 12, 32, 128])
def test_plane_weight_square_closed_form(n):
    pw = O.plane_weight(np.ones((n, n), np.float32))
    m = np.arange(n) - n // 2
    cf = n * n / np.outer(n - np.abs(m), n - np.abs(m))
    np.testing.assert_allclose(pw, cf, rtol=1e-6)


def test_plane_weight_unsupported_size():
    with pytest.raises(ValueError):
        O.plane_weight(np.ones((20, 20), np.float32))


def _gauss(shape, amp, i0, j0, s):
    h, w = shape
    ii, jj = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    di, dj = ii - (h - 1) / 2 - i0, jj - (w - 1) / 2 - j0
    return (amp * np.exp(-(di * di + dj * dj) / (s * s))).astype(np.float32)


@pytest.mark.parametrize("ft", [3, 4, 5, 6])
@pytest.mark.parametrize("shape", [(17, 17), (33, 33), (65, 65), (33, 65)])
@pytest.mark.parametrize("disp", [(0.0, 0.0), (0.35, 0.27), (0.1, -0.4)])
def test_oracle_closed_form_gaussian_position(ft, shape, disp):
    """unit-tests/test_instantaneous_peaks.py:173-234 (< 0.01 px)."""
    if ft == 6:
        h, w = shape
        ii, jj = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        di, dj = ii - (h - 1) / 2 - disp[0], jj - (w - 1) / 2 - disp[1]
        pl = (1000 * np.exp(-0.5 * (di * di / 2.25 + dj * dj / 2.25))).astype(np.float32)
    else:
        pl = _gauss(shape, 1000.0, disp[0], disp[1], 1.5)
    loc, _ = O.lsqpeaklocate_lm(pl, 1, ft)
    assert abs(loc[0] - (shape[0] - 1) / 2 - disp[0]) < 0.01
    assert abs(loc[1] - (shape[1] - 1) / 2 - disp[1]) < 0.01


def test_oracle_detectability_known_answers():
    """unit-tests/test_instantaneous_peaks.py:507-560."""
    z = lambda: np.zeros((33, 33), np.float32)  # noqa: E731
    p = z(); p[16, 16] = 10; p[16, 20] = 4
    assert O.peak_detectability(p) == pytest.approx(2.5)
    p = z(); p[16, 16] = 10
    assert O.peak_detectability(p) == 0.0
    assert O.peak_detectability(np.full((33, 33), -1, np.float32)) == 0.0
    p = z(); p[12, 12] = 7; p[20, 20] = 7
    assert O.peak_detectability(p) == pytest.approx(1.0)
    p = z(); p[16, 16] = 10; p[16, 17] = 9; p[16, 21] = 4
    assert O.peak_detectability(p) == pytest.approx(2.5)
    p = z(); p[16, 16] = 10; p[2, 2] = 9
    assert O.peak_detectability(p) == 0.0


@pytest.mark.parametrize("ft", [3, 4, 5, 6])
def test_oracle_nan_sentinel_flat(ft):
    loc, std = O.lsqpeaklocate_lm(np.zeros((33, 33), np.float32), 1, ft)
    assert np.isnan(loc[0]) and np.isnan(loc[1]) and loc[2] == 0 and np.all(std == 0)


def test_oracle_bulkxcorr_recovers_uniform_shift():
    """A single-pass correlation of an integer-shifted particle pair finds
    the shift (the integer-shift gate of pivtools_cli/lib/test_codelet_gate.c:108-134)."""
    import pivtools_amd.synthetic as S
    import torch

    rng = np.random.default_rng(5)
    x, y, inten = S._positions(rng, (128, 128), 0.05)
    a = S.render((128, 128), x, y, inten).numpy()
    b = S.render((128, 128), x + 3.0, y - 2.0, inten).numpy()
    del torch
    from pivtools_amd.window import compute_window_centers, window_weight
    wc = compute_window_centers((128, 128), (32, 32), 50)
    w = window_weight((32, 32), "gaussian")
    o = O.bulkxcorr2d(a, b, np.zeros((wc.n_win_y, wc.n_win_x), np.float32), wc.win_ctrs_x, wc.win_ctrs_y,
                      (32, 32), w, w, 1, 3)
    inner = (slice(1, -1), slice(1, -1))
    assert np.nanmedian(np.abs(o["pk_x"][0, 0][inner] - 3.0)) < 0.05
    assert np.nanmedian(np.abs(o["pk_y"][0, 0][inner] + 2.0)) < 0.05
