"""Ensemble 'single' run type on the host: grid, padding and weights against
the invariants the reference's own tests state
(unit-tests/test_single_mode_frame_a_sizes.py, window_utils.py docstrings),
and the oracle chain pinned by brute force and closed forms:

* the correlation of a padded pair with the singlepix / bsingle weights
  equals the direct sums over the A block and the sum window;
* the wA * wB pair envelope equals the closed-form trapezoid;
* identical frames give a unit normalised AB peak only with the
  sum-area / A-area scale;
* a uniform displacement is recovered by the whole chain.
"""

import numpy as np
import pytest

import ensemble_oracle as EO
import oracle as O
from pivtools_amd import floor_psd
from pivtools_amd import synthetic as S
from pivtools_amd.window import (build_pass_grids, compute_padding_for_single_mode,
                                 compute_window_centers_single_mode, window_centers_original, window_weight)

SUM = 48
IMAGE_SHAPE = (512, 512)
A_SIZES = (4, 6)


# ---------------------------------------------------------------- weights
@pytest.mark.parametrize("m", A_SIZES)
def test_singlepix_is_a_block_of_ones_in_the_sum_window(m):
    wa = window_weight((m, m), "singlepix", (SUM, SUM))
    assert wa.shape == (SUM, SUM) and wa.dtype == np.float32
    assert wa.sum() == m * m
    assert set(np.unique(wa)) == {0.0, 1.0}
    rows, cols = np.flatnonzero(wa.any(axis=1)), np.flatnonzero(wa.any(axis=0))
    assert len(rows) == m and len(cols) == m
    assert rows[-1] - rows[0] == m - 1 and cols[-1] - cols[0] == m - 1
    r, c = np.nonzero(wa)
    assert r.mean() == pytest.approx((SUM - 1) / 2) and c.mean() == pytest.approx((SUM - 1) / 2)


def test_singlepix_rectangular_and_bsingle():
    wa = window_weight((4, 6), "singlepix", (SUM, SUM))
    assert wa.sum() == 24.0
    assert len(np.flatnonzero(wa.any(axis=1))) == 4 and len(np.flatnonzero(wa.any(axis=0))) == 6
    # odd difference: the block starts at ceil((sum - m) / 2)
    w5 = window_weight((4, 4), "singlepix", (17, 17))
    assert np.flatnonzero(w5.any(axis=1))[0] == 7
    wb = window_weight((4, 4), "bsingle", (SUM, SUM))
    assert wb.shape == (SUM, SUM) and np.all(wb == 1.0)
    with pytest.raises(ValueError):
        window_weight((4, 4), "singlepix")


# ---------------------------------------------------------------- padding
@pytest.mark.parametrize("m,expected", [(4, 22), (6, 21), (16, 16)])
def test_symmetric_padding_for_even_sizes(m, expected):
    assert compute_padding_for_single_mode((m, m), (SUM, SUM)) == (expected,) * 4


def test_padding_ceil_top_left_and_spans_the_sum_window():
    assert compute_padding_for_single_mode((4, 4), (16, 16)) == (6, 6, 6, 6)
    assert compute_padding_for_single_mode((4, 4), (17, 17)) == (7, 6, 7, 6)
    for m in A_SIZES:
        t, b, l, r = compute_padding_for_single_mode((m, m), (SUM, SUM))
        assert t + m + b == SUM and l + m + r == SUM
    with pytest.raises(ValueError):
        compute_padding_for_single_mode((32, 32), (16, 16))


# ---------------------------------------------------------------- grid
@pytest.mark.parametrize("m", A_SIZES)
def test_spacing_follows_the_small_window(m):
    r0 = compute_window_centers_single_mode(IMAGE_SHAPE, (m, m), (SUM, SUM), 0)
    assert r0.win_spacing_x == m and r0.win_spacing_y == m
    assert len(r0.win_ctrs_x) == r0.n_win_x and len(r0.win_ctrs_y) == r0.n_win_y
    assert np.allclose(np.diff(r0.win_ctrs_x), m) and np.allclose(np.diff(r0.win_ctrs_y), m)
    r50 = compute_window_centers_single_mode(IMAGE_SHAPE, (m, m), (SUM, SUM), 50)
    assert r50.win_spacing_x == m // 2 and r50.win_spacing_y == m // 2


def test_docstring_grid_and_count_scaling():
    r = compute_window_centers_single_mode((512, 512), (4, 4), (16, 16), 50)
    assert r.n_win_x == 255 and r.win_spacing_x == 2 and r.padding == (6, 6, 6, 6)
    counts = {m: compute_window_centers_single_mode(IMAGE_SHAPE, (m, m), (SUM, SUM), 0).n_win_x for m in (4, 16)}
    assert counts[4] == pytest.approx(counts[16] * 4, rel=0.02)
    r = compute_window_centers_single_mode(IMAGE_SHAPE, (4, 6), (SUM, SUM), 0)
    assert r.win_spacing_y == 4 and r.win_spacing_x == 6


@pytest.mark.parametrize("m", A_SIZES)
def test_every_sum_window_lies_inside_the_padded_image(m):
    """The correlator skips a window whose sum window leaves the image: with
    the correlator's origin rule floor(c - (S-1)/2 + 0.5) every window of the
    grid starts at >= 0 and ends inside the padded image, and the first
    window starts at 0 (the grid is as dense as the border allows)."""
    r = compute_window_centers_single_mode(IMAGE_SHAPE, (m, m), (SUM, SUM), 0)
    t, b, l, rr = r.padding
    hp, wp = IMAGE_SHAPE[0] + t + b, IMAGE_SHAPE[1] + l + rr
    assert r.win_ctrs_y.min() - (SUM - 1) / 2 >= -0.5 and r.win_ctrs_x.min() - (SUM - 1) / 2 >= -0.5
    assert r.win_ctrs_y.max() + (SUM + 1) / 2 <= hp + 0.5 and r.win_ctrs_x.max() + (SUM + 1) / 2 <= wp + 0.5
    ox = np.array([O.window_origin(c, SUM) for c in r.win_ctrs_x])
    oy = np.array([O.window_origin(c, SUM) for c in r.win_ctrs_y])
    assert ox.min() == 0 and ox.max() + SUM <= wp and oy.min() >= 0 and oy.max() + SUM == hp
    # the A block of each window covers the original pixels it is centred on
    assert np.all(ox + (SUM - m + 1) // 2 - l == np.floor(r.win_ctrs_x - l - (m - 1) / 2 + 0.5))


def test_original_coordinates_and_pass_grids():
    wc = window_centers_original((200, 240), (8, 8), 50, "single", (32, 32))
    wp = compute_window_centers_single_mode((200, 240), (8, 8), (32, 32), 50)
    assert wc.padding == wp.padding == (12, 12, 12, 12)
    np.testing.assert_array_equal(wc.win_ctrs_x, wp.win_ctrs_x - 12)
    # single-mode centres in original coordinates are the standard grid's
    std = window_centers_original((200, 240), (8, 8), 50)
    np.testing.assert_array_equal(wc.win_ctrs_x, std.win_ctrs_x)
    np.testing.assert_array_equal(wc.win_ctrs_y, std.win_ctrs_y)
    g = build_pass_grids((200, 240), [(32, 32), (8, 8)], [50, 50], ["standard", "single"], (32, 32))
    assert g[1].spacing == (4, 4) and g[1].pred_map_x is not None
    with pytest.raises(ValueError):
        window_centers_original((200, 240), (8, 8), 50, "double", (32, 32))


# ---------------------------------------------------------------- oracle pins
def _padded_pair(shape, m, S_, seed, shift=(0.0, 0.0), n=3):
    rng = np.random.default_rng(seed)
    A, B = [], []
    for _ in range(n):
        x, y, inten = S._positions(rng, shape, 0.03)
        A.append(S.render(shape, x, y, inten).numpy())
        B.append(S.render(shape, x + shift[1], y + shift[0], inten).numpy())
    pad = compute_padding_for_single_mode((m, m), (S_, S_))
    return np.stack(A), np.stack(B), pad


def test_single_correlation_is_the_direct_sum_over_the_a_block():
    """AB(d) = sum over the A block of a(p) b(p + d) (inside the plateau no
    circular wrap reaches the sum window), AA(0) / BB(0) = energy of the
    whole sum window, all at the padded window origins."""
    shape, m, S_ = (40, 44), 6, 24
    A, B, pad = _padded_pair(shape, m, S_, 5, shift=(1.0, -2.0))
    wc = compute_window_centers_single_mode(shape, (m, m), (S_, S_), 50)
    Ap, Bp = EO.pad_single(A, pad), EO.pad_single(B, pad)
    wa = window_weight((m, m), "singlepix", (S_, S_))
    wb = window_weight((S_, S_), "bsingle", (S_, S_))
    ab, aa, bb = O.accumulate_triple(Ap, Bp, np.zeros(wc.n_win_y * wc.n_win_x, np.float32), wc.win_ctrs_x,
                                     wc.win_ctrs_y, (S_, S_), wa, wb, wb, wb)
    s0 = (S_ - m + 1) // 2
    c = S_ // 2
    for j in (0, wc.n_win_y // 2, wc.n_win_y - 1):
        for i in (0, wc.n_win_x - 1):
            r0, c0 = O.window_origin(wc.win_ctrs_y[j], S_), O.window_origin(wc.win_ctrs_x[i], S_)
            blk = np.s_[:, r0 + s0:r0 + s0 + m, c0 + s0:c0 + s0 + m]
            full = np.s_[:, r0:r0 + S_, c0:c0 + S_]
            assert aa[j, i, c, c] == pytest.approx(float((Ap[full].astype(np.float64) ** 2).sum()), rel=1e-5)
            assert bb[j, i, c, c] == pytest.approx(float((Bp[full].astype(np.float64) ** 2).sum()), rel=1e-5)
            for dy, dx in ((0, 0), (1, -2), (-3, 4), (s0, -s0)):
                blk_b = np.s_[:, r0 + s0 + dy:r0 + s0 + dy + m, c0 + s0 + dx:c0 + s0 + dx + m]
                want = float((Ap[blk].astype(np.float64) * Bp[blk_b]).sum())
                assert ab[j, i, c + dy, c + dx] == pytest.approx(want, rel=1e-4, abs=1e-4), (j, i, dy, dx)


def test_single_ab_envelope_is_the_trapezoid():
    S_, m = 32, 8
    for fit in (32, 24):
        wa = window_weight((m, m), "singlepix", (S_, S_))
        wb = window_weight((S_, S_), "bsingle", (S_, S_))
        env = floor_psd.pair_envelope(wa, wb, (fit, fit))
        a0 = (S_ - m + 1) // 2
        d = np.arange(fit) - fit // 2
        # overlap of the A block [a0, a0 + m) with the sum window moved by d
        ov = np.maximum(0, np.minimum(a0 + m, S_ + d) - np.maximum(a0, d)) / m
        np.testing.assert_allclose(env, np.outer(ov, ov), atol=1e-12)
        assert np.all(env > 0) and env[fit // 2, fit // 2] == 1.0
        auto = floor_psd.pair_envelope(wb, wb, (fit, fit))
        tri = (S_ - np.abs(d)) / S_
        np.testing.assert_allclose(auto, np.outer(tri, tri), atol=1e-12)


def test_identical_frames_give_unit_peak_only_with_the_area_scale():
    """A == B: AB(0) = energy of the A block, AA(0) = BB(0) = energy of the
    sum window, so sqrt(AA BB) over-normalises AB by about sum area / A
    area; the scale brings the normalised AB peak back to ~1."""
    shape, m, S_ = (96, 96), 8, 32
    A, _, pad = _padded_pair(shape, m, S_, 11, n=16)
    wc = compute_window_centers_single_mode(shape, (m, m), (S_, S_), 0)
    Ap = EO.pad_single(A, pad)
    wa = window_weight((m, m), "singlepix", (S_, S_))
    wb = window_weight((S_, S_), "bsingle", (S_, S_))
    res = {}
    for scale in (1.0, S_ * S_ / (m * m)):
        r = EO.ensemble_pass(Ap, Ap, wc.win_ctrs_x, wc.win_ctrs_y, (S_, S_), wb, bg_method="none",
                             weight_a=wa, ab_scale=scale)
        c = S_ // 2
        res[scale] = np.median(r["planes"][2].reshape(-1, S_, S_)[:, c, c])
    assert res[1.0] == pytest.approx(m * m / (S_ * S_), rel=0.25)
    assert res[S_ * S_ / (m * m)] == pytest.approx(1.0, rel=0.25)


@pytest.mark.parametrize("fit", [None, 24])
def test_oracle_single_chain_recovers_a_uniform_shift(fit):
    shape, m, S_ = (96, 96), 8, 32
    A, B, pad = _padded_pair(shape, m, S_, 21, shift=(-0.6, 1.3), n=24)
    wc = compute_window_centers_single_mode(shape, (m, m), (S_, S_), 0)
    wa = window_weight((m, m), "singlepix", (S_, S_))
    wb = window_weight((S_, S_), "bsingle", (S_, S_))
    r = EO.ensemble_pass(EO.pad_single(A, pad), EO.pad_single(B, pad), wc.win_ctrs_x, wc.win_ctrs_y, (S_, S_), wb,
                         weight_a=wa, fit_win=None if fit is None else (fit, fit), ab_scale=S_ * S_ / (m * m),
                         peak_scale=np.sqrt(S_ * S_ / (m * m)))
    ok = r["nan_reason"] == 0
    assert ok.mean() > 0.8
    assert np.median(r["ux"][ok]) == pytest.approx(1.3, abs=0.05)
    assert np.median(r["uy"][ok]) == pytest.approx(-0.6, abs=0.05)
    assert r["ux"].shape == (wc.n_win_y, wc.n_win_x)
    ph = r["peakheight"][ok]
    assert np.all((ph >= 0) & (ph <= 1)) and np.median(ph) > 0.3


def test_pad_frames_refuses_bad_geometry_before_any_launch():
    from pivtools_amd import _native
    L = _native.core()
    p = 16
    assert L.pivamd_pad_frames(p, 0, 1, 8, 8, 2, 2, 9, 12, p, None) == 1     # Hp < H + top
    assert L.pivamd_pad_frames(p, 1, 1, 8, 8, -1, 2, 12, 12, p, None) == 1   # negative border
    assert L.pivamd_pad_frames(None, 0, 1, 8, 8, 2, 2, 12, 12, p, None) == 1
    assert L.pivamd_pad_frames(p, 0, 0, 8, 8, 2, 2, 12, 12, p, None) == 0    # no frames
