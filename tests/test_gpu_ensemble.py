"""Ensemble accumulation on the device (bulkxcorr2d_accumulate(_triple)
drop-ins) against the oracle and the reference's own identities
(unit-tests/test_ensemble_options.py:115-216)."""

import numpy as np
import pytest

import oracle as O
from pivtools_amd import _native
from pivtools_amd import synthetic as S
from pivtools_amd.window import compute_window_centers, window_weight

pytestmark = pytest.mark.gpu


def _stack(n_pairs, shape, seed=0, shift=(1.3, -0.7), pedestal=0.0):
    A, B = [], []
    for k in range(n_pairs):
        rng = np.random.default_rng(seed + k)
        x, y, inten = S._positions(rng, shape, 0.04)
        A.append(S.render(shape, x, y, inten).numpy() + pedestal)
        B.append(S.render(shape, x + shift[0], y + shift[1], inten).numpy() + pedestal)
    return np.ascontiguousarray(np.stack(A), np.float32), np.ascontiguousarray(np.stack(B), np.float32)


def _triple(A, B, mask, wc, win, wab_a, wab_b, waa, wbb, fit=None, mean=0, norm=0, init=0.0):
    L = _native.bulkxcorr2d_lib()
    N, H, W = A.shape
    ny, nx = wc.n_win_y, wc.n_win_x
    oh, ow = fit if fit else win
    outs = [np.full(ny * nx * oh * ow, init, np.float32) for _ in range(3)]
    fw = None if fit is None else np.array(fit, np.int32)
    err = L.bulkxcorr2d_accumulate_triple(A, B, mask, np.array([H, W], np.int32), N, wc.win_ctrs_x, wc.win_ctrs_y,
                                          np.array([ny, nx], np.int32), wab_a, wab_b, waa, wbb,
                                          np.array(win, np.int32), None if fw is None else fw.ctypes.data,
                                          mean, norm, *outs)
    assert err == 0
    return [o.reshape(ny, nx, oh, ow) for o in outs]


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("n", [16, 32, 64])
@pytest.mark.parametrize("mean,norm", [(0, 0), (1, 0), (1, 1)])
def test_triple_matches_oracle(n, mean, norm):
    shape = (4 * n, 5 * n)
    A, B = _stack(5, shape, seed=n, pedestal=0.3)
    wc = compute_window_centers(shape, (n, n), 75)
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    mask[1, 1] = 1
    wsq = window_weight((n, n), "square")
    wg = window_weight((n, n), "gaussian")
    fit = (n // 2, n // 2)
    got = _triple(A, B, mask, wc, (n, n), wg, wsq, wsq, wsq, fit, mean, norm, init=0.5)
    ref = O.accumulate_triple(A, B, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), wg, wsq, wsq, wsq, fit, mean, norm)
    for g, r in zip(got, ref):
        assert np.all(g[1, 1] == 0.5)          # masked window untouched
        assert _rel(g - 0.5, r) < 1e-5


def test_triple_flags_off_is_fp64_xcorr_and_aa_zero_lag_is_energy():
    """unit-tests/test_ensemble_options.py:115-136."""
    n = 32
    shape = (64, 64)
    A, B = _stack(1, shape, seed=3)
    wc = compute_window_centers(shape, (n, n), 50)
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    w = window_weight((n, n), "square")
    ab, aa, bb = _triple(A, B, mask, wc, (n, n), w, w, w, w)
    r0 = O.window_origin(wc.win_ctrs_y[0], n)
    c0 = O.window_origin(wc.win_ctrs_x[0], n)
    a = A[0, r0:r0 + n, c0:c0 + n].astype(np.float64)
    b = B[0, r0:r0 + n, c0:c0 + n].astype(np.float64)
    ref = np.fft.fftshift(np.real(np.fft.ifft2(np.fft.fft2(b) * np.conj(np.fft.fft2(a))))) / 1.0
    assert _rel(ab[0, 0], ref) < 1e-5
    assert aa[0, 0][n // 2, n // 2] == pytest.approx(np.sum(a * a), rel=1e-5)


def test_per_pair_norm_gain_invariance_and_unit_peaks():
    """unit-tests/test_ensemble_options.py:169-216."""
    n = 32
    shape = (96, 96)
    A, B = _stack(3, shape, seed=9)
    wc = compute_window_centers(shape, (n, n), 50)
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    w = window_weight((n, n), "square")
    ab1, aa1, bb1 = _triple(A, B, mask, wc, (n, n), w, w, w, w, None, 1, 1)
    ab2, aa2, bb2 = _triple(np.ascontiguousarray(A * 7.0), np.ascontiguousarray(B * 0.25), mask, wc, (n, n),
                            w, w, w, w, None, 1, 1)
    assert _rel(ab2, ab1) < 1e-5 and _rel(aa2, aa1) < 1e-5
    np.testing.assert_allclose(aa1[..., n // 2, n // 2], 3.0, rtol=1e-5)
    np.testing.assert_allclose(bb1[..., n // 2, n // 2], 3.0, rtol=1e-5)


def test_window_mean_offset_invariance():
    """unit-tests/test_ensemble_options.py:139-166."""
    n = 32
    shape = (96, 96)
    A, B = _stack(2, shape, seed=12)
    wc = compute_window_centers(shape, (n, n), 50)
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    w = window_weight((n, n), "square")
    r1 = _triple(A, B, mask, wc, (n, n), w, w, w, w, None, 1, 0)
    r2 = _triple(np.ascontiguousarray(A + 5.0), np.ascontiguousarray(B + 5.0), mask, wc, (n, n), w, w, w, w,
                 None, 1, 0)
    for a, b in zip(r1, r2):
        assert _rel(b, a) < 1e-4


def test_accumulate_ab_only_matches_oracle():
    n = 16
    shape = (64, 80)
    A, B = _stack(4, shape, seed=21)
    wc = compute_window_centers(shape, (n, n), 50)
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    wa = window_weight((n, n), "gaussian")
    wb = window_weight((n, n), "square")
    L = _native.bulkxcorr2d_lib()
    N, H, W = A.shape
    out = np.zeros(wc.n_win_y * wc.n_win_x * n * n, np.float32)
    assert L.bulkxcorr2d_accumulate(A, B, mask, np.array([H, W], np.int32), N, wc.win_ctrs_x, wc.win_ctrs_y,
                                    np.array([wc.n_win_y, wc.n_win_x], np.int32), wa, wb,
                                    np.array([n, n], np.int32), None, out) == 0
    ref = np.zeros_like(out)
    O.lib().orc_bulkxcorr2d_accumulate(A, B, mask, np.array([H, W], np.int32), N, wc.win_ctrs_x, wc.win_ctrs_y,
                                       np.array([wc.n_win_y, wc.n_win_x], np.int32), wa, wb,
                                       np.array([n, n], np.int32), None, ref)
    assert _rel(out, ref) < 1e-5


@pytest.mark.parametrize("shared", [True, False])
def test_triple_many_pairs_crosses_flush(shared):
    """130 pairs in one call: the per-lane FP32 partials are flushed into the
    FP64 sums every 64 images (two full flushes and a ragged one)."""
    n = 16
    shape = (64, 80)
    A, B = _stack(130, shape, seed=7, pedestal=0.1)
    wc = compute_window_centers(shape, (n, n), 50)
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    wsq = window_weight((n, n), "square")
    wab = wsq if shared else window_weight((n, n), "gaussian")     # AB weight of frame A
    got = _triple(A, B, mask, wc, (n, n), wab, wsq, wsq, wsq, None, 1, 0)
    ref = O.accumulate_triple(A, B, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), wab, wsq, wsq, wsq, None, 1, 0)
    for g, r in zip(got, ref):
        assert _rel(g, r.reshape(g.shape)) < 1e-5
