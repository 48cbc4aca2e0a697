"""Pins for the analytic noise-floor shape (pivtools_amd/floor_psd.py).

The closed form for E|DFT(w (u - m))|^2 is checked against the brute-force
quadratic form diag(F G C G^T F^H) with explicit matrices on small lattices
(C = C_y (x) C_x from the kernel taps, G = diag(w)(I - 1 w^T / sum w)),
which shares no algebra with the closed form.
"""

import numpy as np
import pytest

from pivtools_amd import floor_psd as FP


def _toeplitz(n, lags, r):
    c = np.zeros((n, n))
    for t, v in zip(lags, r):
        for i in range(n):
            j = i + t
            if 0 <= j < n:
                c[i, j] = v
    return c


def brute_force(fx, fy, w, kernel, hole):
    sh, sw = w.shape
    lx, rx = FP.tap_autocorr(fx, kernel)
    ly, ry = FP.tap_autocorr(fy, kernel)
    c = np.kron(_toeplitz(sh, ly, ry), _toeplitz(sw, lx, rx))
    wf = w.ravel()
    g = np.diag(wf)
    if hole:
        g = g @ (np.eye(wf.size) - np.outer(np.ones(wf.size), wf) / wf.sum())
    cov = g @ c @ g.T
    yy, xx = np.meshgrid(np.arange(sh), np.arange(sw), indexing="ij")
    ky, kx = np.meshgrid(np.arange(sh), np.arange(sw), indexing="ij")
    phase = 2 * np.pi * (np.outer(ky.ravel(), yy.ravel()) / sh + np.outer(kx.ravel(), xx.ravel()) / sw)
    f = np.exp(-1j * phase)
    return np.real(np.einsum("ka,ab,kb->k", f, cov, np.conj(f))).reshape(sh, sw)


@pytest.mark.parametrize("kernel", ["cubic", "lanczos"])
@pytest.mark.parametrize("hole", [True, False])
@pytest.mark.parametrize("f", [(0.0, 0.0), (0.3, 0.1), (0.5, 0.45)])
def test_expected_spectrum_matches_quadratic_form(kernel, hole, f):
    rng = np.random.default_rng(int(100 * f[0] + 10 * f[1]) + hole)
    w = rng.uniform(0.2, 1.0, (8, 10))
    es = FP.expected_auto_spectrum(f[0], f[1], w, kernel, hole)
    bf = brute_force(f[0], f[1], w, kernel, hole)
    np.testing.assert_allclose(es, bf, rtol=1e-10, atol=1e-10 * np.abs(bf).max())
    if hole:
        assert abs(bf[0, 0]) < 1e-9 * np.abs(bf).max()


def test_taps_sum_to_one_and_interpolate():
    for f in (0.0, 0.25, 0.5, 0.9):
        for k in ("cubic", "lanczos"):
            _, c = FP.TAPS[k](f)
            assert abs(c.sum() - 1.0) < 1e-12
    # half shift: taps symmetric about the interval centre
    _, c = FP.keys_taps(0.5)
    assert c[0] == c[3] and c[1] == c[2]
    _, c = FP.lanczos3_taps(0.5)
    np.testing.assert_allclose(c, c[::-1], rtol=1e-14)


def test_square_unshifted_no_hole_is_flat():
    """f = 0 means identity taps: white noise stays white, so with the square
    weight and no mean removal the floor is flat; after normalisation P = 1."""
    env = FP.pair_envelope(np.ones((32, 32)), np.ones((32, 32)), (32, 32))
    fs, grid = FP.build_p_grid(np.ones((32, 32)), "cubic", False, np.ones((32, 32)), n_fracs=2)
    np.testing.assert_allclose(grid[0, 0], 1.0, rtol=1e-12)
    assert env[16, 16] == 1.0 and env.min() > 0
    # (N - |m|)(N - |n|) / N^2 closed form of the square envelope
    m = np.arange(32) - 16
    np.testing.assert_allclose(env, np.outer(32 - np.abs(m), 32 - np.abs(m)) / 1024.0, rtol=1e-12)


def test_interp_nodes_and_midpoints():
    fs = np.linspace(0, 0.5, 3)
    grid = np.arange(3 * 3 * 4, dtype=float).reshape(3, 3, 2, 2)
    p = FP.interp_p(fs, grid, [0.0, 0.25, 0.5, 0.125], [0.0, 0.5, 0.25, 0.0])
    np.testing.assert_allclose(p[0], grid[0, 0].ravel())
    np.testing.assert_allclose(p[1], grid[2, 1].ravel())
    np.testing.assert_allclose(p[2], grid[1, 2].ravel())
    np.testing.assert_allclose(p[3], 0.5 * (grid[0, 0] + grid[0, 1]).ravel())
    np.testing.assert_allclose(FP.frac_distance([0.2, 0.7, -1.5, 2.0]), [0.2, 0.3, 0.5, 0.0])
