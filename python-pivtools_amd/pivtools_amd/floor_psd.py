"""Analytic coloured noise floor P(k; fx, fy) for the k-space fit (host side).

Restates pivtools_cli/piv/piv_backend/kspace_floor_psd.py:84-275 (and the
tap weights of interpolation_noise_psd.py:37-127).  The fitter models the
white-noise floor of F_ref as N0 * P(k) with P the expected auto spectrum of
one window buffer g = w (u - m) after the whole chain:

* u is white noise seen through the warp kernel at fractional shift
  (fx, fy): separable covariance C = C_y (x) C_x, C_a = Toeplitz(r_a) with
  r_a the autocorrelation of the kernel taps;
* m = (w . u) / sum(w) is the weighted window mean (removed only when the
  background method contains window_mean: ``hole``);
* E|G(k)|^2 = T1(k) - 2 Re[conj(W) V] / sum(w) + q |W|^2 / sum(w)^2 with
  T1(k) = sum_tau r_y r_x A(tau) e^{-i k.tau}, A the zero-padded weight
  autocorrelation, v = C w, V = DFT(w v), q = w . v, W = DFT(w);
* then the production post-chain: inverse DFT, fftshift, central crop to the
  plane size, divide by the stored envelope, centred DFT, magnitude.

P is computed once per pass on a (fy, fx) grid and interpolated per window;
it is host work (tens of small DFTs), not part of the per-pair path.
"""

from __future__ import annotations

import numpy as np

KEYS_A = -0.75


def keys_taps(f: float):
    """Keys a = -0.75 weights at fractional shift f on offsets -1..2."""
    a, f2 = KEYS_A, f * f
    f3 = f2 * f
    w = np.array([a * f3 - 2 * a * f2 + a * f,
                  (a + 2) * f3 - (a + 3) * f2 + 1.0,
                  -(a + 2) * f3 + (2 * a + 3) * f2 - a * f,
                  -a * f3 + a * f2])
    return np.arange(-1, 3), w


def _lanczos3(t: float) -> float:
    if abs(t) < 1e-12:
        return 1.0
    if abs(t) >= 3.0:
        return 0.0
    x = np.pi * t
    return float(np.sin(x) / x * np.sin(x / 3.0) / (x / 3.0))


def lanczos3_taps(f: float):
    """Normalised Lanczos-3 weights at fractional shift f on offsets -2..3."""
    offs = np.arange(-2, 4)
    raw = np.array([_lanczos3(f - d) for d in offs])
    s = raw.sum()
    if s < 1e-12:
        raw = np.zeros(6)
        raw[2] = 1.0
        return offs, raw
    return offs, raw / s


TAPS = {"cubic": keys_taps, "lanczos": lanczos3_taps}


def tap_autocorr(f: float, kernel: str = "cubic"):
    """(lags, r) with r(tau) = sum_i c_i c_{i+tau} of the kernel taps."""
    if kernel not in TAPS:
        raise ValueError(f"unknown kernel {kernel!r}")
    _, c = TAPS[kernel](f)
    n = len(c)
    r = np.correlate(c, c, mode="full")          # lags -(n-1) .. n-1
    return np.arange(-(n - 1), n), r


def _shift(a: np.ndarray, tau: int, axis: int) -> np.ndarray:
    """out(x) = a(x + tau) with zeros outside."""
    out = np.zeros_like(a)
    n = a.shape[axis]
    if abs(tau) >= n:
        return out
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    if tau >= 0:
        src[axis], dst[axis] = slice(tau, n), slice(0, n - tau)
    else:
        src[axis], dst[axis] = slice(0, n + tau), slice(-tau, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


def expected_auto_spectrum(fx: float, fy: float, w: np.ndarray, kernel: str = "cubic",
                           hole: bool = True) -> np.ndarray:
    """E|DFT(w (u - m))|^2 on the (sh, sw) lattice, natural k order."""
    w = np.asarray(w, dtype=np.float64)
    sh, sw = w.shape
    tx, rx = tap_autocorr(fx, kernel)
    ty, ry = tap_autocorr(fy, kernel)
    vx = sum(r * _shift(w, int(t), 1) for t, r in zip(tx, rx) if r != 0.0)
    v = sum(r * _shift(vx, int(t), 0) for t, r in zip(ty, ry) if r != 0.0)
    ky = 2 * np.pi * np.arange(sh)[:, None] / sh
    kx = 2 * np.pi * np.arange(sw)[None, :] / sw
    t1 = np.zeros((sh, sw))
    for tyv, ryv in zip(ty, ry):
        if ryv == 0.0:
            continue
        wy = _shift(w, int(tyv), 0)
        for txv, rxv in zip(tx, rx):
            if rxv == 0.0:
                continue
            a = float(np.sum(w * _shift(wy, int(txv), 1)))
            if a != 0.0:
                t1 += ryv * rxv * a * np.cos(ky * tyv + kx * txv)
    if not hole:
        return t1
    sw_ = w.sum()
    q = float(np.sum(w * v))
    W = np.fft.fft2(w)
    V = np.fft.fft2(w * v)
    es = t1 - (2.0 / sw_) * np.real(np.conj(W) * V) + (q / sw_ ** 2) * np.abs(W) ** 2
    es[0, 0] = 0.0
    return es


def floor_shape(fx: float, fy: float, env_auto: np.ndarray, w: np.ndarray, kernel: str = "cubic",
                hole: bool = True) -> np.ndarray:
    """|F_ref| floor shape (h, w), centred k order, up to a scale."""
    env_auto = np.asarray(env_auto, dtype=np.float64)
    h, wd = env_auto.shape
    sh, sw = np.asarray(w).shape
    es = expected_auto_spectrum(fx, fy, w, kernel, hole)
    r = np.fft.fftshift(np.fft.ifft2(es).real)
    y0, x0 = (sh - h) // 2, (sw - wd) // 2
    plane = r[y0:y0 + h, x0:x0 + wd] / env_auto
    return np.abs(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(plane))))


def build_p_grid(env_auto: np.ndarray, kernel: str, hole: bool, weight: np.ndarray, n_fracs: int = 21):
    """(fs, P[iy, ix, h, w]) on fs = linspace(0, 0.5, n_fracs), unit mean at (0, 0)."""
    fs = np.linspace(0.0, 0.5, n_fracs)
    h, wd = np.asarray(env_auto).shape
    grid = np.empty((n_fracs, n_fracs, h, wd))
    for iy, fy in enumerate(fs):
        for ix, fx in enumerate(fs):
            grid[iy, ix] = floor_shape(float(fx), float(fy), env_auto, weight, kernel, hole)
    grid /= grid[0, 0].mean()
    return fs, grid


def interp_p(fs: np.ndarray, grid: np.ndarray, fx, fy) -> np.ndarray:
    """Bilinear (fx, fy) interpolation per window -> (n, h*w)."""
    fx = np.asarray(fx, dtype=np.float64).ravel()
    fy = np.asarray(fy, dtype=np.float64).ravel()
    nf = len(fs)
    ix = np.clip(np.searchsorted(fs, fx) - 1, 0, nf - 2)
    iy = np.clip(np.searchsorted(fs, fy) - 1, 0, nf - 2)
    tx = np.clip((fx - fs[ix]) / (fs[ix + 1] - fs[ix]), 0.0, 1.0)[:, None]
    ty = np.clip((fy - fs[iy]) / (fs[iy + 1] - fs[iy]), 0.0, 1.0)[:, None]
    g = grid.reshape(nf, nf, -1)
    return ((1 - ty) * (1 - tx) * g[iy, ix] + (1 - ty) * tx * g[iy, ix + 1]
            + ty * (1 - tx) * g[iy + 1, ix] + ty * tx * g[iy + 1, ix + 1])


def frac_distance(x):
    """Distance to the nearest integer (0 at integers, 0.5 at half pixels)."""
    x = np.asarray(x, dtype=np.float64)
    return np.abs(x - np.round(x))


def pair_envelope(wa: np.ndarray, wb: np.ndarray, corr: tuple) -> np.ndarray:
    """Linear pair-count envelope (wa * wb correlation on the full linear
    support), zero lag at corr // 2, unit at zero lag
    (single_pass_accumulator.py:52-102)."""
    wa = np.asarray(wa, dtype=np.float64)
    wb = np.asarray(wb, dtype=np.float64)
    h, w = wa.shape
    ch, cw = corr
    full = np.fft.fftshift(np.fft.ifft2(np.fft.fft2(wa, s=(2 * h, 2 * w))
                                        * np.conj(np.fft.fft2(wb, s=(2 * h, 2 * w)))).real)
    env = full[h - ch // 2:h - ch // 2 + ch, w - cw // 2:w - cw // 2 + cw]
    return env / env[ch // 2, cw // 2]
