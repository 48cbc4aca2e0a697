"""Synthetic correlation triplets for the k-space fit, built two ways.

* ``exact_triplet``: planes whose centred spectra are, by construction,
  F_AA = F_BB = F_ref and F_AB = F_ref * T(k) (flat floor) or
  g * decay * (F_ref - N0 * P) * phase (coloured floor), with
  T(k) = g exp(-2 pi^2 k^T S k - b4x kx^4 - b4y ky^4) exp(-2 pi i k.mu),
  the characteristic function of a Gaussian displacement PDF.  The fit must
  return (mu, S, g, N0) to rounding.  Nyquist rows and columns of F_ref are
  zero so the planes are exactly real.
* ``gaussian_triplet``: real-space sampled Gaussian planes as the reference
  tests build them (auto width 2 sigma_p^2, cross width 2 sigma_p^2 + S,
  peak at the centre + mu), recovered to a physical tolerance.
"""

import numpy as np


def kgrid(h, w):
    return np.meshgrid(np.fft.fftfreq(h), np.fft.fftfreq(w), indexing="ij")


def ref_spectrum(h, w, width=0.15):
    ky, kx = kgrid(h, w)
    f = np.exp(-(kx ** 2 + ky ** 2) / (2 * width ** 2))
    f[ky == -0.5] = 0.0
    f[kx == -0.5] = 0.0
    return f


def transfer(h, w, mu, sig, g=1.0, b4=(0.0, 0.0)):
    ky, kx = kgrid(h, w)
    sxx, syy, sxy = sig
    dec = np.exp(-2 * np.pi ** 2 * (sxx * kx ** 2 + 2 * sxy * kx * ky + syy * ky ** 2)
                 - b4[0] * kx ** 4 - b4[1] * ky ** 4)
    return g * dec * np.exp(-2j * np.pi * (kx * mu[0] + ky * mu[1]))


def plane_from_spectrum(spec):
    """Centred real plane R with fftshift(fft2(ifftshift(R))) == fftshift(spec)."""
    r = np.fft.fftshift(np.fft.ifft2(spec))
    assert np.abs(r.imag).max() <= 1e-12 * max(np.abs(r.real).max(), 1e-300)
    return r.real


def floor_shape(h, w):
    """A smooth, even, positive floor shape P in natural order."""
    ky, kx = kgrid(h, w)
    return 1.0 + 0.6 * np.exp(-(kx ** 2 + ky ** 2) / 0.25)


def exact_triplet(h, w, mu, sig, g=1.0, b4=(0.0, 0.0), n0=0.0, amp=1.0):
    """(R_AA, R_BB, R_AB, P_centred or None)."""
    fref = amp * ref_spectrum(h, w)
    t = transfer(h, w, mu, sig, g, b4)
    if n0:
        p = floor_shape(h, w)
        fab = t * (fref - n0 * p)
        fab[fref == 0.0] = 0.0
        pc = np.fft.fftshift(p)
    else:
        fab = t * fref
        pc = None
    raa = plane_from_spectrum(fref)
    return raa, raa.copy(), plane_from_spectrum(fab), pc


def gaussian_plane(shape, cov, centre, amp=1.0):
    h, w = shape
    y, x = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64), indexing="ij")
    dx, dy = x - centre[1], y - centre[0]
    det = cov[0][0] * cov[1][1] - cov[0][1] ** 2
    q = (cov[1][1] * dx * dx - 2 * cov[0][1] * dx * dy + cov[0][0] * dy * dy) / det
    return amp * np.exp(-0.5 * q)


def gaussian_triplet(shape, sp=(2.0, 2.0), stress=(0.0, 0.0, 0.0), mu=(0.0, 0.0), noise=0.0, seed=0):
    h, w = shape
    auto = [[2 * sp[0] ** 2, 0.0], [0.0, 2 * sp[1] ** 2]]
    cross = [[2 * sp[0] ** 2 + stress[0], stress[2]], [stress[2], 2 * sp[1] ** 2 + stress[1]]]
    raa = gaussian_plane(shape, auto, (h / 2.0, w / 2.0))
    rab = gaussian_plane(shape, cross, (h / 2.0 + mu[1], w / 2.0 + mu[0]))
    rbb = raa.copy()
    if noise:
        rng = np.random.default_rng(seed)
        raa = raa + rng.normal(0, noise, shape)
        rbb = rbb + rng.normal(0, noise, shape)
        rab = rab + rng.normal(0, noise, shape)
    return raa, rbb, rab


def noisy_batch(corr=(24, 24), n=40, seed=7, noise=0.002):
    """A batch of genuinely different windows with every seventh masked, as
    the reference's parity gate draws it (test_kspace_c_fit.py:56-85)."""
    rng = np.random.default_rng(seed)
    aa, bb, ab = [], [], []
    for i in range(n):
        sp = (1.8 + 0.4 * rng.random(), 1.8 + 0.4 * rng.random())
        st = (0.3 * rng.random(), 0.3 * rng.random(), 0.0)
        mu = (rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5))
        a, b, c = gaussian_triplet(corr, sp, st, mu, noise=noise, seed=100 + i)
        aa.append(a)
        bb.append(b)
        ab.append(c)
    mask = np.zeros(n, dtype=bool)
    mask[::7] = True
    return np.stack(aa), np.stack(bb), np.stack(ab), mask


def p_win_centred(corr, n):
    """Positive smooth P in the centred layout, one per window (the
    reference gate's probe, test_kspace_c_fit.py:88-99)."""
    ky, kx = np.meshgrid(np.linspace(-1, 1, corr[0]), np.linspace(-1, 1, corr[1]), indexing="ij")
    return np.tile((1.0 + 0.6 * np.exp(-(kx ** 2 + ky ** 2) / 0.25)).ravel(), n).reshape(n, *corr)
