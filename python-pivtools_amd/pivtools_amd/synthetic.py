"""Seeded synthetic PIV image pairs for the benchmark workloads.

Particles are Gaussians rendered on a local patch of +-3 sigma with
sigma = d/4 (the e^-2 diameter convention of the reference's test
generator, unit-tests/synthetic_piv.py:22-70), peak intensity uniform in
[0.5, 1].  Frame B moves each particle by the displacement field evaluated
at its frame-A position.  Rendering is a scatter-add in torch so the same
code fills host memory for tests and HBM for the benchmark.

Workloads (NumPy PCG64 seeds = 1000 * workload + pair index):
  single_pass_1024  1024 x 1024 f32, 0.02 ppp, dx = 3.2 + 1.5 sin(2 pi y/256),
                    dy = 0.8, additive N(0, 0.01)
  multipass_2560    2160 x 2560, 0.03 ppp, Lamb-Oseen vortex (rc 200 px,
                    peak 8 px) + uniform 2 px in x; 12-bit counts with
                    Poisson shot noise (gain 3000 counts per unit
                    intensity), clipped to 4095, stored uint16
  ensemble_1024     1024 x 1024 f32, particles as single_pass_1024; mean
                    profile u(y) = 6 ln(1 + y/20) px plus a zero-mean field
                    of 256 random Fourier modes (wavelengths 16-512 px,
                    rms 0.6 px in u, 0.4 px in v, u-v correlation -0.3),
                    redrawn for every pair
"""

from __future__ import annotations

import math

import numpy as np
import torch

DIAMETER = 2.5
SIGMA = DIAMETER / 4.0
GAIN_12BIT = 3000.0


def render(shape, x, y, inten, sigma=SIGMA, device="cpu", dtype=torch.float32, cutoff=3.0):
    """Sum of Gaussians I*exp(-r^2 / (2 sigma^2)) over a +-ceil(3 sigma) patch."""
    h, w = shape
    x = torch.as_tensor(x, dtype=torch.float64, device=device)
    y = torch.as_tensor(y, dtype=torch.float64, device=device)
    inten = torch.as_tensor(inten, dtype=torch.float64, device=device)
    rad = int(math.ceil(cutoff * sigma))
    off = torch.arange(-rad, rad + 1, device=device, dtype=torch.int64)
    ix = torch.floor(x).to(torch.int64)
    iy = torch.floor(y).to(torch.int64)
    px = ix[:, None, None] + off[None, None, :]
    py = iy[:, None, None] + off[None, :, None]
    r2 = (px.to(torch.float64) - x[:, None, None]) ** 2 + (py.to(torch.float64) - y[:, None, None]) ** 2
    val = inten[:, None, None] * torch.exp(-r2 / (2.0 * sigma * sigma))
    ok = (px >= 0) & (px < w) & (py >= 0) & (py < h)
    idx = (py * w + px)[ok]
    img = torch.zeros(h * w, dtype=torch.float64, device=device)
    img.index_add_(0, idx, val[ok])
    return img.view(h, w).to(dtype)


def _positions(rng, shape, ppp, margin=3.0):
    h, w = shape
    n = rng.poisson(ppp * h * w)
    x = rng.uniform(-margin, w + margin, n)
    y = rng.uniform(-margin, h + margin, n)
    inten = rng.uniform(0.5, 1.0, n)
    return x, y, inten


def single_pass_pair(seed=0, shape=(1024, 1024), ppp=0.02, noise=0.01, device="cpu"):
    """The single-pass workload: sinusoidal shear in x, uniform 0.8 px in y."""
    rng = np.random.default_rng(seed)
    x, y, inten = _positions(rng, shape, ppp)
    dx = 3.2 + 1.5 * np.sin(2 * np.pi * y / 256.0)
    dy = np.full_like(y, 0.8)
    a = render(shape, x, y, inten, device=device)
    b = render(shape, x + dx, y + dy, inten, device=device)
    na = torch.as_tensor(rng.normal(0, noise, shape), dtype=torch.float32, device=device)
    nb = torch.as_tensor(rng.normal(0, noise, shape), dtype=torch.float32, device=device)
    return a + na, b + nb


def single_pass_truth(xs, ys):
    """Displacement (dx, dy) of the single-pass workload at points (x, y)."""
    ys = np.asarray(ys, np.float64)
    return 3.2 + 1.5 * np.sin(2 * np.pi * ys / 256.0) + 0 * np.asarray(xs), 0.8 + 0 * ys


def lamb_oseen(x, y, shape, rc=200.0, peak=8.0, uniform=(2.0, 0.0)):
    """Lamb-Oseen vortex centred on the image, max |u_theta| = peak, plus uniform flow."""
    h, w = shape
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    rx, ry = x - cx, y - cy
    r = np.sqrt(rx * rx + ry * ry)
    # u_theta(r) = G/(2 pi r) (1 - exp(-r^2/rc^2)); its maximum is at r = 1.12091 rc
    rm = 1.1209141 * rc
    fmax = (1.0 - math.exp(-(rm / rc) ** 2)) / rm
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.where(r > 1e-9, (1.0 - np.exp(-(r / rc) ** 2)) / np.maximum(r, 1e-9), 0.0)
    ut = peak * f / fmax
    with np.errstate(invalid="ignore", divide="ignore"):
        ux = np.where(r > 1e-9, -ut * ry / np.maximum(r, 1e-9), 0.0)
        uy = np.where(r > 1e-9, ut * rx / np.maximum(r, 1e-9), 0.0)
    return ux + uniform[0], uy + uniform[1]


def multipass_pair(seed=1000, shape=(2160, 2560), ppp=0.03, device="cpu", gain=GAIN_12BIT):
    """One uint16 pair of the multipass workload (torch tensors on `device`)."""
    rng = np.random.default_rng(seed)
    x, y, inten = _positions(rng, shape, ppp)
    dx, dy = lamb_oseen(x, y, shape)
    a = render(shape, x, y, inten, device=device)
    b = render(shape, x + dx, y + dy, inten, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    out = []
    for img in (a, b):
        counts = torch.poisson(torch.clamp(img, min=0) * gain, generator=g)
        out.append(torch.clamp(counts, 0, 4095).to(torch.int16).view(torch.int16))
    # int16 storage of 0..4095 is bit-identical to uint16
    return out[0], out[1]


def fourier_fluctuation(rng, n_modes=256, lam=(16.0, 512.0), rms=(0.6, 0.4), corr=-0.3):
    """A random-Fourier-mode fluctuation field (u', v') as a callable of (x, y).

    Two independent unit-rms fields f1, f2 (equal-amplitude modes, random
    direction and phase, log-uniform wavelength); u' = rms_u f1 and
    v' = rms_v (corr f1 + sqrt(1 - corr^2) f2)."""
    def field():
        k = 2 * np.pi / np.exp(rng.uniform(np.log(lam[0]), np.log(lam[1]), n_modes))
        th = rng.uniform(0, 2 * np.pi, n_modes)
        ph = rng.uniform(0, 2 * np.pi, n_modes)
        return k * np.cos(th), k * np.sin(th), ph
    m1, m2 = field(), field()
    amp = math.sqrt(2.0 / n_modes)

    def ev(m, x, y):
        kx, ky, ph = m
        return amp * np.cos(np.outer(x, kx) + np.outer(y, ky) + ph).sum(axis=1)

    def uv(x, y):
        f1, f2 = ev(m1, x, y), ev(m2, x, y)
        return rms[0] * f1, rms[1] * (corr * f1 + math.sqrt(1 - corr * corr) * f2)
    return uv


def ensemble_pair(seed=2000, shape=(1024, 1024), ppp=0.02, noise=0.01, log_scale=6.0, device="cpu"):
    """One f32 pair of the ensemble workload (log-law mean + random modes)."""
    rng = np.random.default_rng(seed)
    x, y, inten = _positions(rng, shape, ppp)
    up, vp = fourier_fluctuation(rng)(x, y)
    dx = log_scale * np.log1p(np.maximum(y, 0.0) / 20.0) + up
    dy = vp
    a = render(shape, x, y, inten, device=device)
    b = render(shape, x + dx, y + dy, inten, device=device)
    na = torch.as_tensor(rng.normal(0, noise, shape), dtype=torch.float32, device=device)
    nb = torch.as_tensor(rng.normal(0, noise, shape), dtype=torch.float32, device=device)
    return a + na, b + nb
