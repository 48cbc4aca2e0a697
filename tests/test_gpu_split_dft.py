"""Split-precision MFMA DFT form of the correlation (fp16 hi/lo operands,
FP32 accumulation, N = 16, 32, 48, 64) against the FP64 oracle and against
the FP32 FFT form.

Tolerances: correlation planes within 1e-5 relative L2 of FP64 (the bar the
FP32 form is held to against the reference); integer peaks equal to the FP32
form's except on near-ties; sub-pixel positions within 1e-3 px; the whole
multipass loop within 0.01 px RMS of the FP32 loop.
"""

import numpy as np
import pytest
import torch

import oracle as O
from gpu_util import dropin_bulkxcorr2d
from pivtools_amd import _native
from pivtools_amd import synthetic as S
from pivtools_amd.multipass import MultipassPIV
from pivtools_amd.window import compute_window_centers, window_weight

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


class _mode:
    def __init__(self, m):
        self.m = m

    def __enter__(self):
        self.prev = _native.core().pivamd_set_dft_mode(self.m)

    def __exit__(self, *a):
        _native.core().pivamd_set_dft_mode(self.prev)


def _pair(n, shape, seed=3):
    rng = np.random.default_rng(seed)
    x, y, inten = S._positions(rng, shape, 0.03)
    dx = 0.21 * n * np.sin(2 * np.pi * y / shape[0])
    dy = 0.13 * n * np.cos(2 * np.pi * x / shape[1])
    a = S.render(shape, x, y, inten).numpy() * 3000.0
    b = S.render(shape, x + dx, y + dy, inten).numpy() * 3000.0
    return a, b


@pytest.mark.parametrize("n", [16, 32, 48, 64])
def test_split_planes_vs_fp64_and_fft(n):
    shape = (6 * n + 3, 7 * n + 9)
    a, b = _pair(n, shape, seed=n)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    mask[1, 2] = 1.0
    print(f"split dft planes n={n}", flush=True)
    with _mode(1):
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    fft = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    assert got["err"] == 0 and fft["err"] == 0
    p64, _ = O.planes_fp64(a, b, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w.astype(np.float64),
                           w.astype(np.float64))
    live = mask.ravel() == 0
    g = got["planes"].reshape(-1, n, n)[live].astype(np.float64)
    f = fft["planes"].reshape(-1, n, n)[live].astype(np.float64)
    ref = p64.reshape(-1, n, n)[live]
    err_split = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    err_fft = np.linalg.norm(f - ref) / np.linalg.norm(ref)
    print(f"  rel L2 vs FP64: split {err_split:.2e}  fft {err_fft:.2e}", flush=True)
    assert err_split < 1e-5
    # integer peaks and sub-pixel positions against the FP32 form
    gi = np.argmax(g.reshape(len(g), -1), axis=1)
    fi = np.argmax(f.reshape(len(f), -1), axis=1)
    assert np.mean(gi != fi) <= 0.002
    gx, fx = got["pk_x"].ravel()[live], fft["pk_x"].ravel()[live]
    gy, fy = got["pk_y"].ravel()[live], fft["pk_y"].ravel()[live]
    both = np.isfinite(gx) & np.isfinite(fx)
    assert np.mean(np.isfinite(gx) != np.isfinite(fx)) <= 0.002
    d = np.hypot(gx[both] - fx[both], gy[both] - fy[both])
    assert np.median(d) < 1e-4 and np.mean(d > 1e-3) <= 0.002
    # masked window untouched in both forms
    assert np.all(got["pk_x"][0, 0, 1, 2] == 0.0)


def test_split_zero_and_constant_windows_are_defined():
    """Zero-energy windows (black or constant) give the FFT form's NaN
    sentinel, not garbage, under the exponent scaling."""
    n = 32
    shape = (3 * n, 3 * n)
    a = np.zeros(shape, np.float32)
    b = np.full(shape, 7.0, np.float32)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "square")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    with _mode(1):
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3)
    fft = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3)
    np.testing.assert_array_equal(np.isnan(got["pk_x"]), np.isnan(fft["pk_x"]))
    assert np.all(np.isnan(got["pk_x"]))


def test_split_multipass_loop_vs_fft_loop():
    shape = (256, 320)
    A, B = [], []
    for k in range(3):
        a, b = S.multipass_pair(seed=500 + k, shape=shape)
        A.append(a)
        B.append(b)
    A = torch.stack(A).to(dev)
    B = torch.stack(B).to(dev)
    sched, ov = [(64, 64), (32, 32), (16, 16)], [50, 50, 50]
    r_fft = MultipassPIV(shape, sched, ov).run(A, B)[-1]
    r_split = MultipassPIV(shape, sched, ov, dft="split").run(A, B)[-1]
    torch.cuda.synchronize()
    assert _native.core().pivamd_set_dft_mode(0) == 0          # mode restored after the run
    fm, sm = r_fft.nan_mask.cpu().numpy().astype(bool), r_split.nan_mask.cpu().numpy().astype(bool)
    assert np.mean(fm != sm) <= 0.005
    ok = ~fm & ~sm
    d = np.hypot((r_fft.ux - r_split.ux).cpu().numpy()[ok], (r_fft.uy - r_split.uy).cpu().numpy()[ok])
    assert np.sqrt(np.mean(d ** 2)) < 0.01, np.sqrt(np.mean(d ** 2))


# Register-chained 16-point forms: mode 1 (fp16 hi/lo), 2 (bf16 hi/lo, about
# 16 significant bits, so planes within 1e-4 of FP64), 3 (bf16 hi/mid/lo).
GEMM16_TOL = {1: 1e-5, 2: 1e-4, 3: 1e-5}


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_gemm16_planes_vs_fp64_and_fft(mode):
    n = 16
    shape = (9 * n + 5, 11 * n + 3)
    a, b = _pair(n, shape, seed=40 + mode)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    mask[2, 3] = 1.0
    print(f"gemm16 planes mode={mode}", flush=True)
    with _mode(mode):
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    fft = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    assert got["err"] == 0 and fft["err"] == 0
    p64, _ = O.planes_fp64(a, b, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w.astype(np.float64),
                           w.astype(np.float64))
    live = mask.ravel() == 0
    g = got["planes"].reshape(-1, n, n)[live].astype(np.float64)
    f = fft["planes"].reshape(-1, n, n)[live].astype(np.float64)
    ref = p64.reshape(-1, n, n)[live]
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    err_fft = np.linalg.norm(f - ref) / np.linalg.norm(ref)
    print(f"  rel L2 vs FP64: gemm16 {err:.2e}  fft {err_fft:.2e}", flush=True)
    assert err < GEMM16_TOL[mode]
    # window by window too: no window far off (a wrong lane or register map
    # would spoil some windows, not the average)
    nrm = np.linalg.norm(ref.reshape(len(g), -1), axis=1)
    per = np.linalg.norm((g - ref).reshape(len(g), -1), axis=1) / nrm
    per_fft = np.linalg.norm((f - ref).reshape(len(f), -1), axis=1) / nrm
    k = int(np.argmax(per))
    print(f"  worst window {k}: gemm16 {per[k]:.2e} fft {per_fft[k]:.2e} (fft worst {per_fft.max():.2e}); "
          f"plane norm {nrm[k]:.3e}, median norm {np.median(nrm):.3e}", flush=True)
    # a window with almost no correlation energy has a large relative error in
    # either form; hold each window to the FFT's error there
    assert np.all(per < np.maximum(20 * GEMM16_TOL[mode], 10 * per_fft))
    gi = np.argmax(g.reshape(len(g), -1), axis=1)
    fi = np.argmax(f.reshape(len(f), -1), axis=1)
    assert np.mean(gi != fi) <= 0.002
    gx, fx = got["pk_x"].ravel()[live], fft["pk_x"].ravel()[live]
    assert np.mean(np.isfinite(gx) != np.isfinite(fx)) <= 0.002
    assert np.all(got["pk_x"][0, 0, 2, 3] == 0.0)                # masked window untouched
    # the plane dump equals what the fit saw: integer peaks from the records
    gy = got["pk_y"].ravel()[live]
    both = np.isfinite(gx) & np.isfinite(fx)
    d = np.hypot(gx[both] - fx[both], gy[both] - fft["pk_y"].ravel()[live][both])
    assert np.median(d) < (1e-4 if mode != 2 else 1e-3)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_gemm16_zero_and_constant_windows(mode):
    n = 16
    shape = (4 * n, 4 * n)
    a = np.zeros(shape, np.float32)
    b = np.full(shape, 7.0, np.float32)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "square")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    with _mode(mode):
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3)
    assert np.all(np.isnan(got["pk_x"]))


def test_gemm16_u16_loop_vs_fft_loop():
    """16-bit frames straight into a 16-point first pass (pivamd_xcorr_u16),
    then a warped 16-point pass, for every 16-point form."""
    shape = (192, 256)
    A, B = [], []
    for k in range(2):
        a, b = S.multipass_pair(seed=700 + k, shape=shape, device=dev)
        A.append(a)
        B.append(b)
    A = torch.stack(A)
    B = torch.stack(B)
    assert A.dtype == torch.int16
    sched, ov = [(16, 16), (16, 16)], [50, 50]
    r_fft = MultipassPIV(shape, sched, ov).run(A, B)[-1]
    fm = r_fft.nan_mask.cpu().numpy().astype(bool)
    for dft in ("split", "split-bf16", "split-bf16x3"):
        print(f"gemm16 u16 loop {dft}", flush=True)
        r = MultipassPIV(shape, sched, ov, dft=dft).run(A, B)[-1]
        torch.cuda.synchronize()
        sm = r.nan_mask.cpu().numpy().astype(bool)
        assert np.mean(fm != sm) <= 0.005
        ok = ~fm & ~sm
        d = np.hypot((r_fft.ux - r.ux).cpu().numpy()[ok], (r_fft.uy - r.uy).cpu().numpy()[ok])
        rms = float(np.sqrt(np.mean(d ** 2)))
        print(f"  rms vs fft {rms:.2e}", flush=True)
        assert rms < 0.01


def test_staged_form_at_16_still_selectable():
    """The LDS-staged MFMA form stays reachable at 16 points as a variant."""
    L = _native.core()
    n = 16
    shape = (5 * n, 6 * n)
    a, b = _pair(n, shape, seed=77)
    wc = compute_window_centers(shape, (n, n), 50)
    w = window_weight((n, n), "gaussian")
    mask = np.zeros((wc.n_win_y, wc.n_win_x), np.float32)
    prev = L.pivamd_xcorr_set_variant(16, 16)
    assert prev >= 0
    try:
        got = dropin_bulkxcorr2d(a, b, mask, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w, w, 1, 3, want_planes=True)
    finally:
        L.pivamd_xcorr_set_variant(16, prev)
    p64, _ = O.planes_fp64(a, b, wc.win_ctrs_x, wc.win_ctrs_y, (n, n), w.astype(np.float64),
                           w.astype(np.float64))
    g = got["planes"].reshape(-1, n, n).astype(np.float64)
    assert np.linalg.norm(g - p64.reshape(-1, n, n)) / np.linalg.norm(p64) < 1e-5
