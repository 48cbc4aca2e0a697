"""Pins for the k-space fit oracle (oracle/kspace_oracle.c).

Every check is independent of the oracle's own formulas: numpy's FFT, planes
whose spectra are built from the characteristic function of a known Gaussian
displacement PDF, numpy medians and the reference's output contract
(kspace_lm_fit.c:669-886, test_kspace_c_fit.py:202-270).
"""

import numpy as np
import pytest

import oracle as O
from kspace_synth import (exact_triplet, floor_shape, gaussian_triplet, kgrid, noisy_batch,
                          p_win_centred, ref_spectrum)

SIZES = [8, 12, 16, 24, 32, 48, 64, 96, 128]


@pytest.mark.parametrize("shape", [(n, n) for n in SIZES] + [(24, 32), (16, 48)])
def test_centred_fft_matches_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    r = rng.normal(size=shape)
    ret, f = O.kspace_fft2_centered(r)
    assert ret == 0
    ref = np.fft.fft2(np.fft.ifftshift(r))
    np.testing.assert_allclose(f, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_unsupported_axis():
    ret, _ = O.kspace_fft2_centered(np.zeros((10, 16)))
    assert ret == -1
    z = np.zeros((1, 10, 10))
    ret, _ = O.kspace_fit(z, z, z, np.zeros(1, bool), (10, 10))
    assert ret == -1


@pytest.mark.parametrize("corr", [(32, 32), (24, 24), (16, 32)])
@pytest.mark.parametrize("case", [
    dict(mu=(0.7, -1.3), sig=(0.30, 0.12, 0.05), g=0.8),
    dict(mu=(-2.25, 0.4), sig=(0.05, 0.45, -0.08), g=1.3),
    dict(mu=(3.1, 2.6), sig=(0.0, 0.0, 0.0), g=1.0),
])
def test_exact_transfer_recovered(corr, case):
    h, w = corr
    raa, rbb, rab, _ = exact_triplet(h, w, case["mu"], case["sig"], case["g"])
    ret, o = O.kspace_fit(raa[None], rbb[None], rab[None], np.zeros(1, bool), corr)
    assert ret == 0 and o["status"][0] == 0 and o["conv"][0] == 1
    g = o["gauss"][0]
    np.testing.assert_allclose(g[14:16] - [w / 2 + 1, h / 2 + 1], case["mu"], atol=1e-7)
    np.testing.assert_allclose(g[9:12], np.maximum(case["sig"], [0, 0, -np.inf]), atol=1e-7)
    np.testing.assert_allclose(o["gain"][0], case["g"], rtol=1e-7)
    assert abs(o["N0"][0]) < 1e-7
    assert o["cost_per_pt"][0] < 1e-18
    # n_valid: every bin with F_ref > 1e-12 except DC
    assert o["n_valid"][0] == int(np.sum(ref_spectrum(h, w) > 1e-12)) - 1


@pytest.mark.parametrize("kx4,ky4", [(1, 0), (0, 1), (1, 1)])
def test_exact_quartic_recovered(kx4, ky4):
    b4 = (40.0 * kx4, 25.0 * ky4)
    raa, rbb, rab, _ = exact_triplet(32, 32, (0.4, -0.9), (0.2, 0.1, 0.03), 0.9, b4=b4)
    ret, o = O.kspace_fit(raa[None], rbb[None], rab[None], np.zeros(1, bool), (32, 32), kx4=kx4, ky4=ky4)
    assert ret == 0 and o["status"][0] == 0
    np.testing.assert_allclose(o["gauss"][0, 9:12], (0.2, 0.1, 0.03), atol=1e-6)
    if kx4:
        np.testing.assert_allclose(o["b4x"][0], b4[0], rtol=1e-6)
    else:
        assert np.isnan(o["b4x"][0])
    if ky4:
        np.testing.assert_allclose(o["b4y"][0], b4[1], rtol=1e-6)
    else:
        assert np.isnan(o["b4y"][0])


def test_exact_coloured_floor_recovered():
    h = w = 32
    raa, rbb, rab, pc = exact_triplet(h, w, (1.2, -0.6), (0.25, 0.15, 0.0), 0.7, n0=0.004)
    ret, o = O.kspace_fit(raa[None], rbb[None], rab[None], np.zeros(1, bool), (h, w), p_win=pc[None])
    assert ret == 0 and o["status"][0] == 0
    np.testing.assert_allclose(o["N0"][0], 0.004, rtol=1e-6)
    np.testing.assert_allclose(o["gauss"][0, 9:11], (0.25, 0.15), atol=1e-6)
    np.testing.assert_allclose(o["gain"][0], 0.7, rtol=1e-6)


def test_regressor_and_n0_seed_against_numpy():
    """Dv = P / max(F_ref, 1e-30) in natural order and the N0 seed =
    clip(median(F_ref / P) over |k| >= 0.35, 1e-3, 10), with F_ref from numpy."""
    rng = np.random.default_rng(3)
    h, w = 24, 24
    raa, rbb, rab = gaussian_triplet((h, w), (1.9, 2.1), (0.1, 0.2, 0.0), (0.3, -0.2), noise=0.01, seed=4)
    pc = np.fft.fftshift(floor_shape(h, w)) * rng.uniform(0.5, 1.5, (h, w))
    ret, dv, n0 = O.kspace_debug_prep(raa, rbb, rab, pc)
    assert ret == 0
    fa = np.abs(np.fft.fft2(np.fft.ifftshift(raa)))
    fb = np.abs(np.fft.fft2(np.fft.ifftshift(rbb)))
    fr = np.sqrt(fa * fb)
    pn = np.fft.ifftshift(pc)
    np.testing.assert_allclose(dv, pn / np.maximum(fr, 1e-30), rtol=1e-11)
    ky, kx = kgrid(h, w)
    tail = kx ** 2 + ky ** 2 >= 0.35 ** 2
    expect = np.clip(np.median(fr[tail] / np.maximum(pn[tail], 1e-30)), 1e-3, 10.0)
    assert abs(n0 - expect) <= 1e-12 * expect
    ret, dv, n0 = O.kspace_debug_prep(raa, rbb, rab, None)
    np.testing.assert_allclose(dv, 1.0 / np.maximum(fr, 1e-30), rtol=1e-11)
    assert n0 == 0.0


def test_mu_seed_is_exact_on_axis_aligned_gaussian():
    """The 3-point log-Gaussian estimate is exact on sampled axis-aligned
    Gaussians, so the initial guess carries the true displacement."""
    h = w = 32
    mu = (0.37, -0.81)
    raa, rbb, rab = gaussian_triplet((h, w), (2.0, 1.7), (0.3, 0.1, 0.0), mu)
    ret, o = O.kspace_fit(raa[None], rbb[None], rab[None], np.zeros(1, bool), (h, w))
    assert ret == 0
    ig = o["init"][0]
    np.testing.assert_allclose(ig[14:16] - [w / 2 + 1, h / 2 + 1], mu, atol=1e-9)
    assert ig[0] == raa[h // 2, w // 2] and ig[1] == rbb[h // 2, w // 2] and ig[2] == rab.max()
    assert ig[9] == 1.0 and ig[10] == 0.2 and np.all(np.isnan(ig[6:9]))


def test_gaussian_planes_recover_stress():
    """Real-space Gaussian planes (auto 2 Sp, cross 2 Sp + S): the fit returns
    S and mu to the accuracy that truncation of the planes allows."""
    corr = (32, 32)
    for st, mu in [((0.3, 0.1, 0.05), (0.6, -0.4)), ((0.0, 0.25, 0.0), (-1.1, 0.9))]:
        raa, rbb, rab = gaussian_triplet(corr, (1.6, 1.6), st, mu)
        ret, o = O.kspace_fit(raa[None], rbb[None], rab[None], np.zeros(1, bool), corr)
        assert ret == 0 and o["status"][0] == 0
        np.testing.assert_allclose(o["gauss"][0, 9:12], st, atol=2e-3)
        np.testing.assert_allclose(o["gauss"][0, 14:16] - [17, 17], mu, atol=1e-3)


def test_output_contract_masked_and_gates():
    corr = (24, 24)
    aa, bb, ab, mask = noisy_batch(corr, 14)
    aa[3] = 0.0                                    # amplitude gate -> low SNR
    ret, o = O.kspace_fit(aa, bb, ab, mask, corr, p_win=p_win_centred(corr, 14))
    assert ret == 0
    st, gf, ig = o["status"], o["gauss"], o["init"]
    assert np.all(st[mask] == -1)
    assert np.all(ig[mask] == 0.0)
    assert np.all(np.isnan(gf[mask][:, 6:9]))
    assert np.all(gf[mask][:, 12] == 13.0) and np.all(gf[mask][:, 13] == 13.0)
    assert np.all(np.isnan(o["gain"][mask])) and np.all(o["n_valid"][mask] == 0)
    assert st[3] == 2 and np.isnan(o["gain"][3])
    assert np.sum(st == 0) >= 10
    # too few valid bins: a plane whose spectrum is zero away from DC
    flat = np.full((1, 24, 24), 0.5)
    ret, o2 = O.kspace_fit(flat, flat, flat, np.zeros(1, bool), corr)
    assert o2["status"][0] == 1 and o2["n_valid"][0] == 0


def test_all_masked_batch_mirrors_defaults():
    corr = (24, 24)
    aa, bb, ab, _ = noisy_batch(corr, 6)
    ret, o = O.kspace_fit(aa, bb, ab, np.ones(6, bool), corr)
    assert ret == 0 and np.all(o["status"] == -1)
    np.testing.assert_array_equal(o["init"], o["gauss"])


def test_noisy_batch_statuses_converge():
    corr = (24, 24)
    aa, bb, ab, mask = noisy_batch(corr, 40)
    for p in (None, p_win_centred(corr, 40)):
        ret, o = O.kspace_fit(aa, bb, ab, mask, corr, p_win=p)
        assert ret == 0
        ok = o["status"] == 0
        assert ok.sum() >= 30
        assert np.all(o["gauss"][ok, 9:11] >= 0)


def test_multipass_ensemble_oracle_recovers_a_uniform_shift():
    """The oracle's two-pass ensemble chain on pairs displaced by a uniform
    (dy, dx) = (-1.3, 4.6) px: pass 0 measures it directly, pass 1 warps
    by it (a shared predictor) and measures a residual near 0, and the
    added-back field equals the shift; with even-integer rounding the
    warp moves whole pixels and the residual carries the rest."""
    import ensemble_oracle as EO
    from pivtools_amd import synthetic as S
    from pivtools_amd.window import build_pass_grids

    shape = (128, 128)
    A, B = [], []
    for k in range(6):
        rng = np.random.default_rng(70 + k)
        x, y, inten = S._positions(rng, shape, 0.05)
        A.append(S.render(shape, x, y, inten).numpy())
        B.append(S.render(shape, x + 4.6, y - 1.3, inten).numpy())
    A, B = np.stack(A), np.stack(B)
    grids = build_pass_grids(shape, [(32, 32), (32, 32)], [50, 50])
    for rounding in (False, True):
        out = EO.ensemble_multipass(A, B, grids, rounding=rounding, final_infill="none")
        for p, r in enumerate(out):
            ux, uy = r["ux"][1:-1, 1:-1], r["uy"][1:-1, 1:-1]
            assert abs(np.nanmedian(ux) - 4.6) < 0.05 and abs(np.nanmedian(uy) + 1.3) < 0.05, (rounding, p)
        pred = out[1]["predictor"]
        if rounding:
            assert np.all(pred[1] == 4.0) and np.all(pred[0] == -2.0)
        else:
            assert abs(np.median(pred[1]) - 4.6) < 0.05
