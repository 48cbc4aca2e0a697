"""CPU checks of the multipass oracle glue (no GPU)."""

import numpy as np
import pytest

import multipass as M
from pivtools_amd.window import build_pass_grids, compute_window_centers


def test_window_counts_of_the_workloads():
    """Window counts from the grid formula (pivtools_core/window_utils.py:217-255)."""
    g = build_pass_grids((2160, 2560), [(128, 128), (64, 64), (32, 32), (16, 16)], [50] * 4)
    assert [np.prod(x.n_win) for x in g] == [1248, 5214, 21306, 85811]
    assert compute_window_centers((1024, 1024), (32, 32), 50).n_win_x == 63
    assert compute_window_centers((1024, 1024), (32, 32), 75).n_win_x == 125
    assert g[1].ksize == (3, 3) and abs(g[1].sd - 0.65) < 1e-12


def _median_brute(ux, uy, eps, thr, size):
    ny, nx = ux.shape
    out = np.zeros((ny, nx), bool)
    h = size // 2
    for r in range(ny):
        for c in range(nx):
            nbx, nby = [], []
            for dr in range(-h, h + 1):
                for dc in range(-h, h + 1):
                    if dr == 0 and dc == 0:
                        continue
                    rr, cc = r + dr, c + dc
                    if 0 <= rr < ny and 0 <= cc < nx:
                        nbx.append(ux[rr, cc])
                        nby.append(uy[rr, cc])
            mu, mv = np.median(nbx), np.median(nby)
            r0 = np.hypot(ux[r, c] - mu, uy[r, c] - mv)
            rm = np.median(np.hypot(np.array(nbx) - mu, np.array(nby) - mv))
            pres = sum(1 for dr in (-1, 0, 1) for dc in (-1, 0, 1) if 0 <= r + dr < ny and 0 <= c + dc < nx)
            out[r, c] = (r0 / (rm + eps) > thr) or pres < 6
    return out


def test_median_test_vs_bruteforce():
    rng = np.random.default_rng(0)
    ux = rng.normal(3, 0.2, (17, 23)).astype(np.float32)
    uy = rng.normal(-1, 0.2, (17, 23)).astype(np.float32)
    ux[5, 7] += 6
    uy[10, 3] -= 5
    got = M.median_outliers(ux, uy)
    ref = _median_brute(ux.astype(np.float64), uy.astype(np.float64), 0.1, 2.0, 5)
    assert np.array_equal(got, ref)
    assert got[5, 7] and got[10, 3]
    assert got[0, 0] and got[-1, -1]          # corners: only 4 of 9 guard cells


def test_remap_closed_forms():
    ny, nx = 9, 11
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    lin = (0.5 * xx - 0.25 * yy + 2).astype(np.float32)
    field = np.stack([lin[None], (2 * lin)[None]])
    mx = np.array([[1.5, 3.25, 7.0]], np.float32)
    my = np.array([[2.0, 4.75, 6.5]], np.float32)
    out = M.remap_cubic(field, mx, my)

    def keys(t, a=-0.75):
        t = abs(t)
        if t <= 1:
            return (a + 2) * t ** 3 - (a + 3) * t ** 2 + 1
        if t < 2:
            return a * t ** 3 - 5 * a * t ** 2 + 8 * a * t - 4 * a
        return 0.0

    for k in range(3):
        x, y = float(mx[0, k]), float(my[0, k])
        ix, iy = int(np.floor(x)), int(np.floor(y))
        v = 0.0
        for r in range(iy - 1, iy + 3):
            for c in range(ix - 1, ix + 3):
                v += keys(y - r) * keys(x - c) * lin[min(max(r, 0), ny - 1), min(max(c, 0), nx - 1)]
        assert abs(out[0, 0, 0, k] - v) < 2e-5
    # symmetric stencils (half-pixel positions) reproduce a linear field
    mh = np.array([[1.5, 3.0]], np.float32)
    out_h = M.remap_cubic(field, mh, np.array([[2.5, 4.5]], np.float32))
    np.testing.assert_allclose(out_h[0, 0], 0.5 * mh - 0.25 * np.array([[2.5, 4.5]]) + 2, atol=1e-5)
    const = np.full((2, 1, ny, nx), 3.5, np.float32)
    np.testing.assert_allclose(M.remap_cubic(const, mx, my), 3.5, atol=1e-6)
    # integer positions sample the field exactly
    mi = np.array([[0.0, 4.0, 10.0]], np.float32)
    mj = np.array([[0.0, 3.0, 8.0]], np.float32)
    np.testing.assert_array_equal(M.remap_cubic(field, mi, mj)[0, 0, 0], lin[[0, 3, 8], [0, 4, 10]])


def test_gate_reason_invariant():
    rng = np.random.default_rng(2)
    pkx = rng.normal(0, 5, (2, 6, 7)).astype(np.float32)
    pky = rng.normal(0, 5, (2, 6, 7)).astype(np.float32)
    pkx[0, 1, 1] = np.nan
    pky[0, 1, 1] = np.nan
    pkh = np.abs(rng.normal(0, 1, (2, 6, 7))).astype(np.float32)
    u, nm, reason, pk = M.gate(pkx, pky, pkh, None, None, 8.0, 8.0)
    assert reason[0, 1, 1] == 1
    u, nm, reason, pk = M.median_stage(u, pk, nm, reason)
    assert np.array_equal(reason != 0, nm)


def test_infill_nearest_tie_rule_matches_scipy():
    """The device rule (nearest, ties to smallest column then row) equals
    scipy's EDT donor choice on random masks."""
    rng = np.random.default_rng(7)
    for t in range(20):
        ny, nx = rng.integers(5, 25, 2)
        sp = [(1.0, 1.0), (16.0, 16.0), (8.0, 16.0)][t % 3]
        u = rng.normal(size=(2, 1, ny, nx)).astype(np.float32)
        holes = rng.uniform(size=(ny, nx)) < 0.3
        holes[0, 0] = False
        u[:, 0][:, holes] = np.nan
        ref = M.infill_nearest(u, sp)
        ys, xs = np.nonzero(~holes)
        for r, c in zip(*np.nonzero(holes)):
            d = ((ys - r) * sp[0]) ** 2 + ((xs - c) * sp[1]) ** 2
            cand = np.nonzero(d == d.min())[0]
            k = cand[np.lexsort((ys[cand], xs[cand]))[0]]
            assert ref[0, 0, r, c] == u[0, 0, ys[k], xs[k]]
