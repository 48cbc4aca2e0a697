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
            nbx, nby = np.array(nbx), np.array(nby)
            keep = np.isfinite(nbx) & np.isfinite(nby)
            nbx, nby = nbx[keep], nby[keep]
            mu, mv = np.median(nbx), np.median(nby)
            r0 = np.hypot(ux[r, c] - mu, uy[r, c] - mv)
            rm = np.median(np.hypot(nbx - mu, nby - mv))
            pres = sum(1 for dr in (-1, 0, 1) for dc in (-1, 0, 1) if 0 <= r + dr < ny and 0 <= c + dc < nx
                       and np.isfinite(ux[r + dr, c + dc]))
            out[r, c] = (r0 / (rm + eps) > thr) or pres < 6
    return out


@pytest.mark.parametrize("size", [3, 5, 7, 9])
def test_median_test_vs_bruteforce_any_size(size):
    """outlier_detection.py:87-243 accepts any odd size >= 3."""
    rng = np.random.default_rng(10 + size)
    ux = rng.normal(3, 0.2, (15, 19)).astype(np.float32)
    uy = rng.normal(-1, 0.2, (15, 19)).astype(np.float32)
    ux[rng.uniform(size=ux.shape) < 0.08] = np.nan
    uy[np.isnan(ux)] = np.nan
    ux[7, 9] += 5
    got = M.median_outliers(ux, uy, size=size)
    ref = _median_brute(ux.astype(np.float64), uy.astype(np.float64), 0.1, 2.0, size)
    fin = np.isfinite(ux)
    assert np.array_equal(got[fin], ref[fin])
    assert got[7, 9]


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


def _keys64(t, a=-0.75):
    t = abs(t)
    if t <= 1:
        return (a + 2) * t ** 3 - (a + 3) * t ** 2 + 1
    if t < 2:
        return a * t ** 3 - 5 * a * t ** 2 + 8 * a * t - 4 * a
    return 0.0


def _remap_fp64_padded(f, x, y, pad=6):
    """FP64 bicubic value at lattice position (x, y) computed on an
    explicitly edge-padded copy of f (np.pad mode='edge'): an independent
    statement of the replicate border, without index clamping."""
    fp = np.pad(f.astype(np.float64), pad, mode="edge")
    ix, iy = int(np.floor(x)), int(np.floor(y))
    v = 0.0
    for r in range(iy - 1, iy + 3):
        for c in range(ix - 1, ix + 3):
            v += _keys64(y - r) * _keys64(x - c) * fp[r + pad, c + pad]
    return v


def test_remap_rounds_to_the_nearest_1_32_lattice_point_ties_to_even():
    """cv2 converts the float maps to X = cvRound(32 x) (round half to even)
    and interpolates at X / 32.  Positions a little off the lattice give the
    lattice point's value bit for bit (a floor rule fails for positions just
    below a lattice point), and positions on the tie k/32 + 1/64 take the
    even one of k and k + 1 (a round-half-up rule fails for even k)."""
    rng = np.random.default_rng(21)
    ny, nx = 10, 12
    f = rng.normal(size=(ny, nx)).astype(np.float32)
    field = f[None, None]
    ks = np.arange(64, 200)                     # lattice indices 2.0 .. 6.2 px
    on = (ks / 32.0).astype(np.float32)
    y0 = np.full_like(on, np.float32(4.25))
    ref = M.remap_cubic(field, on[None], y0[None])[0, 0, 0]
    for d in (-1 / 128, -1 / 96, 1 / 96, 1 / 128):
        off = (ks / 32.0 + d).astype(np.float32)
        np.testing.assert_array_equal(M.remap_cubic(field, off[None], y0[None])[0, 0, 0], ref)
    tie = (ks / 32.0 + 1 / 64).astype(np.float32)
    even = ks + (ks % 2)                        # round half to even
    want = M.remap_cubic(field, (even / 32.0).astype(np.float32)[None], y0[None])[0, 0, 0]
    got = M.remap_cubic(field, tie[None], y0[None])[0, 0, 0]
    np.testing.assert_array_equal(got, want)
    # the same along y
    gy = M.remap_cubic(field, np.full_like(tie, np.float32(5.5))[None], tie[None])[0, 0, 0]
    wy = M.remap_cubic(field, np.full_like(tie, np.float32(5.5))[None], (even / 32.0).astype(np.float32)[None])
    np.testing.assert_array_equal(gy, wy[0, 0, 0])
    # and every one of them against FP64 Keys weights at the rounded position
    for k, v in zip(even[:20], got[:20]):
        assert abs(v - _remap_fp64_padded(f, k / 32.0, 4.25)) < 2e-5


def test_remap_border_is_replicate():
    """Stencils that reach each border (and positions outside the field)
    read the edge values: against an FP64 evaluation on an edge-padded copy,
    which a zero, mirror or off-by-one clamp fails."""
    rng = np.random.default_rng(22)
    ny, nx = 9, 11
    f = rng.normal(size=(ny, nx)).astype(np.float32)
    pts = [(-0.75, 4.0), (0.3, 4.5), (nx - 1.6, 3.25), (nx - 0.2, 2.0), (nx + 0.9, 6.5),
           (5.5, -1.4), (4.25, 0.6), (6.0, ny - 1.3), (3.75, ny + 0.4), (-1.2, -1.7), (nx + 0.3, ny + 0.8),
           (0.0, 0.0), (nx - 1.0, ny - 1.0)]
    mx = np.array([[p[0] for p in pts]], np.float32)
    my = np.array([[p[1] for p in pts]], np.float32)
    out = M.remap_cubic(f[None, None], mx, my)[0, 0, 0]
    for (x, y), v in zip(pts, out):
        xl = np.rint(np.float32(x) * 32) / 32.0
        yl = np.rint(np.float32(y) * 32) / 32.0
        ref = _remap_fp64_padded(f, xl, yl)
        assert abs(v - ref) < 2e-5 * max(1.0, abs(ref)), (x, y, v, ref)
    # far outside: the corner value exactly weighted by a partition of unity
    far = M.remap_cubic(f[None, None], np.array([[-5.0]], np.float32), np.array([[-5.0]], np.float32))
    assert abs(far[0, 0, 0, 0] - f[0, 0]) < 1e-6
    # at integer positions on the border the value is the sample itself
    edge = M.remap_cubic(f[None, None], np.array([[0.0, nx - 1.0]], np.float32),
                         np.array([[ny - 1.0, 0.0]], np.float32))[0, 0, 0]
    np.testing.assert_array_equal(edge, [f[ny - 1, 0], f[0, nx - 1]])


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


def _smooth_field(ny=9, nx=11, seed=31):
    rng = np.random.default_rng(seed)
    pkx = (2.0 + 0.02 * rng.normal(size=(1, ny, nx))).astype(np.float32)
    pky = (-1.0 + 0.02 * rng.normal(size=(1, ny, nx))).astype(np.float32)
    pkh = np.ones((1, ny, nx), np.float32)
    return pkx, pky, pkh


MEDIAN = [{"type": "median_2d", "epsilon": 0.1, "threshold": 2.0, "size": 5}]


def test_nan_reason_known_answers_and_precedence():
    """Every single-peak code of nan_reason_codes.py:18-25 where the stage
    that sets it is known by construction (cpu_instantaneous.py:576-757):
    masked -1 overrides everything (also a masked large displacement), the
    fit's NaN sentinel 1, the large-displacement gate 6 (win/4 after pass 0),
    a primary-peak outlier 10, a NaN peak height 13 on an otherwise valid
    vector, and a fit failure the median test also flags keeps 1 (the
    outlier code is written only where no earlier stage claimed the vector)."""
    pkx, pky, pkh = _smooth_field()
    mask = np.zeros((9, 11), bool)
    mask[1, 1] = True
    pkx[0, 1, 1] = 100.0                 # masked and large: masked wins
    pkx[0, 2, 2] = np.nan                # fit failure
    pky[0, 3, 3] = 9.5                   # |dy| > 8: large displacement
    pkx[0, 4, 5] = 6.0                   # isolated wild vector inside the gate
    pkh[0, 6, 7] = np.nan                # valid vector, NaN peak height
    u, nm, reason, ph = M.gate(pkx, pky, pkh, None, mask, 8.0, 8.0)
    assert reason[0, 2, 2] == 1 and reason[0, 3, 3] == 6 and reason[0, 1, 1] == 6
    u, nm, reason, ph = M.outlier_stage(u, ph, nm, reason, mask, MEDIAN)
    want = np.zeros((1, 9, 11), np.int8)
    want[0, 1, 1] = -1
    want[0, 2, 2] = 1
    want[0, 3, 3] = 6
    want[0, 4, 5] = 10
    want[0, 6, 7] = 13
    # the four corners see 4 of their 9 guard cells, fewer than the 6 the
    # median test requires (outlier_detection.py:20-25, 221-240): outliers too
    want[0, [0, 0, -1, -1], [0, -1, 0, -1]] = 10
    np.testing.assert_array_equal(reason, want)
    np.testing.assert_array_equal(nm, want != 0)
    assert np.isnan(u[:, nm & ~mask[None]]).all() and (u[:, 0, 1, 1] == 0).all()
    assert np.isnan(ph[nm]).all() and not np.isnan(ph[~nm]).any()


def test_nan_reason_unclassified_and_masked_block():
    """Code 9: a vector in nan_mask that no stage claimed -- here a NaN
    predictor with the outlier test off (outlier_methods = []).  And the
    reference's masked-block test (unit-tests/test_nan_reason.py:97-106): a
    block of masked windows is exactly -1 and -1 appears nowhere else."""
    pkx, pky, pkh = _smooth_field()
    pred = np.zeros((2, 1, 9, 11), np.float32)
    pred[1, 0, 5, 3] = np.nan
    mask = np.zeros((9, 11), bool)
    mask[1:3, 1:4] = True
    u, nm, reason, ph = M.gate(pkx, pky, pkh, pred, mask, 8.0, 8.0)
    u, nm, reason, ph = M.outlier_stage(u, ph, nm, reason, mask, [])
    assert reason[0, 5, 3] == 9
    np.testing.assert_array_equal(reason[0] == -1, mask)
    np.testing.assert_array_equal(reason != 0, nm)
    assert set(np.unique(reason)) == {-1, 0, 9}


def test_nan_reason_secondary_peak_codes():
    """Several peaks with secondary substitution (cpu_instantaneous.py:664-722):
    a primary outlier whose second peak fits the field is rescued (0); one
    whose second peak is an outlier too gets 12; one whose second peak is
    the fit's NaN sentinel gets 13; a primary fit failure with a good second
    peak is rescued; a masked window stays -1."""
    pkx1, pky1, pkh1 = _smooth_field()
    P = 2
    pkx = np.stack([pkx1, pkx1 + 0.01], axis=1)
    pky = np.stack([pky1, pky1 - 0.01], axis=1)
    pkh = np.stack([pkh1, 0.5 * pkh1], axis=1)
    mask = np.zeros((9, 11), bool)
    mask[7, 9] = True
    pkx[0, 0, 2, 3] = 6.5                        # primary outlier, second peak good
    pkx[0, 0, 4, 5], pkx[0, 1, 4, 5] = 6.5, -4.0  # both outliers
    pkx[0, 0, 6, 2], pkx[0, 1, 6, 2] = 6.0, np.nan
    pkx[0, 0, 3, 8] = np.nan                     # primary fit failure, second good
    u, nm, reason, ph, choice, _ = M.peaks_stage(pkx, pky, pkh, None, mask, 8.0, 8.0, MEDIAN, True)
    assert reason[0, 2, 3] == 0 and choice[0, 2, 3] == 2 and not nm[0, 2, 3]
    assert abs(u[1, 0, 2, 3] - pkx[0, 1, 2, 3]) < 1e-6
    assert reason[0, 4, 5] == 12
    assert reason[0, 6, 2] == 13
    assert reason[0, 3, 8] == 0 and choice[0, 3, 8] == 2
    assert reason[0, 7, 9] == -1
    np.testing.assert_array_equal(reason != 0, nm)
    assert P == 2


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


def test_vector_masks_summed_area_equals_the_window_loop():
    """The device loop's summed-area vector masks equal the reference's
    per-window count, at every threshold, on blob and stripe masks."""
    from pivtools_amd.window import vector_masks

    shape = (160, 200)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    masks = [((yy - 70) ** 2 + (xx - 90) ** 2 < 30 ** 2), (xx % 37 < 5) | (yy > 150),
             np.random.default_rng(2).random(shape) < 0.3]
    sched, ov = [(64, 64), (32, 32), (16, 16)], [50, 75, 50]
    for pm in masks:
        for thr in (0.0, 0.25, 0.5, 1.0):
            a = vector_masks(shape, sched, ov, pm, thr)
            b = M.vector_masks_loop(shape, sched, ov, pm, thr)
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x, y)


def test_vector_mask_threshold_is_strict():
    """Exactly half a window masked is not more than 0.5 (hand count)."""
    shape = (64, 64)
    pm = np.zeros(shape, bool)
    pm[:, :16] = True                       # the first 32x32 window: 16 of 32 columns
    m = M.vector_masks_loop(shape, [(32, 32)], [50], pm, 0.5)[0]
    assert not m[0, 0]
    m = M.vector_masks_loop(shape, [(32, 32)], [50], pm, 0.49)[0]
    assert m[0, 0] and not m[0, 1]          # the second window (x 16..47) holds none


def _lstsq_fill(field, fill, spacing):
    """Brute force: every first-difference equation that touches a fill
    node, written out and solved with np.linalg.lstsq."""
    ny, nx = field.shape
    unk = {(r, c): i for i, (r, c) in enumerate(zip(*np.nonzero(fill)))}
    rows, rhs = [], []
    for (r, c), (r2, c2), h in ([((r, c), (r + 1, c), spacing[0]) for r in range(ny - 1) for c in range(nx)]
                                + [((r, c), (r, c + 1), spacing[1]) for r in range(ny) for c in range(nx - 1)]):
        if not (fill[r, c] or fill[r2, c2]):
            continue
        row = np.zeros(len(unk))
        b = 0.0
        for (rr, cc), s in (((r, c), -1.0 / h), ((r2, c2), 1.0 / h)):
            if fill[rr, cc]:
                row[unk[(rr, cc)]] += s
            else:
                b -= s * field[rr, cc]
        rows.append(row)
        rhs.append(b)
    x = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
    out = field.astype(np.float64).copy()
    for (r, c), i in unk.items():
        out[r, c] = x[i]
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_harmonic_fill_against_lstsq_and_closed_forms(seed):
    """The oracle's harmonic fill and the device path's host-side linear map
    (predictor_fill.harmonic_operator) against explicit least squares, a
    constant field and a linear field under an interior mask."""
    from pivtools_amd.predictor_fill import harmonic_operator

    rng = np.random.default_rng(seed)
    ny, nx, sp_ = 7, 9, (16.0, 8.0)
    fill = rng.random((ny, nx)) < 0.3
    fill[0, 0] = True
    f = rng.normal(0, 1, (ny, nx))
    ref = _lstsq_fill(f, fill, sp_)
    got = M.fill_harmonic(f[None], fill, sp_)[0]
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10)
    fi, bi, gt = harmonic_operator(fill, sp_)
    x = f.ravel().copy()
    x[fi] = gt.T @ f.ravel()[bi]
    np.testing.assert_allclose(x.reshape(ny, nx), ref, rtol=0, atol=1e-10)
    np.testing.assert_allclose(M.fill_harmonic(np.full((1, ny, nx), 2.5), fill, sp_), 2.5, atol=1e-12)
    inner = np.zeros((ny, nx), bool)
    inner[2:5, 3:7] = True
    yy, xx = np.mgrid[0:ny, 0:nx]
    lin = 0.3 + 0.05 * xx * sp_[1] - 0.02 * yy * sp_[0]
    np.testing.assert_allclose(M.fill_harmonic(lin[None], inner, sp_)[0], lin, atol=1e-12)


def test_harmonic_operator_refuses_an_all_masked_grid():
    from pivtools_amd.predictor_fill import harmonic_operator

    with pytest.raises(ValueError):
        harmonic_operator(np.ones((4, 5), bool), (8.0, 8.0))
    fi, bi, gt = harmonic_operator(np.zeros((4, 5), bool), (8.0, 8.0))
    assert fi.size == 0


def test_predictor_bcs_overwrite_rows_beyond_the_position():
    """base.py:364-416 by hand: a bottom BC at y_position 20 of a 64-row
    image sets rows with centre >= 43, a top BC at 5 rows with centre <= 5;
    a centre on each threshold pins the inclusive comparisons."""
    pad = np.zeros((2, 1, 6, 4), np.float32)
    ya = np.array([-10.0, 5.0, 20.0, 43.0, 50.0, 70.0])
    M.apply_predictor_bcs(pad, ya, 64, [{"edge": "bottom", "y_position": 20, "ux": 1.5, "uy": -0.5},
                                        {"edge": "top", "y_position": 5, "ux": 0.25, "uy": 0.0}])
    np.testing.assert_array_equal(pad[1, 0, :, 0], [0.25, 0.25, 0, 1.5, 1.5, 1.5])
    np.testing.assert_array_equal(pad[0, 0, :, 0], [0, 0, 0, -0.5, -0.5, -0.5])


def _lstsq_fill_general(field, fill, spacing, order, exclude):
    """Brute force for any order with excluded cells: the explicit difference
    equations touching a fill cell and no excluded cell, by np.linalg.lstsq."""
    ny, nx = field.shape
    unk = {(r, c): i for i, (r, c) in enumerate(zip(*np.nonzero(fill)))}
    st = (-1.0, 1.0) if order == 1 else (1.0, -2.0, 1.0)
    eqs = []
    for r in range(ny):
        for c in range(nx):
            for axis, h in ((0, spacing[0]), (1, spacing[1])):
                cells = [(r + j, c) if axis == 0 else (r, c + j) for j in range(order + 1)]
                if any(a >= ny or b >= nx for a, b in cells):
                    continue
                eqs.append((cells, [s / h ** order for s in st]))
    rows, rhs = [], []
    for cells, co in eqs:
        if not any(fill[a, b] for a, b in cells) or any(exclude[a, b] for a, b in cells):
            continue
        row = np.zeros(len(unk))
        bv = 0.0
        for (a, b), s in zip(cells, co):
            if fill[a, b]:
                row[unk[(a, b)]] += s
            else:
                bv -= s * field[a, b]
        rows.append(row)
        rhs.append(bv)
    x = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
    out = field.astype(np.float64).copy()
    for (a, b), i in unk.items():
        out[a, b] = x[i]
    return out


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("seed", [0, 1])
def test_least_squares_infill_against_lstsq_and_exact_fields(order, seed):
    rng = np.random.default_rng(10 + seed)
    ny, nx, sp_ = 8, 9, (16.0, 12.0)
    fill = rng.random((ny, nx)) < 0.2
    excl = (rng.random((ny, nx)) < 0.1) & ~fill
    f = rng.normal(0, 1, (ny, nx))
    ref = _lstsq_fill_general(f, fill, sp_, order, excl)
    got = M.fill_least_squares(f[None], fill, sp_, order, excl)[0]
    np.testing.assert_allclose(got[~excl], ref[~excl], rtol=0, atol=1e-9)
    yy, xx = np.mgrid[0:ny, 0:nx]
    const = np.full((1, ny, nx), -0.75)
    np.testing.assert_allclose(M.fill_least_squares(const, fill, sp_, order, excl), const, atol=1e-12)
    if order == 2:
        # the biharmonic fill reproduces a linear field for any hole, open edges included
        lin = 0.4 + 0.03 * xx * sp_[1] - 0.01 * yy * sp_[0]
        np.testing.assert_allclose(M.fill_least_squares(lin[None], fill, sp_, 2, excl)[0], lin, atol=1e-10)


def test_infill_ls_fills_nan_vectors_only():
    rng = np.random.default_rng(3)
    u = rng.normal(0, 1, (2, 2, 6, 7)).astype(np.float32)
    u[:, 0, 2, 3] = np.nan
    u[1, 1, 0, 0] = np.nan
    ex = np.zeros((6, 7), bool)
    ex[4, 4] = True
    out = M.infill_ls(u, (8.0, 8.0), 2, exclude=ex)
    assert np.isfinite(out).all()
    # a vector with one NaN component is filled whole (both components)
    keep = np.broadcast_to(~np.isnan(u).any(axis=0), u.shape)
    np.testing.assert_array_equal(out[keep], u[keep])
    assert out[0, 1, 0, 0] != u[0, 1, 0, 0]


def test_sigma_outliers_against_a_neighbour_loop():
    """The 8-neighbour mean/std (uniform_filter form) against explicit
    neighbour lists in FP64; nodes within 1e-4 of the threshold are skipped
    (float32 against float64 can differ there)."""
    rng = np.random.default_rng(4)
    ny, nx, k = 13, 17, 2.0
    v = rng.normal(1, 0.3, (ny, nx)).astype(np.float32)
    v[3, 4] = 9.0
    v[0, 5] = np.nan
    got = M.sigma_outliers(v, k)
    checked = 0
    for r in range(ny):
        for c in range(nx):
            nb = [float(v[rr, cc]) for rr in range(r - 1, r + 2) for cc in range(c - 1, c + 2)
                  if (rr, cc) != (r, c) and 0 <= rr < ny and 0 <= cc < nx and np.isfinite(v[rr, cc])]
            if not np.isfinite(v[r, c]) or not nb:
                assert got[r, c]
                continue
            m = np.mean(nb)
            sd = np.sqrt(max(np.mean(np.square(nb)) - m * m, 0.0))
            d = abs(float(v[r, c]) - m) - k * sd
            if abs(d) < 1e-4:
                continue
            assert got[r, c] == (d > 0 or sd == 0), (r, c)
            checked += 1
    assert checked > 150 and got[3, 4]


def test_div_vort_against_explicit_differences():
    rng = np.random.default_rng(6)
    ny, nx = 10, 12
    ux = rng.normal(0, 1, (ny, nx)).astype(np.float32)
    uy = rng.normal(0, 1, (ny, nx)).astype(np.float32)
    ux[4, 5] = 25.0
    div = np.zeros((ny, nx))
    vort = np.zeros((ny, nx))
    for r in range(ny):
        for c in range(nx):
            dudx = 0.5 * (ux[r, (c + 1) % nx] - ux[r, (c - 1) % nx])
            dudy = 0.5 * (ux[(r + 1) % ny, c] - ux[(r - 1) % ny, c])
            dvdx = 0.5 * (uy[r, (c + 1) % nx] - uy[r, (c - 1) % nx])
            dvdy = 0.5 * (uy[(r + 1) % ny, c] - uy[(r - 1) % ny, c])
            div[r, c], vort[r, c] = dudx + dvdy, dvdx - dudy

    def thr(x):
        s = np.sort(x.ravel())
        med = (s[(s.size - 1) // 2] + s[s.size // 2]) / 2
        d = np.sort(np.abs(x.ravel() - med))
        return 6 * 1.4826 * (d[(d.size - 1) // 2] + d[d.size // 2]) / 2

    want = (np.abs(div) > thr(div)) | (np.abs(vort) > thr(vort))
    got = M.div_vort_outliers(ux, uy)
    assert want.sum() >= 4
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(M.div_vort_outliers(ux, uy, 1e9, 1e9), np.zeros((ny, nx), bool))


def test_peaks_stage_without_substitution_equals_the_single_peak_stage():
    rng = np.random.default_rng(8)
    N, P, ny, nx = 2, 3, 11, 13
    pkx = rng.normal(0, 1.5, (N, P, ny, nx)).astype(np.float32)
    pky = rng.normal(0, 1.5, (N, P, ny, nx)).astype(np.float32)
    pkh = rng.uniform(0.05, 1, (N, P, ny, nx)).astype(np.float32)
    pkx[0, 0, 3, 3] = np.nan
    pred = rng.normal(0, 0.5, (2, N, ny, nx)).astype(np.float32)
    m = np.zeros((ny, nx), bool)
    m[5, 6] = True
    methods = [{"type": "median_2d"}, {"type": "peak_mag", "threshold": 0.1}]
    u, nm, rs, ppm, ch, _ = M.peaks_stage(pkx, pky, pkh, pred, m, 4.0, 4.0, methods, False)
    u1, nm1, rs1, pk1 = M.gate(pkx[:, 0], pky[:, 0], pkh[:, 0], pred, m, 4.0, 4.0)
    u1, nm1, rs1, pk1 = M.outlier_stage(u1, pk1, nm1, rs1, m, methods)
    np.testing.assert_array_equal(nm, nm1)
    np.testing.assert_array_equal(rs, rs1)
    np.testing.assert_array_equal(u, u1)
    np.testing.assert_array_equal(ppm, pk1)
    np.testing.assert_array_equal(ch, np.where(nm, 0, 1))


def test_secondary_peak_substitution_by_hand():
    """Hand-built windows: an outlier rescued by its second peak, one whose
    further peaks are an outlier and then gated out, a gated primary with a
    good second peak, and a masked window."""
    N, P, ny, nx = 1, 3, 9, 9
    pkx = np.full((N, P, ny, nx), 1.0, np.float32)
    pky = np.full((N, P, ny, nx), 0.5, np.float32)
    pkh = np.full((N, P, ny, nx), 0.8, np.float32)
    pkx[0, :, 4, 4] = [3.5, 1.02, 0.0]          # outlier, then a consistent second peak
    pkx[0, :, 2, 2] = [3.5, -3.0, 9.0]          # outlier, outlier, gated out (|x| > 4)
    pkx[0, :, 6, 6] = [9.0, 0.98, 0.0]          # gated primary, good second peak
    m = np.zeros((ny, nx), bool)
    m[0, 8] = True
    u, nm, rs, ppm, ch, pha = M.peaks_stage(pkx, pky, pkh, None, m, 4.0, 4.0, [{"type": "median_2d"}], True)
    assert not nm[0, 4, 4] and ch[0, 4, 4] == 2 and rs[0, 4, 4] == 0 and u[1, 0, 4, 4] == np.float32(1.02)
    assert nm[0, 2, 2] and ch[0, 2, 2] == 0 and rs[0, 2, 2] == 13 and np.isnan(u[1, 0, 2, 2])
    assert not nm[0, 6, 6] and ch[0, 6, 6] == 2 and u[1, 0, 6, 6] == np.float32(0.98)
    assert nm[0, 0, 8] and rs[0, 0, 8] == -1 and u[1, 0, 0, 8] == 0.0 and ch[0, 0, 8] == 0
    assert np.isnan(pha[0, 0, 6, 6]) and np.isnan(pha[0, 2, 2, 2])
    assert np.array_equal(rs != 0, nm)
    # without substitution the same windows stay invalid
    u2, nm2, rs2, _, ch2, _ = M.peaks_stage(pkx, pky, pkh, None, m, 4.0, 4.0, [{"type": "median_2d"}], False)
    assert nm2[0, 4, 4] and rs2[0, 4, 4] == 10 and rs2[0, 6, 6] == 6


def test_oracle_loop_converges_on_poiseuille_flow():
    """unit-tests/test_multipass_convergence.py:398-430, the end-to-end pin
    of the predictor remap and warp: three passes 64 -> 32 -> 16 on the
    Poiseuille pair recover u_x to a median error below 0.3 px and u_y to
    below 0.15 px over the interior rows, on every pass after the first."""
    import poiseuille as P

    a, b = P.pair()
    grids = build_pass_grids(P.SHAPE, P.SCHEDULE, P.OVERLAP)
    res = M.run_multipass(a[None].astype(np.float32), b[None].astype(np.float32), grids, finder=6,
                          outlier_methods=P.OUTLIERS, mid_infill="nearest", final_infill="biharmonic",
                          keep_passes=True)
    for p in (1, 2):
        ux, uy, nm = res[p]["ux"][0], res[p]["uy"][0], res[p]["nan_mask"][0]
        ex, ey = P.errors(ux, uy, ~nm & np.isfinite(ux))
        assert ex < 0.3 and ey < 0.15, (p, ex, ey)
    assert ex < 0.05


@pytest.mark.parametrize("weights", ["distance", "uniform"])
@pytest.mark.parametrize("spacing", [(16.0, 16.0), (8.0, 12.0)])
def test_knn_infill_oracle_against_sklearn(weights, spacing):
    """The oracle's k-NN infill (donors ranked by distance, then row-major
    index) equals sklearn's KNeighborsRegressor -- the library the reference
    calls (infilling.py:609-695) -- at every hole whose k-th and (k+1)-th
    donor distances differ (where they tie, a KD-tree may keep another donor
    of the same distance); and a linear field is reproduced by neither
    weighting only approximately, so the values are checked, not the field."""
    rng = np.random.default_rng(31)
    N, ny, nx = 2, 17, 23
    u = rng.normal(size=(2, N, ny, nx)).astype(np.float32)
    holes = rng.uniform(size=(N, ny, nx)) < 0.15
    holes[0, 0:4, 0:5] = True
    u[:, holes] = np.nan
    ex = np.zeros((ny, nx), bool)
    ex[8, 10:13] = True
    for k in (1, 5, 32):
        a = M.infill_knn(u, spacing, k=k, weights=weights, exclude=ex, impl="brute")
        b = M.infill_knn(u, spacing, k=k, weights=weights, exclude=ex, impl="sklearn")
        yy, xx = np.mgrid[0:ny, 0:nx]
        for i in range(N):
            fill = (np.isnan(u[0, i]) | np.isnan(u[1, i])) & ~ex
            donor = ~fill & ~ex
            dpts = np.c_[(xx * spacing[1])[donor], (yy * spacing[0])[donor]]
            for r, c in zip(*np.nonzero(fill)):
                d = np.sort(np.hypot(dpts[:, 0] - c * spacing[1], dpts[:, 1] - r * spacing[0]))
                if k < len(d) and d[k - 1] == d[k]:
                    continue
                np.testing.assert_allclose(a[:, i, r, c], b[:, i, r, c], rtol=1e-6, atol=1e-7)
            assert np.isfinite(a[:, i][:, fill]).all()
            np.testing.assert_array_equal(a[:, i][:, ~fill], u[:, i][:, ~fill])


RBF_EPS = {"linear": None, "thin_plate_spline": None, "cubic": None, "quintic": None, "multiquadric": 0.4,
           "inverse_multiquadric": 0.4, "inverse_quadratic": 0.4, "gaussian": 0.4}


def _rbf_case(seed, ny=15, nx=19, frac=0.12):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:ny, 0:nx]
    u = np.stack([np.sin(0.3 * xx + 0.2 * yy) + 0.1 * rng.normal(size=(ny, nx)),
                  np.cos(0.25 * yy) * 0.7 + 0.1 * rng.normal(size=(ny, nx))])[:, None].astype(np.float32)
    u = np.concatenate([u, u[:, :, ::-1, :]], axis=1)          # two images
    holes = rng.uniform(size=u.shape[1:]) < frac
    holes[0, 0:3, 0:4] = True
    u[:, holes] = np.nan
    ex = np.zeros((ny, nx), bool)
    ex[7, 9:12] = True
    return u, ex


@pytest.mark.parametrize("kernel", sorted(RBF_EPS))
@pytest.mark.parametrize("spacing,smoothing", [((1.0, 1.0), 0.0), ((8.0, 12.0), 0.0), ((1.0, 1.0), 0.05)])
def test_rbf_local_oracle_against_scipy(kernel, spacing, smoothing):
    """The oracle's local RBF infill equals scipy's RBFInterpolator with
    `neighbors` -- what the reference calls (infilling.py:505-597) -- at every
    hole whose k-th and (k+1)-th donor distances differ (where they tie, the
    KD-tree may keep another donor of the same distance).  Smoothing > 0 also
    pins the kernels' signs (without it c * phi and (-c) * (-phi) agree)."""
    u, ex = _rbf_case(7)
    eps = RBF_EPS[kernel]
    ny, nx = u.shape[2:]
    yy, xx = np.mgrid[0:ny, 0:nx]
    # 8 neighbours can lie on two grid lines, where the quadratic monomials
    # of the quintic lose rank (scipy raises "Singular matrix")
    for k in ((20, 64) if kernel == "quintic" else (8, 20, 64)):
        a = M.infill_rbf_local(u, spacing, neighbors=k, kernel=kernel, epsilon=eps, smoothing=smoothing,
                               exclude=ex, impl="brute")
        b = M.infill_rbf_local(u, spacing, neighbors=k, kernel=kernel, epsilon=eps, smoothing=smoothing,
                               exclude=ex, impl="scipy")
        checked = 0
        for i in range(u.shape[1]):
            fill = (np.isnan(u[0, i]) | np.isnan(u[1, i])) & ~ex
            donor = ~fill & ~ex
            dpts = np.c_[(xx * spacing[1])[donor], (yy * spacing[0])[donor]]
            for r, c in zip(*np.nonzero(fill)):
                d = np.sort(np.sqrt((dpts[:, 0] - c * spacing[1]) ** 2 + (dpts[:, 1] - r * spacing[0]) ** 2))
                if k < len(d) and d[k - 1] == d[k]:
                    continue
                np.testing.assert_allclose(a[:, i, r, c], b[:, i, r, c], rtol=2e-6, atol=2e-6)
                checked += 1
            assert np.isfinite(a[:, i][:, fill]).all()
            np.testing.assert_array_equal(a[:, i][:, ~fill], u[:, i][:, ~fill])
        assert checked >= 5, checked


@pytest.mark.parametrize("kernel,field", [("thin_plate_spline", "linear"), ("cubic", "linear"),
                                          ("quintic", "quadratic"), ("linear", "constant"),
                                          ("gaussian", "constant")])
def test_rbf_local_reproduces_its_polynomials(kernel, field):
    """An interpolant with polynomial terms of degree d reproduces every
    polynomial of degree <= d (the RBF coefficients vanish): linear fields
    for the thin-plate spline and cubic, quadratics for the quintic,
    constants for the rest."""
    ny, nx = 13, 16
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    X, Y = xx * 3.0, yy * 2.0
    f = {"linear": 0.3 + 0.02 * X - 0.05 * Y, "quadratic": 0.1 + 0.01 * X * X - 0.004 * X * Y + 0.02 * Y * Y,
         "constant": 0.75 + 0 * X}[field]
    u = np.stack([f, -f])[:, None].astype(np.float32)
    holes = np.zeros((1, ny, nx), bool)
    holes[0, 5:8, 6:9] = True
    holes[0, 0, 0] = holes[0, 12, 15] = True
    u[:, holes] = np.nan
    got = M.infill_rbf_local(u, (2.0, 3.0), neighbors=30, kernel=kernel, epsilon=RBF_EPS[kernel])
    want = np.stack([f, -f])[:, None].astype(np.float32)
    np.testing.assert_allclose(got[:, holes], want[:, holes], rtol=0, atol=2e-6)


def test_rbf_local_needs_four_donors():
    u = np.full((2, 1, 5, 5), np.nan, np.float32)
    u[:, 0, 0, 0:3] = 1.0
    with pytest.raises(ValueError):
        M.infill_rbf_local(u, (1.0, 1.0))
    with pytest.raises(ValueError):
        M.infill_rbf_local(u, (1.0, 1.0), impl="scipy")
