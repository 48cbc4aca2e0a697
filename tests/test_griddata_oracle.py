"""The oracle's griddata_linear infill (oracle/multipass.py
infill_griddata_linear) pinned three ways: linear fields are reproduced
exactly (any triangulation interpolates a plane exactly); at every hole the
oracle's value and scipy's griddata value (the reference's call,
infilling.py:440-503) are both values of globally Delaunay triangles holding
the hole, found here by brute force over all donors; and where that triangle
is unique the two agree."""

import os
import sys

import numpy as np
import pytest
from scipy.interpolate import griddata

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import multipass as M  # noqa: E402


def _scipy(u, mask, spacing):
    """The reference's call: linear griddata in physical coordinates, nearest
    for what is left, only the holes replaced."""
    ny, nx = mask.shape
    yy, xx = np.mgrid[0:ny, 0:nx]
    yy, xx = yy * spacing[0], xx * spacing[1]
    out = u.copy()
    for comp in range(2):
        f = u[comp, 0]
        pts = np.c_[yy[~mask], xx[~mask]]
        lin = griddata(pts, f[~mask], (yy, xx), method="linear")
        near = griddata(pts, f[~mask], (yy, xx), method="nearest")
        g = np.where(np.isnan(lin), near, lin)
        out[comp, 0][mask] = g[mask].astype(np.float32)
    return out


def _delaunay_values(u, mask, spacing, q):
    """Values at hole q of every globally Delaunay triangle holding it (brute
    force over all donors near the hole, whose circumcircles are checked
    against every donor, FP64), or None when q is outside every triangle."""
    ny, nx = mask.shape
    r, c = np.divmod(np.flatnonzero(~mask.ravel()), nx)
    qr, qc = divmod(q, nx)
    Y, X = (r - qr) * float(spacing[0]), (c - qc) * float(spacing[1])
    f0, f1 = u[0, 0][~mask].astype(np.float64), u[1, 0][~mask].astype(np.float64)
    idx = np.flatnonzero((Y * Y + X * X) < (4.5 * max(spacing)) ** 2)
    t = M._triples(idx.size)
    a, b, cc = idx[t[:, 0]], idx[t[:, 1]], idx[t[:, 2]]
    ar = (X[b] - X[a]) * (Y[cc] - Y[a]) - (Y[b] - Y[a]) * (X[cc] - X[a])
    oa = X[b] * Y[cc] - Y[b] * X[cc]
    ob = X[cc] * Y[a] - Y[cc] * X[a]
    oc = X[a] * Y[b] - Y[a] * X[b]
    s = np.sign(ar)
    keep = (ar != 0) & (oa * s >= 0) & (ob * s >= 0) & (oc * s >= 0)
    a, b, cc, ar, oa, ob, oc, s = a[keep], b[keep], cc[keep], ar[keep], oa[keep], ob[keep], oc[keep], s[keep]
    if a.size == 0:
        return None
    vals = []
    for k in range(a.size):
        adx, ady = X[a[k]] - X, Y[a[k]] - Y
        bdx, bdy = X[b[k]] - X, Y[b[k]] - Y
        cdx, cdy = X[cc[k]] - X, Y[cc[k]] - Y
        det = ((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy)
               + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy)
               + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady)) * s[k]
        if (det > 1e-9 * np.abs(det).max()).any():
            continue
        la, lb, lc = oa[k] / ar[k], ob[k] / ar[k], oc[k] / ar[k]
        vals.append((la * f0[a[k]] + lb * f0[b[k]] + lc * f0[cc[k]], la * f1[a[k]] + lb * f1[b[k]] + lc * f1[cc[k]]))
    return np.array(vals) if vals else None


def _case(seed, ny, nx, frac, block=None):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(2, 1, ny, nx)).astype(np.float32)
    mask = rng.uniform(size=(ny, nx)) < frac
    if block is not None:
        mask[block] = True
    g = u.copy()
    g[:, 0][:, mask] = np.nan
    return g, mask


@pytest.mark.parametrize("spacing", [(16.0, 16.0), (16.0, 8.0), (12.0, 20.0)])
def test_linear_fields_exact(spacing):
    ny, nx = 13, 16
    g, mask = _case(1, ny, nx, 0.2, (slice(4, 8), slice(5, 9)))
    yy, xx = np.mgrid[0:ny, 0:nx]
    lin = np.stack([(0.3 * yy - 0.7 * xx + 2)[None], (1.1 * yy + 0.2 * xx - 1)[None]]).astype(np.float32)
    h = lin.copy()
    h[:, 0][:, mask] = np.nan
    mask_inner = mask.copy()
    o = M.infill_griddata_linear(h, spacing)
    # inside the convex hull every triangle reproduces the plane
    sc = _scipy(h, mask, spacing)
    inside = np.abs(sc[0, 0] - lin[0, 0]) < 1e-4
    assert np.abs(o[:, 0][:, mask_inner & inside] - lin[:, 0][:, mask_inner & inside]).max() < 2e-6


@pytest.mark.parametrize("seed,spacing", [(2, (16.0, 16.0)), (3, (16.0, 8.0)), (4, (10.0, 10.0))])
def test_values_are_delaunay_values_and_match_scipy_where_unique(seed, spacing):
    ny, nx = 12, 14
    g, mask = _case(seed, ny, nx, 0.18, (slice(3, 6), slice(4, 7)))
    o = M.infill_griddata_linear(g, spacing)
    s = _scipy(g, mask, spacing)
    n_unique = n_checked = 0
    for q in np.flatnonzero(mask.ravel()):
        r, c = divmod(q, nx)
        vals = _delaunay_values(g, mask, spacing, q)
        ours = np.array([o[0, 0, r, c], o[1, 0, r, c]], np.float64)
        theirs = np.array([s[0, 0, r, c], s[1, 0, r, c]], np.float64)
        if vals is None:            # outside the hull: both take a nearest donor
            continue
        n_checked += 1
        assert np.abs(vals - ours).max(axis=1).min() < 2e-6, (q, ours, vals)
        assert np.abs(vals - theirs).max(axis=1).min() < 2e-6, (q, theirs, vals)
        uniq = np.unique(np.round(vals, 6), axis=0)
        if len(uniq) == 1:
            n_unique += 1
            assert np.abs(ours - theirs).max() < 2e-6
    assert n_checked > 10 and n_unique > 0


def test_planted_mistake_fails():
    """Barycentric weights applied to the wrong vertices are caught."""
    ny, nx = 10, 11
    g, mask = _case(7, ny, nx, 0.15)
    o = M.infill_griddata_linear(g, (16.0, 16.0))
    q = np.flatnonzero(mask.ravel())[0]
    r, c = divmod(q, nx)
    vals = _delaunay_values(g, mask, (16.0, 16.0), q)
    bad = np.array([o[0, 0, r, c] + 0.05, o[1, 0, r, c]])
    assert np.abs(vals - bad).max(axis=1).min() > 1e-3
