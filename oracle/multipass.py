"""CPU oracle of the instantaneous multipass loop -- TEST INFRASTRUCTURE ONLY.

NumPy/SciPy restatement of the host glue of
pivtools_cli/piv/piv_backend/cpu_instantaneous.py:381-1147 around the two C
calls, with the C calls taken from oracle.py (liboracle.so) or, for the
FP64 flavour, from FP64 NumPy planes.

Pins (each checked in tests/test_multipass_oracle.py so that a plausible
mistake fails):
  gate / reason codes   restated from cpu_instantaneous.py:576-757; known
                        answers for every code (-1, 1, 6, 9, 10, 12, 13), their
                        precedence, the corner guard rule, the secondary-peak
                        rescue and unit-tests/test_nan_reason.py's masked block
  median test           outlier_detection.py:87-243 restated; checked against
                        a brute-force per-node loop in the tests
  nearest infill        scipy.ndimage.distance_transform_edt, as the reference
                        calls it (infilling.py:114-141)
  smoothing             scipy.ndimage.gaussian_filter, as the reference calls
                        it (cpu_instantaneous.py:1135-1147)
  predictor remap       cv2.remap(INTER_CUBIC, BORDER_REPLICATE) is not
                        importable here: restated from OpenCV's published
                        algorithm (float map -> 1/32 fixed point with ties to
                        even, Keys a=-0.75 per axis, replicate border, row sums
                        inside the field and term-by-term sums at a border);
                        pinned by positions off the lattice and on its ties,
                        every border against an FP64 evaluation on an
                        edge-padded copy, and closed-form fields
  vector masks          load_images.py:952-985 restated as the reference's
                        per-window loop; checked against hand-counted windows
  predictor fill        infilling.py:808-870: 'nearest' through
                        distance_transform_edt as the reference calls it;
                        'harmonic' as the reference's difference-row least
                        squares (infilling.py:144-260), checked against
                        np.linalg.lstsq on the explicit equations, constant
                        and interior-linear fields
  least-squares infill  infilling.py:144-260 restated (difference rows,
                        normal equations, sparse LU), order 1 and 2 with
                        excluded cells; checked against np.linalg.lstsq on
                        the explicit equations and the fields each order
                        reproduces exactly (constant; linear for order 2)
  predictor BCs         base.py:364-416 restated (row overwrite), checked by
                        hand-computed rows with a centre on each threshold
  k-NN infill           infilling.py:600-695; equals sklearn's
                        KNeighborsRegressor where no tie sits at the k-th
                        distance
  local RBF infill      infilling.py:505-597; equals scipy's RBFInterpolator
                        (neighbors) where no tie sits at the k-th distance,
                        and reproduces the polynomials of its degree
  end to end            the Poiseuille bound of
                        unit-tests/test_multipass_convergence.py:398-430
"""

from __future__ import annotations

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view
from scipy import ndimage as ndi

import oracle as O


def gate(pkx, pky, pkh, pred, mask, tx, ty):
    """cpu_instantaneous.py:576-635 for one peak.  pk* (N, ny, nx); pred
    (2, N, ny, nx) planar [dy, dx] or None; returns u planar, nan_mask,
    nan_reason, peak height with NaN at gated vectors."""
    pkx = pkx.astype(np.float32).copy()
    pky = pky.astype(np.float32).copy()
    pkh = pkh.astype(np.float32).copy()
    N = pkx.shape[0]
    m = np.zeros(pkx.shape[1:], bool) if mask is None else mask.astype(bool)
    mb = np.broadcast_to(m, pkx.shape)
    large = (np.abs(pkx) > tx) | (np.abs(pky) > ty)
    fit_fail = np.isnan(pkx) | np.isnan(pky)
    inv = mb | large
    pkx[inv] = np.nan
    pky[inv] = np.nan
    pkh[inv] = np.nan
    if pred is not None:
        pkx += pred[1]
        pky += pred[0]
    nm = np.isnan(pkx) | np.isnan(pky) | mb
    ux = pkx.copy()
    uy = pky.copy()
    ux[nm] = 0.0
    uy[nm] = 0.0
    reason = np.zeros(pkx.shape, np.int8)
    reason[fit_fail] = 1
    reason[large] = 6
    del N
    return np.stack([uy, ux]).astype(np.float32), nm, reason, pkh


def median_outliers(ux, uy, eps=0.1, thr=2.0, size=5, masked=None):
    """outlier_detection.py:87-243 (combine='norm'), nanmedian via NumPy."""
    ux = ux.astype(np.float32, copy=True)
    uy = uy.astype(np.float32, copy=True)
    if masked is not None:
        ux[masked] = np.nan
        uy[masked] = np.nan

    def nbrs(U):
        pad = size // 2
        Up = np.pad(U, pad, mode="constant", constant_values=np.nan)
        win = sliding_window_view(Up, (size, size)).reshape(U.shape[0], U.shape[1], size * size)
        c = (size * size) // 2
        return np.concatenate([win[..., :c], win[..., c + 1:]], axis=-1)

    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        uxn, uyn = nbrs(ux), nbrs(uy)
        mu = np.nanmedian(uxn, axis=-1).astype(np.float32)
        mv = np.nanmedian(uyn, axis=-1).astype(np.float32)
        r0 = np.sqrt((ux - mu) ** 2 + (uy - mv) ** 2)
        ri = np.sqrt((uxn - mu[..., None]) ** 2 + (uyn - mv[..., None]) ** 2)
        rm = np.nanmedian(ri, axis=-1).astype(np.float32)
        nr = r0 / (rm + np.float32(eps))
    exceeded = (nr > thr) | ~np.isfinite(nr)
    present = np.isfinite(ux) & np.isfinite(uy)
    if masked is not None:
        present |= masked
    npres = ndi.correlate(present.astype(np.float32), np.ones((3, 3), np.float32), mode="constant", cval=0.0)
    out = exceeded | (npres < 6)
    if masked is not None:
        out &= ~masked
    return out


def _detect(ux, uy, pkh, mask, methods):
    """apply_outlier_detection on one image's (ux, uy) (NaN allowed)."""
    m = np.zeros(ux.shape, bool) if mask is None else mask
    om = np.zeros(ux.shape, bool)
    ux = ux.astype(np.float32)
    uy = uy.astype(np.float32)
    if mask is not None:
        ux = np.where(m, np.float32(np.nan), ux)
        uy = np.where(m, np.float32(np.nan), uy)
    for mc in methods:
        t = mc["type"]
        if t == "median_2d":
            om |= median_outliers(ux, uy, mc.get("epsilon", 0.1), mc.get("threshold", 2.0), mc.get("size", 5),
                                  masked=m if mask is not None else None)
        elif t == "peak_mag":
            om |= (pkh < mc.get("threshold", 0.5)) | np.isnan(pkh)
        elif t == "sigma":
            with np.errstate(invalid="ignore"):
                om |= sigma_outliers(np.sqrt(ux ** 2 + uy ** 2), mc.get("sigma_threshold", 2.0))
        elif t == "div_vort":
            om |= div_vort_outliers(ux, uy, mc.get("div_thresh"), mc.get("vort_thresh"))
        else:
            raise ValueError(t)
    return om & ~m


def peaks_stage(pkx, pky, pkh, pred, mask, tx, ty, methods, secondary):
    """Several peaks per window (cpu_instantaneous.py:586-760): pk* (N, P,
    ny, nx).  Gate every peak, add the predictor, test the primary peak,
    then (secondary) move invalid windows to their next peak and re-test
    them; returns u (NaN at invalid, 0 at masked), nan_mask, nan_reason,
    the chosen peak's height (NaN at invalid), peak_choice and the gated
    heights of every peak."""
    pkx = pkx.astype(np.float32).copy()
    pky = pky.astype(np.float32).copy()
    pkh = pkh.astype(np.float32).copy()
    N, P, ny, nx = pkx.shape
    m = np.zeros((ny, nx), bool) if mask is None else mask.astype(bool)
    mb = np.broadcast_to(m, (N, ny, nx))
    large = (np.abs(pkx) > tx) | (np.abs(pky) > ty)
    fit_fail0 = np.isnan(pkx[:, 0]) | np.isnan(pky[:, 0])
    large0 = large[:, 0]
    inv = np.broadcast_to(mb[:, None], pkx.shape) | large
    pkx[inv] = np.nan
    pky[inv] = np.nan
    pkh[inv] = np.nan
    if pred is not None:
        pkx += pred[1][:, None]
        pky += pred[0][:, None]
    ux = pkx[:, 0].copy()
    uy = pky[:, 0].copy()
    nm = np.isnan(ux) | np.isnan(uy) | mb
    ux[nm] = 0.0
    uy[nm] = 0.0
    choice = np.ones((N, ny, nx), np.int64)
    reason = np.zeros((N, ny, nx), np.int8)
    reason[fit_fail0] = 1
    reason[large0] = 6
    if methods:
        for i in range(N):
            om = _detect(ux[i], uy[i], pkh[i, 0], mask, methods)
            nm[i] |= om
            reason[i][om & (reason[i] == 0)] = 10
    if secondary:
        for _ in range(1, P):
            act = nm & (choice < P)
            if not act.any():
                break
            choice[act] += 1
            idx = (choice - 1)[:, None]
            ux = np.take_along_axis(pkx, idx, axis=1)[:, 0]
            uy = np.take_along_axis(pky, idx, axis=1)[:, 0]
            sub_nan = np.isnan(ux) | np.isnan(uy)
            nm[act] = sub_nan[act]
            reason[act] = np.where(sub_nan[act], 13, 0)
            if methods:
                ph = np.take_along_axis(pkh, idx, axis=1)[:, 0]
                for i in range(N):
                    if act[i].any():
                        om = _detect(ux[i], uy[i], ph[i], mask, methods)
                        rf = act[i] & om & ~sub_nan[i]
                        nm[i] |= rf
                        reason[i][rf] = 12
        nm |= mb
    idx = (choice - 1)[:, None]
    ppm = np.take_along_axis(pkh, idx, axis=1)[:, 0]
    nm |= np.isnan(ppm)
    reason[np.isnan(ppm) & (reason == 0)] = 13
    reason[nm & (reason == 0)] = 9
    reason[mb] = -1
    ux = ux.copy()
    uy = uy.copy()
    ux[nm] = np.nan
    uy[nm] = np.nan
    ppm = ppm.copy()
    ppm[nm] = np.nan
    ux[mb] = 0.0
    uy[mb] = 0.0
    choice[nm] = 0
    return np.stack([uy, ux]).astype(np.float32), nm, reason, ppm, choice, pkh


def median_stage(u, pkh, nanm, reason, mask=None, eps=0.1, thr=2.0, size=5):
    """Median test plus the bookkeeping of cpu_instantaneous.py:639-757."""
    return outlier_stage(u, pkh, nanm, reason, mask,
                         [{"type": "median_2d", "epsilon": eps, "threshold": thr, "size": size}])


def sigma_outliers(field, k):
    """outlier_detection.py:244-313 on one scalar field: 8-neighbour mean and
    standard deviation from scipy.ndimage.uniform_filter (the call the
    reference makes), float32 throughout."""
    v = np.asarray(field).astype(np.float32)
    fin = np.isfinite(v)
    v0 = np.where(fin, v, np.float32(0.0)).astype(np.float32)
    cnt = fin.astype(np.float32)
    s9 = ndi.uniform_filter(v0, size=3, mode="constant", cval=0.0) * np.float32(9.0)
    q9 = ndi.uniform_filter(v0 * v0, size=3, mode="constant", cval=0.0) * np.float32(9.0)
    c9 = ndi.uniform_filter(cnt, size=3, mode="constant", cval=0.0) * np.float32(9.0)
    s8 = s9 - v0
    q8 = q9 - v0 * v0
    c8 = c9 - cnt
    with np.errstate(invalid="ignore", divide="ignore"):
        m8 = s8 / c8
        var = q8 / c8 - m8 * m8
    var = np.maximum(var, np.float32(0.0))
    sd = np.sqrt(var)
    with np.errstate(invalid="ignore"):
        return (np.abs(v - m8) > np.float32(k) * sd) | ~fin | ~np.isfinite(m8) | (sd == 0) | (c8 < 1)


def div_vort_outliers(ux, uy, div_thresh=None, vort_thresh=None):
    """outlier_detection.py:316-373: periodic central differences, thresholds
    6 * 1.4826 * MAD of the finite values unless given."""
    h = np.float32(0.5)
    dudx = h * (np.roll(ux, -1, 1) - np.roll(ux, 1, 1))
    dudy = h * (np.roll(ux, -1, 0) - np.roll(ux, 1, 0))
    dvdx = h * (np.roll(uy, -1, 1) - np.roll(uy, 1, 1))
    dvdy = h * (np.roll(uy, -1, 0) - np.roll(uy, 1, 0))
    div = dudx + dvdy
    vort = dvdx - dudy

    def auto(x):
        f = x[np.isfinite(x)]
        if f.size == 0:
            return np.float32(np.inf)
        med = np.median(f)
        return np.float32(6) * (np.float32(1.4826) * np.median(np.abs(f - med)))

    td = auto(div) if div_thresh is None else np.float32(div_thresh)
    tv = auto(vort) if vort_thresh is None else np.float32(vort_thresh)
    with np.errstate(invalid="ignore"):
        return (np.abs(div) > td) | (np.abs(vort) > tv)


def outlier_stage(u, pkh, nanm, reason, mask, methods):
    """apply_outlier_detection (outlier_detection.py:420-500) per image with
    the median_2d and peak_mag methods, then the bookkeeping of
    cpu_instantaneous.py:639-757."""
    N = u.shape[1]
    nanm = nanm.copy()
    reason = reason.copy()
    m = np.zeros(u.shape[2:], bool) if mask is None else mask.astype(bool)
    for i in range(N):
        om = np.zeros(m.shape, bool)
        for mc in methods:
            t = mc["type"]
            if t == "median_2d":
                om |= median_outliers(u[1, i], u[0, i], mc.get("epsilon", 0.1), mc.get("threshold", 2.0),
                                      mc.get("size", 5), masked=m if mask is not None else None)
            elif t == "peak_mag":
                om |= (pkh[i] < mc.get("threshold", 0.5)) | np.isnan(pkh[i])
            elif t in ("sigma", "div_vort"):
                ux = u[1, i].astype(np.float32)
                uy = u[0, i].astype(np.float32)
                if mask is not None:
                    ux = np.where(m, np.float32(np.nan), ux)
                    uy = np.where(m, np.float32(np.nan), uy)
                if t == "sigma":
                    om |= sigma_outliers(np.sqrt(ux ** 2 + uy ** 2), mc.get("sigma_threshold", 2.0))
                else:
                    om |= div_vort_outliers(ux, uy, mc.get("div_thresh"), mc.get("vort_thresh"))
            else:
                raise ValueError(t)
        om &= ~m
        reason[i][om & (reason[i] == 0)] = 10
        nanm[i] |= om
    hn = np.isnan(pkh)
    nanm |= hn
    reason[hn & (reason == 0)] = 13
    reason[nanm & (reason == 0)] = 9
    mb = np.broadcast_to(m, nanm.shape)
    reason[mb] = -1
    uo = u.copy()
    uo[:, nanm] = np.nan
    uo[:, mb] = 0.0
    pk = pkh.copy()
    pk[nanm] = np.nan
    return uo, nanm, reason, pk


def infill_nearest(u, spacing, exclude=None):
    """Nearest-donor fill per image (infilling.py:114-141, 290-333)."""
    out = u.astype(np.float64).copy()
    for i in range(u.shape[1]):
        ux, uy = u[1, i], u[0, i]
        fill = np.isnan(ux) | np.isnan(uy)
        if not fill.any():
            continue
        donor = ~fill & np.isfinite(ux) & np.isfinite(uy)
        if exclude is not None:
            donor &= ~exclude
        if not donor.any():
            raise ValueError("no donor")
        iy, ix = ndi.distance_transform_edt(~donor, sampling=spacing, return_distances=False, return_indices=True)
        for c in range(2):
            f = out[c, i]
            f[fill] = u[c, i][iy, ix][fill]
    return out.astype(np.float32)


def infill_knn(u, spacing, k=32, weights="distance", exclude=None, impl="brute"):
    """infill_knn (infilling.py:609-695) per image: NaN vectors (not excluded)
    become the k-nearest-neighbour regression of the valid vectors in
    physical coordinates (x * px, y * py), k = min(k, donors), weights
    1 / distance or uniform, FP64, stored as float32.  impl="sklearn" calls
    KNeighborsRegressor (kd_tree) as the reference does; impl="brute" ranks
    the donors by (distance, row-major index) and sums in that order -- the
    rule the device follows where several donors share the k-th distance."""
    py, px = float(spacing[0]), float(spacing[1])
    out = u.astype(np.float32).copy()
    ny, nx = u.shape[2:]
    yy, xx = np.mgrid[0:ny, 0:nx]
    X = np.c_[(xx * px).ravel(), (yy * py).ravel()].astype(np.float64)
    for i in range(u.shape[1]):
        ux, uy = u[1, i].ravel(), u[0, i].ravel()
        ex = np.zeros(ny * nx, bool) if exclude is None else np.asarray(exclude, bool).ravel()
        fill = (np.isnan(ux) | np.isnan(uy)) & ~ex
        if not fill.any():
            continue
        donor = ~fill & ~ex & np.isfinite(ux) & np.isfinite(uy)
        nd = int(donor.sum())
        if nd == 0:
            raise ValueError("no donor")
        kk = min(k, nd)
        Xd, Xq = X[donor], X[fill]
        yd = np.c_[ux[donor].astype(np.float64), uy[donor].astype(np.float64)]
        if impl == "sklearn":
            from sklearn.neighbors import KNeighborsRegressor
            knn = KNeighborsRegressor(n_neighbors=kk, weights=weights, algorithm="kd_tree", n_jobs=1)
            pred = knn.fit(Xd, yd).predict(Xq)
        else:
            idx = np.flatnonzero(donor)
            pred = np.empty((len(Xq), 2))
            for q, xq in enumerate(Xq):
                d = np.sqrt(((Xd - xq) ** 2).sum(axis=1))
                order = np.lexsort((idx, d))[:kk]
                wts = np.ones(kk) if weights == "uniform" else 1.0 / d[order]
                sy = sx = sw = 0.0
                for j, o in enumerate(order):
                    sx += yd[o, 0] * wts[j]
                    sy += yd[o, 1] * wts[j]
                    sw += wts[j]
                pred[q] = (sx / sw, sy / sw)
        out[1, i].ravel()[fill] = pred[:, 0].astype(np.float32)
        out[0, i].ravel()[fill] = pred[:, 1].astype(np.float32)
    return out


# scipy.interpolate.RBFInterpolator's kernels, default polynomial degrees and
# scale-invariant kernels (scipy 1.15 _rbfinterp.py:20-48; its kernel
# functions: -r, r^2 log r, r^3, -r^5, -sqrt(r^2 + 1), 1/sqrt(r^2 + 1),
# 1/(r^2 + 1), exp(-r^2)).
def _rbf_phi(kernel, r):
    r = np.asarray(r, np.float64)
    if kernel == "linear":
        return -r
    if kernel == "thin_plate_spline":
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(r == 0.0, 0.0, r * r * np.log(np.where(r == 0.0, 1.0, r)))
    if kernel == "cubic":
        return r ** 3
    if kernel == "quintic":
        return -r ** 5
    if kernel == "multiquadric":
        return -np.sqrt(r * r + 1.0)
    if kernel == "inverse_multiquadric":
        return 1.0 / np.sqrt(r * r + 1.0)
    if kernel == "inverse_quadratic":
        return 1.0 / (r * r + 1.0)
    if kernel == "gaussian":
        return np.exp(-r * r)
    raise ValueError(kernel)


_RBF_MIN_DEGREE = {"multiquadric": 0, "linear": 0, "thin_plate_spline": 1, "cubic": 1, "quintic": 2}
_RBF_SCALE_INVARIANT = {"linear", "thin_plate_spline", "cubic", "quintic"}


def _rbf_powers(degree):
    """Monomial exponents (x, y) in scipy's order (_monomial_powers)."""
    from itertools import combinations_with_replacement

    out = []
    for deg in range(degree + 1):
        for mono in combinations_with_replacement(range(2), deg):
            pw = [0, 0]
            for var in mono:
                pw[var] += 1
            out.append(pw)
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def infill_rbf_local(u, spacing, neighbors=64, kernel="thin_plate_spline", epsilon=None, smoothing=0.0,
                     exclude=None, impl="brute"):
    """infill_rbf_local (infilling.py:505-597) per image: NaN vectors (not
    excluded) become the value of a local RBF interpolant of the valid
    vectors in physical coordinates (x * px, y * py), one interpolant per
    component, stored as float32.  impl="scipy" calls RBFInterpolator with
    `neighbors` as the reference does.  impl="brute" restates it: the k =
    min(neighbors, donors) nearest donors ranked by (distance, row-major
    index) -- the rule the device follows where several share the k-th
    distance -- taken in index order; the system [K(eps r) + s I, P; P^T, 0]
    over monomials of (y - shift) / scale (shift, scale: mid-point and
    half-range of the neighbourhood, a zero half-range taken as 1); solved by
    LAPACK's gesv (numpy.linalg.solve, both components at once); evaluated at
    the hole."""
    if kernel not in ("linear", "thin_plate_spline", "cubic", "quintic", "multiquadric",
                      "inverse_multiquadric", "inverse_quadratic", "gaussian"):
        raise ValueError(kernel)
    if epsilon is None:
        if kernel not in _RBF_SCALE_INVARIANT:
            raise ValueError("`epsilon` must be specified")
        epsilon = 1.0
    eps = float(epsilon)
    powers = _rbf_powers(max(_RBF_MIN_DEGREE.get(kernel, -1), 0))
    nm = len(powers)
    py, px = float(spacing[0]), float(spacing[1])
    out = u.astype(np.float32).copy()
    ny, nx = u.shape[2:]
    yy, xx = np.mgrid[0:ny, 0:nx]
    X = np.c_[(xx * px).ravel(), (yy * py).ravel()].astype(np.float64)
    for i in range(u.shape[1]):
        ux, uy = u[1, i].ravel(), u[0, i].ravel()
        ex = np.zeros(ny * nx, bool) if exclude is None else np.asarray(exclude, bool).ravel()
        fill = (np.isnan(ux) | np.isnan(uy)) & ~ex
        if not fill.any():
            continue
        donor = ~fill & ~ex & np.isfinite(ux) & np.isfinite(uy)
        nd = int(donor.sum())
        if nd < 4:
            raise ValueError(f"Infilling method 'rbf_local' cannot run: {nd} valid donor point(s)")
        Xd, Xq = X[donor], X[fill]
        du, dv = ux[donor].astype(np.float64), uy[donor].astype(np.float64)
        if impl == "scipy":
            from scipy.interpolate import RBFInterpolator

            pu = RBFInterpolator(Xd, du, kernel=kernel, epsilon=epsilon, neighbors=neighbors,
                                 smoothing=smoothing)(Xq)
            pv = RBFInterpolator(Xd, dv, kernel=kernel, epsilon=epsilon, neighbors=neighbors,
                                 smoothing=smoothing)(Xq)
        else:
            k = min(int(neighbors), nd)
            if nm > k:
                raise ValueError("fewer neighbours than monomials")
            idx = np.flatnonzero(donor)
            pu, pv = np.empty(len(Xq)), np.empty(len(Xq))
            for q, xq in enumerate(Xq):
                d = np.sqrt(((Xd - xq) ** 2).sum(axis=1))
                nbr = np.sort(np.lexsort((idx, d))[:k])
                y = Xd[nbr]
                mins, maxs = y.min(axis=0), y.max(axis=0)
                shift, scale = (maxs + mins) / 2, (maxs - mins) / 2
                scale[scale == 0.0] = 1.0
                ye = y * eps
                diff = ye[:, None, :] - ye[None, :, :]
                lhs = np.zeros((k + nm, k + nm))
                lhs[:k, :k] = _rbf_phi(kernel, np.sqrt((diff ** 2).sum(axis=-1)))
                lhs[:k, :k] += np.eye(k) * float(smoothing)
                pm = np.prod(((y - shift) / scale)[:, None, :] ** powers[None, :, :], axis=-1)
                lhs[:k, k:] = pm
                lhs[k:, :k] = pm.T
                rhs = np.zeros((k + nm, 2))
                rhs[:k, 0], rhs[:k, 1] = du[nbr], dv[nbr]
                coef = np.linalg.solve(lhs, rhs)
                de = xq * eps - ye
                vec = np.r_[_rbf_phi(kernel, np.sqrt((de ** 2).sum(axis=1))),
                            np.prod(((xq - shift) / scale)[None, :] ** powers, axis=-1)]
                pu[q], pv[q] = vec @ coef[:, 0], vec @ coef[:, 1]
        out[1, i].ravel()[fill] = np.asarray(pu, np.float64).astype(np.float32)
        out[0, i].ravel()[fill] = np.asarray(pv, np.float64).astype(np.float32)
    return out


def vector_masks_loop(image_shape, window_sizes, overlaps, pixel_mask, threshold=0.5):
    """compute_vector_mask (load_images.py:827-1000, standard mode) as the
    reference's per-window loop."""
    from pivtools_amd.window import compute_window_centers

    H, W = image_shape
    out = []
    for win, ov in zip(window_sizes, overlaps):
        wy, wx = win
        wc = compute_window_centers((H, W), (wy, wx), ov)
        m = np.zeros((wc.n_win_y, wc.n_win_x), bool)
        hy, hx = (wy - 1) / 2.0, (wx - 1) / 2.0
        for iy in range(wc.n_win_y):
            y0 = int(np.floor(wc.win_ctrs_y[iy] - hy + 0.5))
            y1 = int(np.floor(wc.win_ctrs_y[iy] + hy + 0.5))
            for ix in range(wc.n_win_x):
                x0 = int(np.floor(wc.win_ctrs_x[ix] - hx + 0.5))
                x1 = int(np.floor(wc.win_ctrs_x[ix] + hx + 0.5))
                ya, yb, xa, xb = max(0, y0), min(H, y1 + 1), max(0, x0), min(W, x1 + 1)
                if yb > ya and xb > xa:
                    m[iy, ix] = np.sum(pixel_mask[ya:yb, xa:xb]) / (wy * wx) > threshold
        out.append(m)
    return out


def _difference_rows(shape, spacing, order=1):
    """Differences of ``order`` along y and x in physical units: (-1, 1)/h
    or (1, -2, 1)/h^2 at every stencil position inside the grid
    (infilling.py:144-179)."""
    import scipy.sparse as sp

    stencil = (-1.0, 1.0) if order == 1 else (1.0, -2.0, 1.0)
    idx = np.arange(shape[0] * shape[1]).reshape(shape)
    rows, cols, vals = [], [], []
    r = 0
    for axis, h in ((0, spacing[0]), (1, spacing[1])):
        n = shape[axis]
        if n <= order:
            continue
        m = None
        rr = None
        for j, coef in enumerate(stencil):
            sl = [slice(None), slice(None)]
            sl[axis] = slice(j, n - order + j)
            c = idx[tuple(sl)].ravel()
            if rr is None:
                m = c.size
                rr = r + np.arange(m)
            rows.append(rr)
            cols.append(c)
            vals.append(np.full(m, coef / h ** order))
        r += m
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(r, shape[0] * shape[1]))


def fill_least_squares(fields, fill, spacing, order=1, exclude=None):
    """Least-squares difference fill of a stack (K, ny, nx): equations that
    touch a fill node and no excluded node, other nodes fixed, normal
    equations by sparse LU in FP64 (infilling.py:182-260)."""
    from scipy.sparse.linalg import splu

    K, ny, nx = fields.shape
    f = np.asarray(fill, bool).ravel()
    e = np.zeros_like(f) if exclude is None else np.asarray(exclude, bool).ravel()
    k = ~f & ~e
    D = _difference_rows((ny, nx), spacing, order)
    Du = D[:, f].tocsr()
    touches_fill = np.diff(Du.indptr) > 0
    touches_excl = np.diff(D[:, e].tocsr().indptr) > 0
    keep = touches_fill & ~touches_excl
    Du, Dk = Du[keep], D[:, k].tocsr()[keep]
    A = (Du.T @ Du).tocsc()
    Uk = fields.reshape(K, -1)[:, k].T.astype(np.float64)
    Bm = Du.T @ (-(Dk @ Uk))
    X = splu(A).solve(np.asarray(Bm)).reshape(-1, K)
    out = fields.astype(np.float64).reshape(K, -1).copy()
    out[:, f] = X.T
    return out.reshape(K, ny, nx)


def fill_harmonic(fields, fill, spacing):
    return fill_least_squares(fields, fill, spacing, 1)


def infill_ls(u, spacing, order, exclude=None):
    """infill_harmonic / infill_biharmonic per image (infilling.py:336-410):
    NaN vectors filled, excluded cells (and stray non-finite ones) neither
    donors nor filled; the result is stored back as float32."""
    out = u.astype(np.float32).copy()
    for i in range(u.shape[1]):
        ux, uy = u[1, i], u[0, i]
        fill = np.isnan(ux) | np.isnan(uy)
        if not fill.any():
            continue
        excl = ~fill & ~(np.isfinite(ux) & np.isfinite(uy))
        if exclude is not None:
            excl |= exclude
        st = fill_least_squares(np.stack([uy, ux]).astype(np.float64), fill, spacing, order, excl)
        out[0, i] = st[0].astype(np.float32)
        out[1, i] = st[1].astype(np.float32)
    return out


def fill_masked_predictor(u, mask, spacing, method):
    """infilling.py:808-870 on the planar field (2, N, ny, nx)."""
    if method == "none" or mask is None or not mask.any():
        return u
    C, N, ny, nx = u.shape
    stack = u.reshape(C * N, ny, nx).astype(np.float64)
    if method == "harmonic":
        out = fill_harmonic(stack, mask, spacing)
    else:
        donor = ~mask & np.all(np.isfinite(stack), axis=0)
        iy, ix = ndi.distance_transform_edt(~donor, sampling=spacing, return_distances=False, return_indices=True)
        out = stack.copy()
        out[:, mask] = stack[:, iy, ix][:, mask]
    return out.reshape(C, N, ny, nx).astype(np.float32)


def apply_predictor_bcs(pad, ctrs_y_all, H, bcs):
    """base.py:364-416: overwrite the padded rows beyond each BC position."""
    for bc in bcs or []:
        if bc["edge"] == "bottom":
            rows = np.asarray(ctrs_y_all, np.float64) >= float(H - 1 - bc["y_position"])
        else:
            rows = np.asarray(ctrs_y_all, np.float64) <= float(bc["y_position"])
        pad[0, :, rows, :] = bc["uy"]
        pad[1, :, rows, :] = bc["ux"]
    return pad


def pad_edge(u, pre, post):
    return np.pad(u, ((0, 0), (0, 0), (pre[0], post[0]), (pre[1], post[1])), mode="edge")


def smooth(field, sd, ksize):
    """gaussian_filter per component and image (cpu_instantaneous.py:1135-1147)."""
    trunc = (ksize[0] - 1) / (2 * sd)
    out = np.empty_like(field)
    for c in range(field.shape[0]):
        for i in range(field.shape[1]):
            out[c, i] = ndi.gaussian_filter(field[c, i], sigma=sd, truncate=trunc, mode="nearest")
    return out


def _cubic(x):
    A = np.float32(-0.75)
    x = np.float32(x)
    one = np.float32(1)
    w0 = ((A * (x + one) - 5 * A) * (x + one) + 8 * A) * (x + one) - 4 * A
    w1 = ((A + 2) * x - (A + 3)) * x * x + one
    w2 = ((A + 2) * (one - x) - (A + 3)) * (one - x) * (one - x) + one
    w3 = one - w0 - w1 - w2
    return np.array([w0, w1, w2, w3], np.float32)


def remap_cubic(field, map_x, map_y):
    """Bicubic remap with replicate border, field (2, N, nyo, nxo) -> (2, N,
    ny, nx): cv2.remap(INTER_CUBIC, BORDER_REPLICATE) as published
    (cpu_instantaneous.py:1050-1071 calls it on float32 maps).  The map is
    rounded to the 1/32 lattice (X = rint(32 x), ties to even), the stencil
    starts at (X >> 5) - 1, and the per-axis Keys a = -0.75 weights of the
    fraction (X & 31) / 32 form the 4 x 4 products wy[a] * wx[b].  A stencil
    inside the field (0 <= sx <= nxo - 4 and 0 <= sy <= nyo - 4) is summed row
    by row (four left-to-right row sums added in order); one that reaches a
    border reads clamped indices and adds its 16 products one at a time in
    row-major order."""
    C, N, nyo, nxo = field.shape
    ny, nx = map_x.shape
    X = np.rint(map_x.astype(np.float32) * np.float32(32)).astype(np.int64)
    Y = np.rint(map_y.astype(np.float32) * np.float32(32)).astype(np.int64)
    out = np.zeros((C, N, ny, nx), np.float32)
    for r in range(ny):
        for c in range(nx):
            sx, sy = (X[r, c] >> 5) - 1, (Y[r, c] >> 5) - 1
            wx = _cubic(np.float32(X[r, c] & 31) * np.float32(1 / 32))
            wy = _cubic(np.float32(Y[r, c] & 31) * np.float32(1 / 32))
            rows = np.clip(sy + np.arange(4), 0, nyo - 1)
            cols = np.clip(sx + np.arange(4), 0, nxo - 1)
            w2 = (wy[:, None] * wx[None, :]).astype(np.float32)
            prod = field[:, :, rows][:, :, :, cols] * w2          # (C, N, 4, 4) float32 products
            if 0 <= sx <= nxo - 4 and 0 <= sy <= nyo - 4:
                rs = ((prod[..., 0] + prod[..., 1]) + prod[..., 2]) + prod[..., 3]
                acc = ((rs[..., 0] + rs[..., 1]) + rs[..., 2]) + rs[..., 3]
            else:
                acc = np.zeros((C, N), np.float32)
                for a in range(4):
                    for b in range(4):
                        acc = acc + prod[..., a, b]
            out[:, :, r, c] = acc
    return out


def peaks_fp64(planes, pw):
    """Peak search + 3-point fit of FP64 planes (n, h, w), vectorised:
    the quarter-region 3x3 local-max rule of peak_locate_lm.c:796-813 and
    threept_estimate (:35-66) on the 5x5 patch times the plane weight."""
    n, h, w = planes.shape
    P = planes
    i_lo, i_hi, j_lo, j_hi = h // 4, h * 3 // 4, w // 4, w * 3 // 4
    core = P[:, i_lo:i_hi, j_lo:j_hi]
    ismax = np.ones_like(core, bool)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            nb = P[:, i_lo + di:i_hi + di, j_lo + dj:j_hi + dj]
            ismax &= ~(nb > core)
    cand = np.where(ismax & (core > 0), core, -np.inf)
    flat = cand.reshape(n, -1)
    k = np.argmax(flat, axis=1)                      # first maximum in row-major order
    h0 = flat[np.arange(n), k]
    ok = np.isfinite(h0)
    cw = j_hi - j_lo
    i0 = i_lo + k // cw
    j0 = j_lo + k % cw
    rows = np.clip(i0[:, None] + np.arange(-1, 2)[None], 0, h - 1)
    cols = np.clip(j0[:, None] + np.arange(-1, 2)[None], 0, w - 1)
    xv = P[np.arange(n)[:, None], rows, j0[:, None]] * pw[rows, j0[:, None]]
    yv = P[np.arange(n)[:, None], i0[:, None], cols] * pw[i0[:, None], cols]
    eps = np.finfo(np.float32).eps
    lx = np.log(np.maximum(xv, eps))
    ly = np.log(np.maximum(yv, eps))
    dx = 2 * lx[:, 0] - 4 * lx[:, 1] + 2 * lx[:, 2]
    dy = 2 * ly[:, 0] - 4 * ly[:, 1] + 2 * ly[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        oi = np.where(np.abs(dx) > eps, (lx[:, 0] - lx[:, 2]) / dx, 0.0)
        oj = np.where(np.abs(dy) > eps, (ly[:, 0] - ly[:, 2]) / dy, 0.0)
    row = np.where(ok, i0 + oi, np.nan)
    col = np.where(ok, j0 + oj, np.nan)
    return row, col, np.where(ok, h0, 0.0)


def xcorr_pass(A, B, g, w, finder, flavour, mask=None, n_peaks=1):
    """One bulkxcorr2d over all pairs: flavour 'oracle' (float32 C
    restatement) or 'fp64' (FP64 planes, FP64 3-point fit)."""
    ny, nx = g.n_win
    win = g.win
    if flavour == "oracle":
        fm = np.zeros((ny, nx), np.float32) if mask is None else mask.astype(np.float32)
        o = O.bulkxcorr2d(A, B, fm, g.ctrs_x, g.ctrs_y, win, w, w, n_peaks, finder)
        if n_peaks > 1:
            return o["pk_x"], o["pk_y"], o["pk_h"]
        return o["pk_x"][:, 0], o["pk_y"][:, 0], o["pk_h"][:, 0]
    if finder != 3:
        raise ValueError("the FP64 flavour fits with the 3-point estimator")
    if mask is not None:
        raise ValueError("the FP64 flavour runs without window masks")
    pw = O.plane_weight(w).astype(np.float64)
    N = A.shape[0]
    pkx = np.zeros((N, ny, nx))
    pky = np.zeros((N, ny, nx))
    pkh = np.zeros((N, ny, nx))
    for i in range(N):
        a, b, ok = O.windows_fp64(A[i], B[i], g.ctrs_x, g.ctrs_y, win, w.astype(np.float64), w.astype(np.float64))
        fa, fb = np.fft.fft2(a), np.fft.fft2(b)
        planes = np.fft.fftshift(np.real(np.fft.ifft2(fb * np.conj(fa))), axes=(-2, -1)) / 1.0
        row, col, h = peaks_fp64(planes, pw)
        en = 1.0 / np.sqrt(np.sum(a * a, axis=(1, 2)) * np.sum(b * b, axis=(1, 2)))
        pkx[i] = (col - win[1] / 2.0).reshape(ny, nx)
        pky[i] = (row - win[0] / 2.0).reshape(ny, nx)
        pkh[i] = (h * en).reshape(ny, nx)
    return pkx, pky, pkh


def run_multipass(A, B, grids, window_type="gaussian", finder=3, smoothing=True, median=(0.1, 2.0, 5),
                  final_infill="nearest", flavour="oracle", interp=0, keep_passes=False, vector_masks=None,
                  predictor_fill="harmonic", predictor_bcs=None, mid_infill="nearest", outlier_methods=None,
                  num_peaks=1, secondary_peak=False):
    """The whole pass loop on the CPU; A, B (N, H, W) float32."""
    from pivtools_amd.window import window_weight

    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    prev_pad = None
    prev_grid = None
    out = []
    for p, g in enumerate(grids):
        w = window_weight(g.win, window_type)
        pred = None
        if p == 0:
            a_in, b_in = A, B
        else:
            field = prev_pad
            if smoothing and g.sd > 0:
                field = smooth(field, g.sd, g.ksize)
            pred = remap_cubic(field, g.pred_map_x, g.pred_map_y)
            a_in, b_in = O.fused_warp_batch(A, B, field[0], field[1], prev_grid.ctrs_y_all, prev_grid.ctrs_x_all,
                                            interp, False, False)
        m = None
        if vector_masks is not None and vector_masks[p] is not None and np.any(vector_masks[p]):
            m = np.asarray(vector_masks[p], bool)
        pkx, pky, pkh = xcorr_pass(a_in, b_in, g, w, finder, flavour, m, num_peaks)
        div = 2.0 if p == 0 else 4.0
        choice = None
        if num_peaks > 1:
            methods = outlier_methods
            if methods is None:
                methods = [] if median is None else [{"type": "median_2d", "epsilon": median[0],
                                                      "threshold": median[1], "size": median[2]}]
            u, nm, reason, pko, choice, _ = peaks_stage(pkx, pky, pkh, pred, m, g.win[1] / div, g.win[0] / div,
                                                        methods, secondary_peak)
        else:
            u, nm, reason, pkg = gate(pkx, pky, pkh, pred, m, g.win[1] / div, g.win[0] / div)
            if outlier_methods is None:
                u, nm, reason, pko = median_stage(u, pkg, nm, reason, m, *median)
            else:
                u, nm, reason, pko = outlier_stage(u, pkg, nm, reason, m, outlier_methods)
        final = p == len(grids) - 1
        sp_ = (float(g.spacing[0]), float(g.spacing[1]))
        method = mid_infill if not final else final_infill
        if method == "nearest":
            u = infill_nearest(u, sp_, exclude=m)
        elif method in ("harmonic", "biharmonic"):
            u = infill_ls(u, sp_, 1 if method == "harmonic" else 2, exclude=m)
        elif method == "knn":
            u = infill_knn(u, sp_, exclude=m)
        elif method == "rbf_local":
            u = infill_rbf_local(u, sp_, exclude=m)
        elif method == "griddata_linear" or (isinstance(method, dict) and method.get("method") == "griddata_linear"):
            u = infill_griddata_linear(u, sp_, exclude=m)
        elif isinstance(method, dict) and method.get("method") == "rbf_local":
            u = infill_rbf_local(u, sp_, exclude=m, **method.get("parameters", {}))
        elif isinstance(method, dict) and method.get("method") == "knn":
            kp = method.get("parameters", {})
            u = infill_knn(u, sp_, k=kp.get("n_neighbors", 32), weights=kp.get("weights", "distance"),
                           exclude=m)
        if not final:
            src = fill_masked_predictor(u, m, sp_, predictor_fill)
            prev_pad = apply_predictor_bcs(pad_edge(src, g.n_pre, g.n_post), g.ctrs_y_all, A.shape[1],
                                           predictor_bcs)
        if keep_passes or final:
            out.append({"ux": u[1], "uy": u[0], "nan_mask": nm, "nan_reason": reason, "peak_mag": pko,
                        "predictor": pred, "peak_choice": choice})
        prev_grid = g
    return out


def _griddata_scales(spacing):
    """Axis scales for the triangulation: the pitches divided by their common
    integer factor when both are integers (exact predicates), else the
    pitches.  Scaling both axes alike changes neither the Delaunay triangles
    nor the barycentric weights."""
    py, px = float(spacing[0]), float(spacing[1])
    if py == int(py) and px == int(px) and py > 0 and px > 0:
        import math
        g = math.gcd(int(py), int(px))
        return py / g, px / g
    return py, px


def _triples(m):
    """All index triples a < b < c of range(m), in lexicographic order."""
    import itertools
    return np.array(list(itertools.combinations(range(m), 3)), dtype=np.int64).reshape(-1, 3)


def griddata_triangle(X, Y, covered_radius=None, chunk=8192):
    """The triangle the griddata restatement uses for a hole at the origin,
    among candidates (X, Y) in row-major order: the first (a, b, c) in
    lexicographic order with nonzero area, the origin inside or on an edge,
    no candidate strictly inside its circumcircle and, when covered_radius
    is given, its circumcircle strictly within that distance of the origin.
    Returns (a, b, c, area, oa, ob, oc), "uncovered" when the only triangles
    found fail the last test, or None when no triangle holds the origin.
    Triples are examined in lexicographic chunks, stopping at the first hit."""
    m = X.size
    if m < 3:
        return None
    t_all = _triples(m)
    seen_empty = False
    for s0 in range(0, t_all.shape[0], chunk):
        t = t_all[s0:s0 + chunk]
        a, b, c = t[:, 0], t[:, 1], t[:, 2]
        ar = (X[b] - X[a]) * (Y[c] - Y[a]) - (Y[b] - Y[a]) * (X[c] - X[a])
        oa = X[b] * Y[c] - Y[b] * X[c]
        ob = X[c] * Y[a] - Y[c] * X[a]
        oc = X[a] * Y[b] - Y[a] * X[b]
        pos = (ar > 0) & (oa >= 0) & (ob >= 0) & (oc >= 0)
        neg = (ar < 0) & (oa <= 0) & (ob <= 0) & (oc <= 0)
        hold = np.flatnonzero(pos | neg)
        if hold.size == 0:
            continue
        a, b, c, arh = a[hold], b[hold], c[hold], ar[hold]
        adx, ady = X[a][:, None] - X[None, :], Y[a][:, None] - Y[None, :]
        bdx, bdy = X[b][:, None] - X[None, :], Y[b][:, None] - Y[None, :]
        cdx, cdy = X[c][:, None] - X[None, :], Y[c][:, None] - Y[None, :]
        det = ((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy)
               + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy)
               + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))
        det = det * np.sign(arh)[:, None]
        empty = ~(det > 0).any(axis=1)      # a vertex gives det = 0 exactly
        ok = empty.copy()
        if covered_radius is not None:
            bx, by = X[b] - X[a], Y[b] - Y[a]
            cx, cy = X[c] - X[a], Y[c] - Y[a]
            dd = 2.0 * (bx * cy - by * cx)
            b2, c2 = bx * bx + by * by, cx * cx + cy * cy
            ox = (cy * b2 - by * c2) / dd
            oy = (bx * c2 - cx * b2) / dd
            rho = np.sqrt(ox * ox + oy * oy)
            gx, gy = ox + X[a], oy + Y[a]
            ok &= np.sqrt(gx * gx + gy * gy) + rho < covered_radius
        seen_empty |= bool(empty.any())
        idx = np.flatnonzero(ok)
        if idx.size:
            j = hold[idx[0]]
            return (t[j, 0], t[j, 1], t[j, 2], ar[j], oa[j], ob[j], oc[j])
    return "uncovered" if seen_empty else None


GRIDDATA_KS = (24, 48, 96, 192)


def infill_griddata_linear(u, spacing, exclude=None, ks=GRIDDATA_KS):
    """infill_griddata_linear (infilling.py:440-503) per image: NaN vectors
    (not excluded) become the linear interpolation over a Delaunay
    triangulation of the valid vectors in physical coordinates, and cells
    outside the convex hull take the nearest valid vector.

    The restatement, which the device follows step for step, for k = 24,
    48, 96, 192 in turn until a triangle is found:
      - candidates: the k nearest donors of the hole, ranked by (distance,
        row-major index), then taken in row-major order;
      - the triangle: griddata_triangle -- the first (a, b, c) holding the
        hole whose circumcircle is empty of candidates and, unless every
        donor is a candidate, lies strictly within the k-th candidate's
        distance, so that it is empty of all donors and globally Delaunay;
      - value: barycentric weights (ratios of orientation determinants, FP64)
        applied to the vertices' components, stored as float32;
      - no triangle holds the hole: the nearest candidate's value (the first
        in (distance, index) order), as for cells outside the convex hull;
      - triangles hold it but none is covered at k = 192: ValueError.
    On a regular grid four donors around a hole are often cocircular: scipy's
    Qhull then picks one of the equally Delaunay diagonals by its processing
    order.  This rule picks the first in index order; both are Delaunay
    interpolants, and they agree wherever the triangle is unique and on
    linear fields."""
    sy_, sx_ = _griddata_scales(spacing)
    out = u.astype(np.float32).copy()
    ny, nx = u.shape[2:]
    rr, cc = np.mgrid[0:ny, 0:nx]
    rr, cc = rr.ravel(), cc.ravel()
    for i in range(u.shape[1]):
        fy_all, fx_all = u[0, i].ravel(), u[1, i].ravel()
        ex = np.zeros(ny * nx, bool) if exclude is None else np.asarray(exclude, bool).ravel()
        fill = (np.isnan(fx_all) | np.isnan(fy_all)) & ~ex
        if not fill.any():
            continue
        donor = ~fill & ~ex & np.isfinite(fx_all) & np.isfinite(fy_all)
        didx = np.flatnonzero(donor)
        if didx.size == 0:
            raise ValueError("griddata_linear: no donor")
        for q in np.flatnonzero(fill):
            ry = (rr[didx] - rr[q]) * sy_
            rx = (cc[didx] - cc[q]) * sx_
            d = np.sqrt(ry * ry + rx * rx)
            order = np.lexsort((didx, d))
            for k in ks:
                near = order[:k]
                cand = np.sort(near)               # row-major order
                tri = griddata_triangle(rx[cand], ry[cand], None if didx.size <= k else d[near[-1]])
                if not isinstance(tri, str):
                    break
            if isinstance(tri, str):
                raise ValueError(f"griddata_linear: a hole's Delaunay triangle reaches beyond its {ks[-1]} "
                                 "nearest donors")
            if tri is None:
                j = didx[near[0]]
                val = (float(fy_all[j]), float(fx_all[j]))
            else:
                a, b, c, ar, oa, ob, oc = tri
                la, lb, lc = oa / ar, ob / ar, oc / ar
                ia, ib, ic = didx[cand[a]], didx[cand[b]], didx[cand[c]]
                val = (la * float(fy_all[ia]) + lb * float(fy_all[ib]) + lc * float(fy_all[ic]),
                       la * float(fx_all[ia]) + lb * float(fx_all[ib]) + lc * float(fx_all[ic]))
            out[0, i].ravel()[q] = np.float32(val[0])
            out[1, i].ravel()[q] = np.float32(val[1])
    return out
