"""Window grid, window weights and the multipass grid caches.

Host-side restatements of the reference's grid helpers; they produce the
arrays that cross the C ABI (window centres, weights, padded predictor grid)
so the numbers handed to the kernels are the reference's numbers:

* ``compute_window_centers``  pivtools_core/window_utils.py:142-273
* ``compute_padding_for_single_mode`` / ``compute_window_centers_single_mode``
                              pivtools_core/window_utils.py:79-139, 276-444
                              (ensemble 'single' run type)
* ``window_weight``           pivtools_cli/piv/piv_backend/base.py:52-135
* ``vector_masks``            pivtools_core/image_handling/load_images.py:827-1000
                              (compute_vector_mask, standard mode)
* ``PassGrid`` / ``build_pass_grids``
                              pivtools_cli/piv/piv_backend/base.py:137-300
                              (instantaneous: first-pass kernel (0, 0), no BCs)
                              and base.py:410-483 (predictor interpolation maps)
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

BUILT_FFT_SIZES = (8, 12, 16, 24, 32, 48, 64, 96, 128)   # pivtools_core/fft_sizes.py


@dataclass
class WindowCenters:
    win_ctrs_x: np.ndarray
    win_ctrs_y: np.ndarray
    n_win_x: int
    n_win_y: int
    win_spacing_x: int
    win_spacing_y: int
    padding: tuple = (0, 0, 0, 0)    # (top, bottom, left, right) of the single-mode zero border


def compute_window_centers(image_shape, window_size, overlap, validate=True) -> WindowCenters:
    """Window centres: first at (w-1)/2, last at W-(w+1)/2, y anchored to the bottom."""
    H, W = image_shape
    wh, ww = window_size
    if validate:
        if wh > H or ww > W:
            raise ValueError(f"Window size {window_size} exceeds image dimensions {image_shape}")
        if overlap < 0 or overlap >= 100:
            raise ValueError(f"Overlap must be in range [0, 100), got {overlap}")
        if wh < 1 or ww < 1:
            raise ValueError(f"Window size must be positive, got {window_size}")
    sx = max(1, round((1 - overlap / 100) * ww))
    sy = max(1, round((1 - overlap / 100) * wh))
    fx, lx = (ww - 1) / 2.0, W - (ww + 1) / 2.0
    fy, ly = (wh - 1) / 2.0, H - (wh + 1) / 2.0
    nx = max(1, int(np.floor((lx - fx) / sx)) + 1)
    ny = max(1, int(np.floor((ly - fy) / sy)) + 1)
    cx = np.linspace(fx, fx + sx * (nx - 1), nx, dtype=np.float32)
    cy = np.linspace(ly - sy * (ny - 1), ly, ny, dtype=np.float32)
    return WindowCenters(cx, cy, nx, ny, sx, sy)


def compute_padding_for_single_mode(window_size, sum_window) -> tuple:
    """Zero border (top, bottom, left, right) that lets a sum window sit
    around every small-window position: ceil / floor halves of sum - window."""
    (wh, ww), (sh, sw) = window_size, sum_window
    if sh < wh or sw < ww:
        raise ValueError(f"SumWindow {tuple(sum_window)} must be >= window_size {tuple(window_size)} for single mode")
    dy, dx = sh - wh, sw - ww
    return (dy - dy // 2, dy // 2, dx - dx // 2, dx // 2)


def compute_window_centers_single_mode(image_shape, window_size, sum_window, overlap,
                                       validate=True) -> WindowCenters:
    """Single-mode grid in PADDED image coordinates: spacing from the small
    window, first centre at pad + (w-1)/2, last at pad + size - (w+1)/2, y
    anchored to the bottom; every sum window must fit the padded image."""
    H, W = image_shape
    wh, ww = window_size
    sh, sw = sum_window
    if validate:
        if sh < wh or sw < ww:
            raise ValueError(f"SumWindow {tuple(sum_window)} must be >= window_size {tuple(window_size)}")
        if wh > H or ww > W:
            raise ValueError(f"Window size {window_size} exceeds image dimensions {image_shape}")
        if overlap < 0 or overlap >= 100:
            raise ValueError(f"Overlap must be in range [0, 100), got {overlap}")
    pt, pb, pl, pr = compute_padding_for_single_mode(window_size, sum_window)
    sx = max(1, round((1 - overlap / 100) * ww))
    sy = max(1, round((1 - overlap / 100) * wh))
    fx, lx = pl + (ww - 1) / 2.0, pl + W - (ww + 1) / 2.0
    fy, ly = pt + (wh - 1) / 2.0, pt + H - (wh + 1) / 2.0
    nx = max(1, int(np.floor((lx - fx) / sx)) + 1)
    ny = max(1, int(np.floor((ly - fy) / sy)) + 1)
    cx = np.linspace(fx, fx + sx * (nx - 1), nx, dtype=np.float32)
    cy = np.linspace(ly - sy * (ny - 1), ly, ny, dtype=np.float32)
    if cx[-1] + (sw - 1) / 2.0 >= W + pl + pr or cy[-1] + (sh - 1) / 2.0 >= H + pt + pb:
        raise ValueError(f"single-mode window centres exceed the padded image (sum window {tuple(sum_window)})")
    return WindowCenters(cx, cy, nx, ny, sx, sy, (pt, pb, pl, pr))


def window_weight(wsize, window_type: str, sum_window=None) -> np.ndarray:
    """2D window weight (float32), the reference's _window_weight_fun: the
    types an instantaneous or standard ensemble pass uses, and the single-mode
    pair ('singlepix': ones on the wsize block at ceil((sum - wsize)/2) of a
    sum-window array; 'bsingle': ones over the sum window)."""
    m, n = wsize
    t = window_type.lower()
    if t in ("singlepix", "bsingle"):
        if sum_window is None:
            raise ValueError(f"SumWindow must be provided for {window_type!r}")
        sh, sw = int(sum_window[0]), int(sum_window[1])
        if t == "bsingle":
            return np.ones((sh, sw), dtype=np.float32)
        w = np.zeros((sh, sw), dtype=np.float32)
        r0, c0 = -((m - sh) // 2), -((n - sw) // 2)     # ceil((sum - size) / 2)
        w[r0:r0 + m, c0:c0 + n] = 1.0
        return w
    if t == "square":
        return np.ones((m, n), dtype=np.float32)
    if t in ("gaussian", "gaussiantop"):
        wxv = np.arange(-m / 2 + 0.5, m / 2, dtype=np.float32)
        wyv = np.arange(-n / 2 + 0.5, n / 2, dtype=np.float32)
        xm, ym = np.meshgrid(wxv, wyv, indexing="ij")
        w = np.exp(-0.5 * ((1.0 * xm / (m / 2)) ** 2 + (1.0 * ym / (n / 2)) ** 2))
        if t == "gaussiantop":
            if m > 6 and n > 6:
                cm = np.tile(np.array([[1, 0], [0, 0]], dtype=np.float32), (m // 2, n // 2))
                cm = np.roll(np.roll(cm, int(np.ceil(m / 4)), axis=0), int(np.ceil(n / 4)), axis=1)
            else:
                cm = np.zeros((m, n), dtype=np.float32)
                r0, c0 = m // 2 - 1, n // 2 - 1
                cm[r0:r0 + 4, c0:c0 + 4] = 1
            w = w * cm
        return np.ascontiguousarray(w.astype(np.float32))
    if t == "blackman":
        a0, a1, a2 = 0.42659, 0.49656, 0.076849
        dmax = np.sqrt(((m + 1) / 2) ** 2 + ((n + 1) / 2) ** 2)
        w = np.zeros((m, n), dtype=np.float32)
        for ii in range(m):
            for jj in range(n):
                d = np.sqrt((ii - (m - 1) / 2) ** 2 + (jj - (n - 1) / 2) ** 2)
                th = d / dmax * np.pi + np.pi
                w[ii, jj] = a0 - a1 * np.cos(th) + a2 * np.cos(2 * th)
        w[w <= 0] = 0.01
        return w
    raise ValueError(f"Unrecognized window type '{window_type}'")


@dataclass
class PassGrid:
    """Everything one pass of the instantaneous multipass loop needs."""

    win: tuple
    ctrs_x: np.ndarray          # measured centres (n_win_x,)
    ctrs_y: np.ndarray
    spacing: tuple              # (spacing_y, spacing_x)
    ctrs_x_all: np.ndarray      # padded centres (predictor grid)
    ctrs_y_all: np.ndarray
    n_pre: tuple                # (pre_y, pre_x)
    n_post: tuple
    ksize: tuple                # predictor smoothing kernel size
    sd: float                   # predictor smoothing sigma
    pred_map_x: np.ndarray | None = None   # fractional index into previous padded grid
    pred_map_y: np.ndarray | None = None

    @property
    def n_win(self):
        return (len(self.ctrs_y), len(self.ctrs_x))


def _pad_axis(ctrs, spacing, size):
    pre = np.arange(1, ctrs[0] - spacing / 2, spacing)
    if pre.size == 0:
        pre = np.array([1])
    pre = pre - 1
    while len(pre) < 2:
        pre = np.concatenate([[max(0, pre[0] - spacing)], pre])
    post = np.arange(size, ctrs[-1] + spacing / 2, -spacing)
    if post.size == 0:
        post = np.array([size])
    post = post - 1
    while len(post) < 2:
        post = np.concatenate([post, [min(size - 1, post[-1] + spacing)]])
    return np.concatenate([pre, ctrs, post[::-1]]), len(pre), len(post)


def window_centers_original(image_shape, window_size, overlap, run_type="standard",
                            sum_window=None) -> WindowCenters:
    """Window centres in original image coordinates for either run type: a
    single-mode grid has its padding subtracted (cpu_ensemble.py:695-750), so
    every pass shares one coordinate system; ``padding`` is kept for the
    correlation call, which adds it back."""
    if run_type == "standard":
        return compute_window_centers(image_shape, window_size, overlap)
    if run_type != "single":
        raise ValueError(f"run type {run_type!r}: standard or single")
    wc = compute_window_centers_single_mode(image_shape, window_size, sum_window, overlap)
    pt, _, pl, _ = wc.padding
    return WindowCenters(wc.win_ctrs_x - np.float32(pl), wc.win_ctrs_y - np.float32(pt), wc.n_win_x, wc.n_win_y,
                         wc.win_spacing_x, wc.win_spacing_y, wc.padding)


def build_pass_grids(image_shape, window_sizes, overlaps, run_types=None, sum_window=None) -> list[PassGrid]:
    H, W = image_shape
    grids: list[PassGrid] = []
    for p, win in enumerate(window_sizes):
        rt = "standard" if run_types is None else run_types[p]
        wc = window_centers_original(image_shape, win, overlaps[p], rt, sum_window)
        xa, prex, postx = _pad_axis(wc.win_ctrs_x, wc.win_spacing_x, W)
        ya, prey, posty = _pad_axis(wc.win_ctrs_y, wc.win_spacing_y, H)
        if p == 0:
            ksize, sd = (0, 0), 0.0
        else:
            pw, ps = window_sizes[p - 1], grids[p - 1].spacing
            k = np.round(np.array(pw) / np.array(ps)).astype(int) + 1
            ksize = (int(k[0]) + (int(k[0]) % 2 == 0), int(k[1]) + (int(k[1]) % 2 == 0))
            sd = float(np.sqrt(np.prod(ksize)) / 3 * 0.65)
        g = PassGrid(tuple(win), wc.win_ctrs_x.astype(np.float32), wc.win_ctrs_y.astype(np.float32),
                     (wc.win_spacing_y, wc.win_spacing_x), xa.astype(np.float32), ya.astype(np.float32),
                     (prey, prex), (posty, postx), ksize, sd)
        if p > 0:
            prev = grids[p - 1]
            wy, wx = np.meshgrid(g.ctrs_y, g.ctrs_x, indexing="ij")
            ix = np.interp(wx.ravel(), prev.ctrs_x_all, np.arange(len(prev.ctrs_x_all)))
            iy = np.interp(wy.ravel(), prev.ctrs_y_all, np.arange(len(prev.ctrs_y_all)))
            g.pred_map_x = ix.reshape(wx.shape).astype(np.float32)
            g.pred_map_y = iy.reshape(wy.shape).astype(np.float32)
        grids.append(g)
    return grids


def vector_masks(image_shape, window_sizes, overlaps, pixel_mask, threshold: float = 0.5) -> list:
    """Masked windows of every pass, (n_win_y, n_win_x) bool each.

    A window covers rows floor(c - (h-1)/2 + 0.5) .. floor(c + (h-1)/2 + 0.5)
    (and the same in x) clipped to the image; it is masked when its count of
    masked pixels divided by h*w exceeds ``threshold``
    (pivtools_core/image_handling/load_images.py:952-985).  The counts come
    from a summed-area table instead of one slice sum per window; they are
    exact integers either way, and the division and comparison are the
    reference's float64 ones.
    """
    H, W = image_shape
    pm = np.asarray(pixel_mask).astype(bool)
    if pm.shape != (H, W):
        raise ValueError(f"pixel mask {pm.shape} does not match the image shape {(H, W)}")
    sat = np.zeros((H + 1, W + 1), np.int64)
    sat[1:, 1:] = np.cumsum(np.cumsum(pm, axis=0, dtype=np.int64), axis=1)
    out = []
    for win, ov in zip(window_sizes, overlaps):
        wy, wx = (win, win) if np.ndim(win) == 0 else win
        wc = compute_window_centers((H, W), (wy, wx), ov)
        cy = np.asarray(wc.win_ctrs_y, np.float64)
        cx = np.asarray(wc.win_ctrs_x, np.float64)
        y0 = np.floor(cy - (wy - 1) / 2.0 + 0.5).astype(np.int64)
        y1 = np.floor(cy + (wy - 1) / 2.0 + 0.5).astype(np.int64)
        x0 = np.floor(cx - (wx - 1) / 2.0 + 0.5).astype(np.int64)
        x1 = np.floor(cx + (wx - 1) / 2.0 + 0.5).astype(np.int64)
        ya, yb = np.maximum(0, y0), np.minimum(H, y1 + 1)
        xa, xb = np.maximum(0, x0), np.minimum(W, x1 + 1)
        cnt = (sat[yb[:, None], xb[None, :]] - sat[ya[:, None], xb[None, :]]
               - sat[yb[:, None], xa[None, :]] + sat[ya[:, None], xa[None, :]])
        frac = cnt.astype(np.float64) / float(wy * wx)
        nonempty = (yb > ya)[:, None] & (xb > xa)[None, :]
        out.append(nonempty & (frac > threshold))
    return out
