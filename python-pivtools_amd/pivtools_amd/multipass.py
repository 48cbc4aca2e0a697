"""Device-resident instantaneous multipass PIV.

The pass loop of ``InstantaneousCorrelatorCPU.correlate_batch``
(pivtools_cli/piv/piv_backend/cpu_instantaneous.py:381-927) with every step
on the device: for pass p > 0 the previous pass's padded field is smoothed
(scipy gaussian_filter, :1135-1147), remapped onto the new window grid as the
predictor (cv2 bicubic remap, :1050-1071) and used to warp both frames
(fused_symmetric_warp_batch, :1112-1131); the frames are correlated
(bulkxcorr2d, :505-528); the result is gated (N/2 on the first pass, N/4
after), the predictor added back, outliers flagged by the normalised median
test, holes infilled by nearest donor and the field edge-padded into the next
predictor (:555-865).

Masked windows (a per-pass window mask, given directly or computed from a
pixel mask as pivtools_core/image_handling/load_images.py:827-1000 does) are
skipped by the correlation, invalid with reason -1, zero in the outputs and
never infill donors; before a field becomes the next predictor its masked
windows are filled from the unmasked ones ('harmonic', 'nearest' or 'none',
infilling.py:808-870), and optional predictor boundary conditions overwrite
padded rows (base.py:364-416).

Torch provides device memory and the stream; every computation is a HIP
kernel of libpivtools_amd.so launched through its C ABI on torch's current
stream.  Buffers are allocated once per (batch size, schedule) and reused.
"""

from __future__ import annotations

import ctypes
from types import SimpleNamespace
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native
from .predictor_fill import harmonic_operator
from .window import PassGrid, build_pass_grids, window_weight
from .window import vector_masks as _vector_masks

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_sz = ctypes.c_size_t


PEAK_FINDERS = {"gauss3": 3, "gauss4": 4, "gauss5": 5, "gauss6": 6}


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def gaussian_weights_1d(sigma: float, truncate: float) -> tuple[np.ndarray, int]:
    """Half kernel w[0..radius] of scipy.ndimage.gaussian_filter1d."""
    radius = int(truncate * float(sigma) + 0.5)
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    phi = phi / phi.sum()
    return np.ascontiguousarray(phi[radius:], dtype=np.float64), radius


@dataclass
class PassResult:
    """Per-pass device outputs, batched over the pairs: the fields of the
    reference's PIVPassResult (pivtools_cli/piv/piv_result.py:7-30, filled at
    cpu_instantaneous.py:876-891).  ``peak_mag`` holds every peak's height,
    NaN where the peak was gated out (N, num_peaks, ny, nx);
    ``primary_peak_mag`` the chosen peak's height, NaN at invalid vectors;
    ``peak_choice`` the 1-based chosen peak, 0 at invalid vectors."""

    ux: torch.Tensor
    uy: torch.Tensor
    nan_mask: torch.Tensor
    nan_reason: torch.Tensor
    peak_mag: torch.Tensor
    peak_ratio: torch.Tensor
    predictor: torch.Tensor | None
    window_size: tuple
    n_windows: tuple
    primary_peak_mag: torch.Tensor | None = None
    peak_choice: torch.Tensor | None = None
    b_mask: np.ndarray | None = None
    win_ctrs_x: np.ndarray | None = None
    win_ctrs_y: np.ndarray | None = None
    # dump_correlation_planes: {"planes" (N, ny, nx, h, w) raw planes, zero
    # at skipped windows; "pk_loc_x_raw", "pk_loc_y_raw", "pk_height_raw",
    # "sx", "sy", "sxy" (N, P, ny, nx) and "pk_ratio_raw" (N, ny, nx) as the
    # correlation returned them; "peak_finder"} (cpu_instantaneous.py:555-572)
    debug_dump: dict | None = None


@dataclass
class _PassBuf:
    grid: PassGrid
    ctrs: dict = field(default_factory=dict)
    w: torch.Tensor | None = None
    pw: torch.Tensor | None = None
    mask: torch.Tensor | None = None
    out: dict = field(default_factory=dict)
    work: torch.Tensor | None = None
    maps: tuple | None = None
    smooth: tuple | None = None
    fill_op: tuple | None = None       # harmonic predictor fill: (fill_idx, bnd_idx, gt) on the device
    bc_rows: list = field(default_factory=list)   # (row slice of the padded grid, uy, ux)


def _infill_cfg(cfg, final):
    """(method, parameters) of an infill option: a name, or the reference's
    config form {"method": ..., "parameters": {...}} (infilling.py:698-790)."""
    if isinstance(cfg, dict):
        method, params = str(cfg.get("method", "nearest")).lower(), dict(cfg.get("parameters", {}) or {})
        if final and not cfg.get("enabled", True):
            method = "none"
    else:
        method, params = str(cfg).lower(), {}
    allowed = ("nearest", "harmonic", "biharmonic", "knn", "rbf_local", "griddata_linear") + \
        (("none",) if final else ())
    if method not in allowed:
        raise ValueError(f"infill method {method!r}: one of {allowed}")
    if method == "griddata_linear" and str(params.get("method", "linear")) != "linear":
        raise ValueError(f"griddata_linear method {params.get('method')!r}: only 'linear' is the reference's")
    if method == "knn":
        k = int(params.get("n_neighbors", 32))
        if not 1 <= k <= 128:
            raise ValueError(f"knn n_neighbors {k}: 1..128 on the device")
        if params.get("weights", "distance") not in ("distance", "uniform"):
            raise ValueError(f"knn weights {params.get('weights')!r}: distance or uniform")
    if method == "rbf_local":
        k = int(params.get("neighbors", 64))
        if not 1 <= k <= 128:
            raise ValueError(f"rbf_local neighbors {k}: 1..128 on the device")
        kern = params.get("kernel", "thin_plate_spline")
        if kern not in RBF_KERNELS:
            raise ValueError(f"`kernel` must be one of {sorted(RBF_KERNELS)}.")
        eps = params.get("epsilon", None)
        if eps is None:
            if kern not in ("linear", "thin_plate_spline", "cubic", "quintic"):
                raise ValueError("`epsilon` must be specified if `kernel` is not one of "
                                 "{'linear', 'thin_plate_spline', 'cubic', 'quintic'}.")
            eps = 1.0
        sm = float(params.get("smoothing", 0.0) or 0.0)
        if not float(eps) > 0.0 or not sm >= 0.0:
            raise ValueError(f"rbf_local epsilon {eps!r} must be > 0 and smoothing {sm!r} >= 0")
        params = {"neighbors": k, "kernel": kern, "epsilon": float(eps), "smoothing": sm}
    return method, params


# scipy RBFInterpolator kernel names -> pivamd_infill_rbf kernel codes
RBF_KERNELS = {"linear": 0, "thin_plate_spline": 1, "cubic": 2, "quintic": 3, "multiquadric": 4,
               "inverse_multiquadric": 5, "inverse_quadratic": 6, "gaussian": 7}


# Correlation transform of the passes (pivamd_set_dft_mode): the FP32 FFT, or a
# split-precision MFMA DFT; sizes a form does not cover keep the FFT.
DFT_MODES = {"fft": 0, "split": 1, "split-bf16": 2, "split-bf16x3": 3}


class MultipassPIV:
    """Multipass instantaneous PIV for batches of pairs of one image shape."""

    def __init__(self, image_shape, window_sizes, overlaps, *, window_type="gaussian", peak_finder="gauss3",
                 peak_fit_impl="scalar", warp_interpolation="cubic", predictor_smoothing=True,
                 median=(0.1, 2.0, 5), final_infill="nearest", dft="fft", device="cuda", mid_infill="nearest",
                 vector_masks=None, pixel_mask=None, mask_threshold=0.5, predictor_fill="harmonic",
                 predictor_bcs=None, outlier_methods=None, num_peaks=1, secondary_peak=False,
                 dump_correlation_planes=False, warp_arithmetic="fma", max_batch=None):
        """``vector_masks``: one (n_win_y, n_win_x) bool array per pass (True =
        masked window), or ``pixel_mask``: an (H, W) bool image mask turned
        into them with ``mask_threshold``.  ``predictor_fill``: 'harmonic',
        'nearest' or 'none' for the masked windows of each predictor field.
        ``predictor_bcs``: list of {"edge": "top"|"bottom", "y_position",
        "ux", "uy"} overwriting the padded predictor rows beyond y_position.
        ``outlier_methods``: the reference's method list (config
        outlier_detection.methods), e.g. [{"type": "peak_mag", "threshold":
        0.25}, {"type": "median_2d", "epsilon": 0.1, "threshold": 2}]; when
        given it replaces ``median``; an empty list runs no outlier test.
        ``warp_arithmetic``: 'fma' (default) evaluates the bicubic warp's
        interpolants separably with fused multiply-adds (last-bit differences
        in the warped frames, fewer instructions); 'exact' warps with the
        reference's arithmetic, bit for bit.  Lanczos-3 is always exact.
        ``dump_correlation_planes``: keep every pass's raw correlation
        planes and peak outputs in ``PassResult.debug_dump`` (the reference's
        debug dump; N*windows*h*w floats per pass).
        ``max_batch``: when given, every batch-size dependent buffer is
        allocated once for max_batch pairs and each batch size of at most
        max_batch uses the leading part of it (so a caller can capture a
        graph per batch size at the memory cost of one; results of one run
        are overwritten by the next run of any size); ``footprint_bytes()``
        gives that memory without allocating it.
        ``num_peaks``: peaks found and fitted per window; with
        ``secondary_peak`` an invalid vector is replaced by its next peak that
        passes the outlier test (cpu_instantaneous.py:664-722).
        """
        self.H, self.W = image_shape
        if dft not in DFT_MODES:
            raise ValueError(f"dft {dft!r}: one of {sorted(DFT_MODES)} ('fft': FP32 FFT; 'split': split-fp16 "
                             f"MFMA DFT for N <= 64; 'split-bf16', 'split-bf16x3': split-bf16 MFMA DFT for N = 16)")
        self.dft = DFT_MODES[dft]
        self.window_sizes = [tuple(w) if np.ndim(w) else (w, w) for w in window_sizes]
        if len(overlaps) != len(self.window_sizes):
            raise ValueError("one overlap per pass")
        self.overlaps = list(overlaps)
        self.finder = PEAK_FINDERS[peak_finder] if isinstance(peak_finder, str) else int(peak_finder)
        self.impl = 1 if peak_fit_impl == "batch" else 0
        self.interp = 0 if warp_interpolation == "cubic" else 1
        if warp_arithmetic not in ("exact", "fma"):
            raise ValueError(f"warp_arithmetic {warp_arithmetic!r}: 'exact' or 'fma'")
        if warp_arithmetic == "fma" and self.interp == 0:
            self.interp = 2                      # Lanczos-3 keeps the reference's arithmetic
        self.smoothing = bool(predictor_smoothing)
        self.median = median
        self.num_peaks = int(num_peaks)
        if not 1 <= self.num_peaks <= 255:
            raise ValueError(f"num_peaks {num_peaks}: 1..255")
        self.secondary = bool(secondary_peak) and self.num_peaks > 1
        self.dump_planes = bool(dump_correlation_planes)
        self.outliers = {"use_median": 0, "eps": 0.1, "thr": 2.0, "size": 5, "use_peak": 0, "peak_thr": 0.5,
                         "extra": []}
        if outlier_methods is None:
            if median is not None:
                self.outliers.update(use_median=1, eps=float(median[0]), thr=float(median[1]), size=int(median[2]))
        else:
            for mcfg in outlier_methods:
                t = str(mcfg.get("type", "")).lower()
                if t == "median_2d":
                    self.outliers.update(use_median=1, eps=float(mcfg.get("epsilon", 0.1)),
                                         thr=float(mcfg.get("threshold", 2.0)), size=int(mcfg.get("size", 5)))
                elif t == "peak_mag":
                    # outlier_detection.py:463-470: a method entry without a threshold gets 0.5
                    self.outliers.update(use_peak=1, peak_thr=float(mcfg.get("threshold", 0.5)))
                elif t == "sigma":
                    # outlier_detection.py:477-481 (combine 'norm': the vector norm)
                    self.outliers["extra"].append(("sigma", float(mcfg.get("sigma_threshold", 2.0))))
                elif t == "div_vort":
                    dt, vt = mcfg.get("div_thresh"), mcfg.get("vort_thresh")
                    self.outliers["extra"].append(("div_vort", float("nan") if dt is None else float(dt),
                                                   float("nan") if vt is None else float(vt)))
                else:
                    raise ValueError(f"Unknown outlier detection method: {t}")
        if self.outliers["use_median"] and (self.outliers["size"] < 3 or self.outliers["size"] % 2 == 0):
            # outlier_detection.py:170-171
            raise ValueError(f"size must be an odd integer >= 3, got {self.outliers['size']!r}")
        # infilling.py:698-790 apply_infilling: the device path has the
        # nearest-donor and the least-squares (harmonic, biharmonic) methods
        # infilling.py:698-790 apply_infilling: a method name or the config
        # form {"method": ..., "parameters": {...}}
        self.final_infill, self.final_infill_params = _infill_cfg(final_infill, final=True)
        self.mid_infill, self.mid_infill_params = _infill_cfg(mid_infill, final=False)
        self.device = torch.device(device)
        self.lib = _native.core()
        for win in self.window_sizes:
            if not self.lib.pivamd_size_supported(win[0], win[1]):
                raise ValueError(f"window {win} has no device kernel")
        self.grids = build_pass_grids((self.H, self.W), self.window_sizes, self.overlaps)
        self.window_type = window_type
        if vector_masks is not None and pixel_mask is not None:
            raise ValueError("give vector_masks or pixel_mask, not both")
        if pixel_mask is not None:
            vector_masks = _vector_masks((self.H, self.W), self.window_sizes, self.overlaps, pixel_mask,
                                         mask_threshold)
        self.masks = None
        if vector_masks is not None:
            if len(vector_masks) != len(self.grids):
                raise ValueError(f"{len(vector_masks)} vector masks for {len(self.grids)} passes")
            self.masks = []
            for m, g in zip(vector_masks, self.grids):
                m = np.asarray(m).astype(bool)
                if m.shape != g.n_win:
                    raise ValueError(f"vector mask {m.shape} does not match the window grid {g.n_win}")
                self.masks.append(m if m.any() else None)
        if predictor_fill not in ("harmonic", "nearest", "none"):
            raise ValueError(f"predictor_fill {predictor_fill!r}: 'harmonic', 'nearest' or 'none'")
        self.predictor_fill = predictor_fill
        # the harmonic fill's linear map per non-final masked pass (mask geometry only)
        self._fill_host = {}
        if self.masks is not None and predictor_fill == "harmonic":
            for p, (m, g) in enumerate(zip(self.masks[:-1], self.grids[:-1])):
                if m is not None:
                    self._fill_host[p] = harmonic_operator(m, g.spacing)
        self.predictor_bcs = list(predictor_bcs or [])
        for bc in self.predictor_bcs:
            if bc.get("edge") not in ("top", "bottom"):
                raise ValueError(f"predictor BC edge {bc.get('edge')!r}: 'top' or 'bottom'")
        self.max_batch = None if max_batch is None else int(max_batch)
        if self.max_batch is not None and self.max_batch < 1:
            raise ValueError(f"max_batch {max_batch}: at least 1")
        self._const = None
        self._pool = None
        self._bufs: list[_PassBuf] | None = None
        self._n = 0
        # buffer sets per batch size, kept for the life of the object: a HIP
        # graph captured for one batch size keeps pointing at its set
        self._sets: dict = {}

    # -- buffers ---------------------------------------------------------------
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _constants(self):
        """Per-pass tensors that do not depend on the batch size (window
        centres, weights, plane weight, masks, predictor maps and
        smoothing weights), built once and shared by every buffer set."""
        if self._const is not None:
            return self._const
        dev = self.device
        st = self._stream()
        const = []
        for p, g in enumerate(self.grids):
            c = _PassBuf(g)
            c.ctrs = {k: torch.as_tensor(v, device=dev) for k, v in
                      (("x", g.ctrs_x), ("y", g.ctrs_y), ("xa", g.ctrs_x_all), ("ya", g.ctrs_y_all))}
            w = window_weight(g.win, self.window_type)
            c.w = torch.as_tensor(w, device=dev)
            m = self.masks[p] if self.masks is not None else None
            final = p == len(self.grids) - 1
            if m is not None:
                c.mask = torch.as_tensor(m.astype(np.float32), device=dev)
                if p in self._fill_host:
                    c.fill_op = tuple(torch.as_tensor(x, device=dev) for x in self._fill_host[p])
            if not final:
                ya = np.asarray(g.ctrs_y_all, np.float64)
                for bc in self.predictor_bcs:
                    if bc["edge"] == "bottom":
                        rows = np.flatnonzero(ya >= float(self.H - 1 - bc["y_position"]))
                    else:
                        rows = np.flatnonzero(ya <= float(bc["y_position"]))
                    if rows.size:
                        c.bc_rows.append((slice(int(rows[0]), int(rows[-1]) + 1), float(bc["uy"]), float(bc["ux"])))
            c.pw = torch.empty(g.win, dtype=torch.float32, device=dev)
            _native.check(self.lib.pivamd_plane_weight(_ptr(c.w), g.win[0], g.win[1], _ptr(c.pw), st),
                          "plane weight")
            if p > 0:
                c.maps = (torch.as_tensor(g.pred_map_x, device=dev), torch.as_tensor(g.pred_map_y, device=dev))
                if self.smoothing and g.sd > 0:
                    trunc = (g.ksize[0] - 1) / (2 * g.sd)
                    wts, rad = gaussian_weights_1d(g.sd, trunc)
                    c.smooth = (torch.as_tensor(wts, device=dev), rad)
            const.append(c)
        self._const = const
        return const

    def _buffer_specs(self, n):
        """(pass, name, shape, dtype) of every batch-size dependent buffer of
        an n-pair batch; pass -1 holds the warped frame pair."""
        f32, u8 = torch.float32, torch.uint8
        specs = []
        for p, g in enumerate(self.grids):
            ny, nx = g.n_win
            final = p == len(self.grids) - 1
            m = self.masks[p] if self.masks is not None else None
            shape = (n, ny, nx)
            P = self.num_peaks
            pshape = (n, P, ny, nx)
            add = lambda name, shp, dt=f32: specs.append((p, name, tuple(int(x) for x in shp), dt))  # noqa: E731
            for k in ("ratio", "pkh_g", "pkh_o"):
                add(k, shape)
            for k in ("pkx", "pky", "pkh", "sx", "sy", "sxy"):
                add(k, pshape)
            add("choice", shape, u8)
            if self.dump_planes:
                add("planes", (n, ny, nx) + tuple(g.win))
            if P > 1:
                for k in ("pxa", "pya", "pha"):
                    add(k, pshape)
            if self.secondary:
                add("usel", (2,) + shape)
                add("pk_sel", shape)
                add("active", shape, u8)
            for k in ("u", "u2", "u3"):
                add(k, (2,) + shape)
            if m is not None and m.any() and not final and self.predictor_fill != "none":
                add("u4", (2,) + shape)
            add("nanm", shape, u8)
            add("reason", shape, torch.int8)
            add("err", (1,), torch.int32)
            if self.outliers["extra"]:
                add("extra", shape, u8)
                add("om_work", (max(4, self.lib.pivamd_outlier_methods_workspace(n, ny, nx)),), u8)
            nyp, nxp = len(g.ctrs_y_all), len(g.ctrs_x_all)
            for k in ("pad", "pad_s", "pad_t"):
                add(k, (2, n, nyp, nxp))
            if (self.final_infill if final else self.mid_infill) == "knn":
                add("knn_work", (max(4, self.lib.pivamd_infill_knn_workspace(n)),), u8)
            if (self.final_infill if final else self.mid_infill) == "rbf_local":
                add("rbf_work", (max(4, self.lib.pivamd_infill_rbf_workspace(n, ny, nx)),), u8)
            if (self.final_infill if final else self.mid_infill) == "griddata_linear":
                add("gd_work", (max(4, self.lib.pivamd_infill_griddata_workspace(n, ny, nx)),), u8)
            if (self.final_infill if final else self.mid_infill) in ("harmonic", "biharmonic"):
                add("ls_work", (max(4, self.lib.pivamd_infill_ls_workspace(n, ny, nx)),), u8)
            add("work", (max(4, self.lib.pivamd_xcorr_workspace(n, ny, nx, g.win[0], g.win[1], P, 0)),), u8)
            if p > 0:
                add("pred", (2,) + shape)
        specs.append((-1, "warped", (2, n, self.H, self.W), f32))
        return specs

    def footprint_bytes(self, n=None) -> int:
        """Device bytes of the buffers of one n-pair batch (n = max_batch by
        default), without allocating them; with max_batch set this is all
        the batch-size dependent memory the object ever holds."""
        n = self.max_batch if n is None else int(n)
        if n is None:
            raise ValueError("give n or construct with max_batch")
        total = 0
        for _, _, shape, dt in self._buffer_specs(n):
            total += int(np.prod(shape)) * torch.empty((), dtype=dt).element_size()
        return total

    def _alloc(self, n):
        dev = self.device
        const = self._constants()
        if self.max_batch is not None:
            if n > self.max_batch:
                raise ValueError(f"batch of {n} pairs exceeds max_batch={self.max_batch}")
            if self._pool is None:
                # one flat allocation per buffer at the largest batch; every
                # batch size takes views of their leading elements
                self._pool = {(p, k): torch.empty(int(np.prod(shp)), dtype=dt, device=dev)
                              for p, k, shp, dt in self._buffer_specs(self.max_batch)}

            def new(p, k, shp, dt):
                flat = self._pool[(p, k)]
                need = int(np.prod(shp))
                if need > flat.numel():
                    raise RuntimeError(f"buffer {k} of pass {p}: {need} > {flat.numel()} elements")
                return flat[:need].view(shp)
        else:
            def new(p, k, shp, dt):
                return torch.empty(shp, dtype=dt, device=dev)
        bufs = [_PassBuf(c.grid, ctrs=c.ctrs, w=c.w, pw=c.pw, mask=c.mask, maps=c.maps, smooth=c.smooth,
                         fill_op=c.fill_op, bc_rows=c.bc_rows) for c in const]
        warped = None
        for p, k, shp, dt in self._buffer_specs(n):
            t = new(p, k, shp, dt)
            if p < 0:
                warped = t
            elif k == "work":
                bufs[p].work = t
            else:
                bufs[p].out[k] = t
        for b in bufs:
            b.out["err"].zero_()
        self.warped = warped
        self._bufs = bufs
        self._n = n
        self._sets[n] = (bufs, self.warped)

    # -- one batch ---------------------------------------------------------------
    def _begin(self, img_a: torch.Tensor, img_b: torch.Tensor, keep_passes: bool, events):
        """Checks, buffer set of the batch size and the per-run state."""
        if img_a.shape != img_b.shape or img_a.dim() != 3 or tuple(img_a.shape[1:]) != (self.H, self.W):
            raise ValueError(f"expected two (N, {self.H}, {self.W}) stacks, got {tuple(img_a.shape)}")
        n = int(img_a.shape[0])
        if self._bufs is None or self._n != n:
            if n in self._sets:
                self._bufs, self.warped = self._sets[n]
                self._n = n
            else:
                self._alloc(n)
        st = self._stream()
        L = self.lib
        if img_a.dtype != img_b.dtype:
            raise TypeError("the two frames of a pair must have the same dtype")
        if img_a.dtype == torch.float32:
            u16 = False
        elif img_a.dtype in (torch.int16, torch.uint16):
            u16 = True      # the kernels convert the 16-bit counts on load (exact)
        else:
            raise TypeError(f"unsupported image dtype {img_a.dtype}")
        A, B = img_a.contiguous(), img_b.contiguous()
        return SimpleNamespace(A=A, B=B, n=n, u16=u16, keep=keep_passes, events=events, results=[], prev=None,
                               pred=None, ev=None)

    def _pre(self, c, p):
        """Pass p up to its correlation: predictor smoothing and remap and
        the warp (p > 0).  Returns the frames the correlation reads."""
        st, L, n = self._stream(), self.lib, c.n
        A, B = c.A, c.B
        b = self._bufs[p]
        prev = c.prev
        warp_fn = L.pivamd_warp_u16 if c.u16 else L.pivamd_warp
        g = b.grid
        ny, nx = g.n_win
        o = b.out
        final = p == len(self._bufs) - 1
        pred = None
        if p == 0:
            a_in, b_in = A, B
        else:
            field = prev.out["pad"]
            nyp, nxp = field.shape[2], field.shape[3]
            if b.smooth is not None:
                wts, rad = b.smooth
                _native.check(L.pivamd_smooth_axis(_ptr(field), 2 * n, nyp, nxp, 0, _ptr(wts), rad,
                                                   _ptr(prev.out["pad_t"]), st), "smooth")
                _native.check(L.pivamd_smooth_axis(_ptr(prev.out["pad_t"]), 2 * n, nyp, nxp, 1, _ptr(wts), rad,
                                                   _ptr(prev.out["pad_s"]), st), "smooth")
                field = prev.out["pad_s"]
            pred = o["pred"]
            _native.check(L.pivamd_remap_cubic(_ptr(field), n, nyp, nxp, _ptr(b.maps[0]), _ptr(b.maps[1]),
                                               ny, nx, _ptr(pred), st), "predictor remap")
            _native.check(warp_fn(_ptr(A), _ptr(B), _ptr(self.warped[0]), _ptr(self.warped[1]),
                                        _ptr(field[0]), _ptr(field[1]), n, self.H, self.W, nyp, nxp,
                                        _ptr(prev.ctrs["ya"]), _ptr(prev.ctrs["xa"]), self.interp, 0, 0, st),
                          "warp")
            a_in, b_in = self.warped[0], self.warped[1]
        c.pred = pred
        return a_in, b_in

    def _correlate(self, c, p, a_in, b_in):
        """The correlation call of pass p (one launch plus the fit)."""
        st, L, n = self._stream(), self.lib, c.n
        b = self._bufs[p]
        g = b.grid
        ny, nx = g.n_win
        o = b.out
        events = c.events
        xcorr_fn = L.pivamd_xcorr_u16 if c.u16 else L.pivamd_xcorr
        if events is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
        self._zero_skipped(p)
        prev_mode = L.pivamd_set_dft_mode(self.dft)
        fn = xcorr_fn if p == 0 else L.pivamd_xcorr
        rc = fn(_ptr(a_in), _ptr(b_in), _ptr(b.mask), self.H, self.W, n, _ptr(b.ctrs["x"]),
                                     _ptr(b.ctrs["y"]), ny, nx, _ptr(b.w), _ptr(b.w), _ptr(b.pw), 0,
                                     g.win[0], g.win[1], self.num_peaks, self.finder, self.impl, _ptr(o["pkx"]),
                                     _ptr(o["pky"]), _ptr(o["pkh"]), _ptr(o["sx"]), _ptr(o["sy"]),
                                     _ptr(o["sxy"]), _ptr(o["ratio"]), _ptr(o.get("planes")), _ptr(b.work),
                                     b.work.numel(), st)
        L.pivamd_set_dft_mode(prev_mode)
        _native.check(rc, "bulkxcorr2d")
        if events is not None:
            ev[1].record()
            events.append((p, ev))

    def _zero_skipped(self, p):
        """Outputs of windows the correlation skips (masked, out of the
        image) keep zeros; the plane dump holds zero planes there."""
        b = self._bufs[p]
        o = b.out
        if b.mask is not None:
            for k in ("pkx", "pky", "pkh", "sx", "sy", "sxy", "ratio"):
                o[k].zero_()
        if self.dump_planes:
            o["planes"].zero_()

    def _job(self, c, p, a_in, b_in):
        """The correlation call of pass p as a ragged-launch job."""
        b = self._bufs[p]
        g = b.grid
        o = b.out
        J = _native.XcorrJob()
        J.imgA, J.imgB = _ptr(a_in), _ptr(b_in)
        J.u16 = 1 if (p == 0 and c.u16) else 0
        J.H, J.W, J.n_images = self.H, self.W, c.n
        J.mask, J.ctrX, J.ctrY = _ptr(b.mask), _ptr(b.ctrs["x"]), _ptr(b.ctrs["y"])
        J.nWy, J.nWx = g.n_win
        J.wA = J.wB = _ptr(b.w)
        J.planeWeight = _ptr(b.pw)
        J.pkX, J.pkY, J.pkH = _ptr(o["pkx"]), _ptr(o["pky"]), _ptr(o["pkh"])
        J.sx, J.sy, J.sxy, J.ratio = _ptr(o["sx"]), _ptr(o["sy"]), _ptr(o["sxy"]), _ptr(o["ratio"])
        J.planes = _ptr(o.get("planes"))
        J.work, J.work_bytes = _ptr(b.work), b.work.numel()
        return J

    def _post(self, c, p):
        """Pass p after its correlation: gate, outlier tests, infill, the
        padded field for the next pass and the pass result."""
        st, L, n = self._stream(), self.lib, c.n
        b = self._bufs[p]
        g = b.grid
        ny, nx = g.n_win
        o = b.out
        final = p == len(self._bufs) - 1
        pred = c.pred
        keep_passes = c.keep
        results = c.results
        div = 2.0 if p == 0 else 4.0
        nw = ny * nx
        if self.num_peaks == 1:
            _native.check(L.pivamd_gate(_ptr(o["pkx"]), _ptr(o["pky"]), _ptr(o["pkh"]), _ptr(pred),
                                        _ptr(b.mask), n, nw, g.win[1] / div, g.win[0] / div, _ptr(o["u"]),
                                        _ptr(o["nanm"]), _ptr(o["reason"]), _ptr(o["pkh_g"]), st), "gate")
            peak_mag = o["pkh_g"].view(n, 1, ny, nx)
        else:
            _native.check(L.pivamd_gate_peaks(_ptr(o["pkx"]), _ptr(o["pky"]), _ptr(o["pkh"]), _ptr(pred),
                                              _ptr(b.mask), n, self.num_peaks, nw, g.win[1] / div,
                                              g.win[0] / div, _ptr(o["pxa"]), _ptr(o["pya"]), _ptr(o["pha"]),
                                              _ptr(o["u"]), _ptr(o["nanm"]), _ptr(o["reason"]),
                                              _ptr(o["pkh_g"]), _ptr(o["choice"]), st), "gate")
            peak_mag = o["pha"]
        oc = self.outliers
        extra = self._extra_methods(o, o["u"], b.mask, n, ny, nx, st)
        if not self.secondary:
            _native.check(L.pivamd_outlier_stage(_ptr(o["u"]), _ptr(o["pkh_g"]), _ptr(b.mask), n, ny, nx,
                                                 oc["use_median"], oc["eps"], oc["thr"], oc["size"],
                                                 oc["use_peak"], oc["peak_thr"], _ptr(extra), _ptr(o["nanm"]),
                                                 _ptr(o["reason"]), _ptr(o["u2"]), _ptr(o["pkh_o"]),
                                                 _ptr(o["choice"]), st), "outlier stage")
        else:
            # first round on the primary peak, then one substitution round per further peak
            _native.check(L.pivamd_outlier_detect(_ptr(o["u"]), _ptr(o["pkh_g"]), _ptr(b.mask), n, ny, nx,
                                                  oc["use_median"], oc["eps"], oc["thr"], oc["size"],
                                                  oc["use_peak"], oc["peak_thr"], _ptr(extra), None,
                                                  _ptr(o["nanm"]), _ptr(o["reason"]), st), "outlier stage")
            for _ in range(1, self.num_peaks):
                _native.check(L.pivamd_peak_select(_ptr(o["pxa"]), _ptr(o["pya"]), _ptr(o["pha"]), n,
                                                   self.num_peaks, nw, _ptr(o["nanm"]), _ptr(o["reason"]),
                                                   _ptr(o["choice"]), _ptr(o["usel"]), _ptr(o["pk_sel"]),
                                                   _ptr(o["active"]), st), "peak select")
                ex = self._extra_methods(o, o["usel"], b.mask, n, ny, nx, st)
                _native.check(L.pivamd_outlier_detect(_ptr(o["usel"]), _ptr(o["pk_sel"]), _ptr(b.mask), n, ny,
                                                      nx, oc["use_median"], oc["eps"], oc["thr"], oc["size"],
                                                      oc["use_peak"], oc["peak_thr"], _ptr(ex),
                                                      _ptr(o["active"]), _ptr(o["nanm"]), _ptr(o["reason"]),
                                                      st), "secondary outlier stage")
            _native.check(L.pivamd_peak_final(_ptr(o["usel"]), _ptr(o["pk_sel"]), _ptr(b.mask), n, nw,
                                              _ptr(o["nanm"]), _ptr(o["reason"]), _ptr(o["choice"]),
                                              _ptr(o["u2"]), _ptr(o["pkh_o"]), st), "peak final")
        u = o["u2"]
        method = self.final_infill if final else self.mid_infill
        if method == "nearest":
            o["err"].zero_()
            _native.check(L.pivamd_infill_nearest(_ptr(u), _ptr(b.mask), n, ny, nx, float(g.spacing[0]),
                                                  float(g.spacing[1]), _ptr(o["u3"]), _ptr(o["err"]), st),
                          "infill")
            u = o["u3"]
        elif method == "knn":
            o["err"].zero_()
            kp = self.final_infill_params if final else self.mid_infill_params
            wk = o["knn_work"]
            _native.check(L.pivamd_infill_knn(_ptr(u), _ptr(b.mask), n, ny, nx, float(g.spacing[0]),
                                              float(g.spacing[1]), int(kp.get("n_neighbors", 32)),
                                              1 if kp.get("weights", "distance") == "uniform" else 0,
                                              _ptr(o["u3"]), _ptr(wk), wk.numel(), _ptr(o["err"]), st), "infill")
            u = o["u3"]
        elif method == "rbf_local":
            o["err"].zero_()
            kp = self.final_infill_params if final else self.mid_infill_params
            wk = o["rbf_work"]
            _native.check(L.pivamd_infill_rbf(_ptr(u), _ptr(b.mask), n, ny, nx, float(g.spacing[0]),
                                              float(g.spacing[1]), kp["neighbors"], RBF_KERNELS[kp["kernel"]],
                                              kp["epsilon"], kp["smoothing"], _ptr(o["u3"]), _ptr(wk),
                                              wk.numel(), _ptr(o["err"]), st), "infill")
            u = o["u3"]
        elif method == "griddata_linear":
            o["err"].zero_()
            wk = o["gd_work"]
            _native.check(L.pivamd_infill_griddata(_ptr(u), _ptr(b.mask), n, ny, nx, float(g.spacing[0]),
                                                   float(g.spacing[1]), _ptr(o["u3"]), _ptr(wk), wk.numel(),
                                                   _ptr(o["err"]), st), "infill")
            u = o["u3"]
        elif method in ("harmonic", "biharmonic"):
            o["err"].zero_()
            wk = o["ls_work"]
            _native.check(L.pivamd_infill_ls(_ptr(u), _ptr(b.mask), n, ny, nx, float(g.spacing[0]),
                                             float(g.spacing[1]), 1 if method == "harmonic" else 2,
                                             _ptr(o["u3"]), _ptr(wk), wk.numel(), _ptr(o["err"]), st), "infill")
            u = o["u3"]
        if not final:
            nyp, nxp = len(g.ctrs_y_all), len(g.ctrs_x_all)
            src = u
            if "u4" in o:
                # the predictor's copy of the field with its masked windows filled;
                # the pass result keeps zeros there
                src = o["u4"]
                if self.predictor_fill == "nearest":
                    _native.check(L.pivamd_fill_masked_nearest(_ptr(u), _ptr(b.mask), n, ny, nx,
                                                               float(g.spacing[0]), float(g.spacing[1]),
                                                               _ptr(src), _ptr(o["err"]), st), "predictor fill")
                else:
                    src.copy_(u)
                    fi, bi, gt = b.fill_op
                    _native.check(L.pivamd_fill_linear(_ptr(src), 2 * n, ny * nx, fi.numel(), _ptr(fi),
                                                       bi.numel(), _ptr(bi), _ptr(gt), st), "predictor fill")
            _native.check(L.pivamd_pad_edge(_ptr(src), n, ny, nx, g.n_pre[0], g.n_pre[1], nyp, nxp,
                                            _ptr(o["pad"]), st), "pad")
            for rows, uy_bc, ux_bc in b.bc_rows:
                o["pad"][0, :, rows, :].fill_(uy_bc)
                o["pad"][1, :, rows, :].fill_(ux_bc)
        if keep_passes or final:
            results.append(PassResult(ux=u[1], uy=u[0], nan_mask=o["nanm"], nan_reason=o["reason"],
                                      peak_mag=peak_mag, peak_ratio=o["ratio"], predictor=pred,
                                      window_size=g.win, n_windows=(ny, nx), primary_peak_mag=o["pkh_o"],
                                      peak_choice=o["choice"],
                                      b_mask=(self.masks[p] if self.masks is not None and self.masks[p] is not None
                                              else np.zeros((ny, nx), bool)),
                                      win_ctrs_x=g.ctrs_x, win_ctrs_y=g.ctrs_y,
                                      debug_dump=None if not self.dump_planes else {
                                          "planes": o["planes"], "pk_loc_x_raw": o["pkx"],
                                          "pk_loc_y_raw": o["pky"], "pk_height_raw": o["pkh"],
                                          "pk_ratio_raw": o["ratio"], "sx": o["sx"], "sy": o["sy"],
                                          "sxy": o["sxy"], "peak_finder": int(self.finder)}))
        c.prev = b

    def run(self, img_a: torch.Tensor, img_b: torch.Tensor, keep_passes: bool = False, events=None):
        """Correlate a batch of pairs (N, H, W) already on the device.

        ``img_a``/``img_b`` are float32, or 16-bit counts (uint16 data in an
        int16 or uint16 tensor).  Returns the list of ``PassResult`` (all
        passes if ``keep_passes``, else the final one) as device tensors.
        ``events`` (optional list) receives, per pass, a pair of recorded
        torch.cuda.Events bracketing the correlation launch on this stream.
        """
        c = self._begin(img_a, img_b, keep_passes, events)
        for p in range(len(self._bufs)):
            a_in, b_in = self._pre(c, p)
            self._correlate(c, p, a_in, b_in)
            self._post(c, p)
        return c.results

    def _extra_methods(self, o, u, mask, n, ny, nx, st):
        """Flags of the sigma / div_vort methods on the planar field u, or None."""
        if not self.outliers["extra"]:
            return None
        L = self.lib
        extra = o["extra"]
        extra.zero_()
        wk = o["om_work"]
        for m in self.outliers["extra"]:
            if m[0] == "sigma":
                rc = L.pivamd_outlier_sigma(_ptr(u), _ptr(mask), n, ny, nx, m[1], 0, _ptr(extra), _ptr(wk),
                                            wk.numel(), st)
            else:
                rc = L.pivamd_outlier_div_vort(_ptr(u), _ptr(mask), n, ny, nx, m[1], m[2], _ptr(extra), _ptr(wk),
                                               wk.numel(), st)
            _native.check(rc, f"{m[0]} outliers")
        return extra

    def check_infill(self) -> None:
        """Raise as the reference's infill would for the last batch: a field
        with no donor (nearest; carried on as zero displacement), a
        least-squares fill that missed the 1e-6 residual acceptance, or fill
        cells no difference equation ties to a valid cell (left at 0)."""
        last = len(self.window_sizes) - 1
        for p, b in enumerate(self._bufs or []):
            e = int(b.out["err"].item())
            if (self.final_infill if p == last else self.mid_infill) == "rbf_local" and e:
                if e & 1:
                    raise ValueError(f"pass {p}: Infilling method 'rbf_local' cannot run: fewer than 4 valid "
                                     "donor points (or fewer than the polynomial terms)")
                if e & 2:
                    raise np.linalg.LinAlgError(f"pass {p}: rbf_local: Singular matrix.")
                raise ValueError(f"pass {p}: rbf_local: the neighbour search exceeded its buffer "
                                 "(grid pitch too anisotropic)")
            if (self.final_infill if p == last else self.mid_infill) == "griddata_linear" and e:
                if e & 1:
                    raise ValueError(f"pass {p}: Infilling method 'griddata_linear' cannot run: no valid donor")
                raise ValueError(f"pass {p}: griddata_linear: a hole's Delaunay triangle reaches beyond its 192 "
                                 "nearest donors" if e & 2 else
                                 f"pass {p}: griddata_linear: the neighbour search exceeded its buffer")
            if e & 1:
                raise ValueError(f"pass {p}: infilling has no donor points: 100% of vectors invalid")
            if e & 4:
                raise ValueError(f"pass {p}: least-squares infill: a group of fill cells is not linked "
                                 "to any valid vector")
            if e & 2:
                raise np.linalg.LinAlgError(f"pass {p}: least-squares infill did not reach a relative "
                                            "residual of 1e-6")


def run_ragged(items, keep_passes: bool = False):
    """Several pass loops of one schedule over batches of different image
    sizes (``items``: (MultipassPIV, img_a, img_b) triples on the current
    stream), pass by pass: each pass correlates every batch in ONE ragged
    launch (``pivamd_xcorr_ragged``: the windows of all batches share one
    grid of workgroups) between the batches' own warps and glue.  Each
    batch's results equal its own ``run``.  Returns one result list per item.
    The schedules must match pass for pass (square windows; a rectangular
    pass runs each batch's own correlation) and the peak settings must agree;
    the split-DFT mode is not used."""
    if not items:
        return []
    mps = [it[0] for it in items]
    m0 = mps[0]
    for m in mps[1:]:
        if m.window_sizes != m0.window_sizes:
            raise ValueError("ragged batches need the same window schedule")
        if (m.num_peaks, m.finder, m.impl) != (m0.num_peaks, m0.finder, m0.impl):
            raise ValueError("ragged batches need the same peak settings")
        if m.device != m0.device:
            raise ValueError("ragged batches need one device")
    if len(items) > 8:
        raise ValueError("at most 8 batches per ragged launch")
    if any(m.dft for m in mps):
        raise ValueError("the ragged launch runs the FP32 FFT forms (dft='fft')")
    ctx = [m._begin(a, b, keep_passes, None) for m, a, b in items]
    L = m0.lib
    st = m0._stream()
    for p, win in enumerate(m0.window_sizes):
        ins = [m._pre(c, p) for m, c in zip(mps, ctx)]
        if win[0] == win[1]:
            for m in mps:
                m._zero_skipped(p)
            jobs = (_native.XcorrJob * len(mps))(*[m._job(c, p, a, b) for m, c, (a, b) in zip(mps, ctx, ins)])
            _native.check(L.pivamd_xcorr_ragged(jobs, len(mps), win[0], win[1], m0.num_peaks, m0.finder, m0.impl,
                                                st), "ragged bulkxcorr2d")
        else:
            for m, c, (a, b) in zip(mps, ctx, ins):
                m._correlate(c, p, a, b)
        for m, c in zip(mps, ctx):
            m._post(c, p)
    return [c.results for c in ctx]
