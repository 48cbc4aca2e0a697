"""Device-resident ensemble PIV pass (standard and 'single' run types).

Follows the reference's streaming ensemble pass:

* per batch (``cpu_ensemble.py:826-1084``): image sums for the mean images
  and ``bulkxcorr2d_accumulate_triple`` of every pair into the AB, AA, BB
  plane sums (weights wA / wB for AB, wB for both autos);
* finalize (``single_pass_accumulator.py:508-1290``): averages, background
  correlation of the mean images (``correlation`` base method), optional
  plane-mean removal (``+dc_zero``) and envelope divide, normalisation by
  the auto peaks, the coloured noise-floor shape P (flat for ``flat``), the
  k-space LM fit, and displacement / Reynolds-stress extraction.

The ``single`` run type (``cpu_ensemble.py:188-229,695-750,963-1010``,
``single_pass_accumulator.py:388-506,700-795,855-880,1191-1210``) puts a
small A window (``singlepix``: ones on the window block) inside a sum
window (``bsingle``: all ones) on a grid spaced by the small window: the
frames get a zero border so a sum window fits around every position (one
device pass that also converts 16-bit counts), AB correlates wA with wB and
both autos wB with wB at the sum size, cropped centrally to the fitting
window; finalize divides AA/BB by the wB*wB envelope and AB by the wA*wB
one (when the envelope divide is on), scales AB by sum area / A area, takes
the coloured floor of the wB chain, and scales the peak height by
sqrt(sum area / A area).

Every step is a HIP kernel of libpivtools_amd.so on torch's current stream;
only the floor shape P is host work (one small analytic model per pass).
The plane sums stay in HBM for the whole ensemble; the fitter inputs are
FP64.

``EnsembleMultipass`` chains passes as the reference's ensemble does
(``cpu_ensemble.py:912-960,1445-1660``, ``single_pass_accumulator.py:
820-890,1146-1400``): pass p > 0 warps every pair with the previous pass's
field as one shared predictor (edge-padded, optionally smoothed and rounded
to even integers), accumulates the warped pairs on its own grid, takes the
per-window half-shift fractions of the remapped predictor into the coloured
floor, adds the predictor back to the fitted displacements, and infills
failed fits before they feed the next pass.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _native, floor_psd
from .window import build_pass_grids, window_centers_original, window_weight

BG_METHODS = {"correlation", "window_mean", "correlation+window_mean", "correlation+dc_zero", "none"}
RUN_TYPES = ("standard", "single")


@dataclass
class EnsembleResult:
    """Per-window outputs of one ensemble pass (device tensors, (ny, nx))."""

    ux: torch.Tensor
    uy: torch.Tensor
    UU: torch.Tensor
    VV: torch.Tensor
    UV: torch.Tensor
    peakheight: torch.Tensor
    nan_reason: torch.Tensor
    gauss: torch.Tensor          # (n_windows, 16) FP64 rows of the k-space fit
    status: torch.Tensor
    init: torch.Tensor
    diag: dict
    planes: tuple                # FP64 (AA, BB, AB) normalised planes
    n_images: int


class EnsemblePIV:
    def __init__(self, image_shape, window_size=(32, 32), overlap=75, window_type="square",
                 bg_method="correlation", kspace_floor="coloured", kspace_shape="gaussian",
                 envelope_divide=False, per_pair_normalization=False, warp_kernel="cubic",
                 mask=None, device=None, run_type="standard", sum_window=(32, 32), fit_window=None):
        if run_type not in RUN_TYPES:
            raise ValueError(f"run type {run_type!r}: one of {RUN_TYPES}")
        if bg_method not in BG_METHODS:
            raise ValueError(f"background method {bg_method!r} not in {sorted(BG_METHODS)}")
        if kspace_floor not in ("coloured", "flat"):
            raise ValueError(f"kspace_floor {kspace_floor!r}")
        if per_pair_normalization and bg_method != "window_mean":
            raise ValueError("per-pair normalisation requires the window_mean background method")
        L = _native.core()
        self.single = run_type == "single"
        self.run_type = run_type
        self.win = (int(window_size[0]), int(window_size[1]))
        if self.single:
            self.comp = (int(sum_window[0]), int(sum_window[1]))
            self.corr = self.comp if fit_window is None else (int(fit_window[0]), int(fit_window[1]))
            if self.corr[0] > self.comp[0] or self.corr[1] > self.comp[1]:
                raise ValueError(f"fitting window {self.corr} exceeds the sum window {self.comp}")
            if not (L.pivamd_size_supported(self.corr[0], self.corr[0]) and
                    L.pivamd_size_supported(self.corr[1], self.corr[1])):
                raise ValueError(f"fitting window {self.corr}: each axis must be a built FFT size")
        else:
            if fit_window is not None:
                raise ValueError("a fitting window applies to the single run type only")
            self.comp = self.corr = self.win
        if self.comp[0] != self.comp[1] or not L.pivamd_size_supported(self.comp[0], self.comp[1]):
            raise ValueError(f"correlation size {self.comp} has no device ensemble kernel (square built sizes)")
        self.L = L
        self.dev = torch.device(device) if device is not None else torch.device("cuda")
        self.H, self.W = (int(image_shape[0]), int(image_shape[1]))
        self.overlap = overlap
        self.grid = window_centers_original((self.H, self.W), self.win, overlap, run_type, self.comp)
        self.padding = self.grid.padding            # (top, bottom, left, right), zero in standard mode
        pt, pb, pl, pr = self.padding
        self.pad = (pt, pl)
        self.Hp, self.Wp = self.H + pt + pb, self.W + pl + pr
        self.ny, self.nx = int(self.grid.n_win_y), int(self.grid.n_win_x)
        self.nwin = self.ny * self.nx
        self.bg_method = bg_method
        self.mean_sub = bg_method in ("window_mean", "correlation+window_mean")
        self.bg_correlation = bg_method.startswith("correlation")
        self.dc_zero = bg_method.endswith("dc_zero")
        self.per_pair_norm = per_pair_normalization
        self.floor = kspace_floor
        self.kx4 = int("kx4" in kspace_shape)
        self.ky4 = int("ky4" in kspace_shape)
        self.envelope_divide = envelope_divide
        self.warp_kernel = warp_kernel
        d, f32 = self.dev, torch.float32
        if self.single:
            wa = window_weight(self.win, "singlepix", self.comp)
            wb = window_weight(self.comp, "bsingle", self.comp)
            area_sum, area_a = self.comp[0] * self.comp[1], self.win[0] * self.win[1]
            self.ab_scale = float(np.float32(area_sum / area_a))
            self.peak_scale = float(np.sqrt(area_sum / area_a))
        else:
            wa = wb = window_weight(self.win, window_type)
            self.ab_scale = self.peak_scale = 1.0
        self.wa_host, self.w_host = wa, wb          # w_host: the B / auto weight (floor chain)
        self.wa = torch.as_tensor(wa, device=d)
        self.w = self.wa if wb is wa else torch.as_tensor(wb, device=d)
        self.sum_w = float(np.float32(wb.sum(dtype=np.float32)))
        # the correlation takes padded coordinates (the padding is zero in standard mode)
        self.ctr_x = torch.as_tensor((self.grid.win_ctrs_x + np.float32(pl)).astype(np.float32), device=d)
        self.ctr_y = torch.as_tensor((self.grid.win_ctrs_y + np.float32(pt)).astype(np.float32), device=d)
        self.mask_host = (np.zeros(self.nwin, bool) if mask is None
                          else np.asarray(mask, bool).reshape(-1))
        self.mask_f = torch.as_tensor(self.mask_host.astype(np.float32), device=d)
        self.mask_u8 = torch.as_tensor(self.mask_host.astype(np.uint8), device=d)
        P = self.corr[0] * self.corr[1]
        self.sums = [torch.zeros(self.nwin * P, dtype=f32, device=d) for _ in range(3)]   # AB, AA, BB
        # image sums at the padded size (the border stays zero, so the mean
        # images come out already padded for the background correlation)
        self.sum_a = torch.zeros(self.Hp * self.Wp, dtype=torch.float64, device=d)
        self.sum_b = torch.zeros(self.Hp * self.Wp, dtype=torch.float64, device=d)
        nbytes = int(L.pivamd_ensemble_workspace(self.ny, self.nx, self.comp[0], self.comp[1]))
        self.work = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=d)
        self._f32 = [None, None]
        self.n_images = 0
        self.predictor = None      # (2, ny, nx) [dy, dx] at the window centres (pass > 0)

    def set_predictor(self, pred: torch.Tensor) -> None:
        """The predictor the pairs of this pass were warped with, at this
        pass's window centres (planar (2, ny, nx), dy then dx): it sets the
        coloured floor's per-window half-shift fractions and is added back to
        the fitted displacements (single_pass_accumulator.py:834-878,
        1146-1162)."""
        pred = torch.as_tensor(pred, device=self.dev, dtype=torch.float32).reshape(2, self.ny, self.nx)
        if not bool(torch.isfinite(pred).all()):
            raise ValueError("non-finite predictor: cannot derive the coloured-floor fractions")
        self.predictor = pred.contiguous()

    # ------------------------------------------------------------------
    def _stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _buffer(self, slot: int, numel: int) -> torch.Tensor:
        buf = self._f32[slot]
        if buf is None or buf.numel() < numel:
            buf = self._f32[slot] = torch.empty(numel, dtype=torch.float32, device=self.dev)
        return buf[:numel]

    def _frames(self, x: torch.Tensor, slot: int) -> torch.Tensor:
        """float32 (n, Hp, Wp) frames for the correlation: 16-bit counts are
        converted, and in single mode the zero border is added in the same
        device pass."""
        if x.dtype not in (torch.float32, torch.int16, torch.uint16):
            raise TypeError(f"unsupported image dtype {x.dtype}")
        x = x.contiguous()
        n = int(x.shape[0])
        if self.single:
            out = self._buffer(slot, n * self.Hp * self.Wp).view(n, self.Hp, self.Wp)
            _native.check(self.L.pivamd_pad_frames(x.data_ptr(), int(x.dtype != torch.float32), n, self.H, self.W,
                                                   self.pad[0], self.pad[1], self.Hp, self.Wp, out.data_ptr(),
                                                   self._stream()), "pad_frames")
            return out
        if x.dtype == torch.float32:
            return x
        out = self._buffer(slot, x.numel()).view(x.shape)
        _native.check(self.L.pivamd_u16_to_f32(x.data_ptr(), out.data_ptr(), x.numel(), self._stream()),
                      "u16_to_f32")
        return out

    def _accumulate(self, a, b, n, outs, per_pair_norm):
        st = self._stream()
        C, F = self.comp, self.corr
        _native.check(self.L.pivamd_accumulate(
            a.data_ptr(), b.data_ptr(), self.mask_f.data_ptr(), self.Hp, self.Wp, n, self.ctr_x.data_ptr(),
            self.ctr_y.data_ptr(), self.ny, self.nx, self.wa.data_ptr(), self.w.data_ptr(), self.w.data_ptr(),
            self.w.data_ptr(), C[0], C[1], F[0], F[1], int(self.mean_sub), int(per_pair_norm),
            1, self.sum_w, self.sum_w, self.sum_w, outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr(),
            self.work.data_ptr(), self.work.numel(), st), "accumulate_triple")

    def accumulate(self, img_a: torch.Tensor, img_b: torch.Tensor) -> None:
        """Add a batch of pairs (n, H, W) (float32 or 12-bit counts in uint16)."""
        if img_a.shape != img_b.shape or img_a.dim() != 3 or tuple(img_a.shape[1:]) != (self.H, self.W):
            raise ValueError(f"expected (n, {self.H}, {self.W}) frames, got {tuple(img_a.shape)}")
        n = int(img_a.shape[0])
        a = self._frames(img_a, 0)
        b = self._frames(img_b, 1)
        st = self._stream()
        hw = self.Hp * self.Wp
        _native.check(self.L.pivamd_sum_images(a.data_ptr(), n, hw, self.sum_a.data_ptr(), st), "sum_images")
        _native.check(self.L.pivamd_sum_images(b.data_ptr(), n, hw, self.sum_b.data_ptr(), st), "sum_images")
        self._accumulate(a, b, n, self.sums, self.per_pair_norm)
        self.n_images += n

    # ------------------------------------------------------------------
    def floor_shape(self):
        """Floor shape P in the centred layout: one shared (h, w) shape when
        every window sits at half-shift fraction 0 (no predictor, or an
        even-integer one), else one shape per window (n, h * w) from the
        fractions frac(pred / 2) of the predictor at the window centres."""
        corr = self.corr
        env_auto = floor_psd.pair_envelope(self.w_host, self.w_host, corr)
        env_for_floor = env_auto if self.envelope_divide else np.ones_like(env_auto)
        if self.predictor is None:
            fx = fy = np.zeros(1)
        else:
            pr = self.predictor.double().cpu().numpy()
            fx = floor_psd.frac_distance(pr[1].ravel() / 2.0)
            fy = floor_psd.frac_distance(pr[0].ravel() / 2.0)
        n_fracs = 2 if max(fx.max(), fy.max()) == 0.0 else 21
        fs, grid = floor_psd.build_p_grid(env_for_floor, self.warp_kernel, "window_mean" in self.bg_method,
                                          self.w_host, n_fracs=n_fracs)
        if n_fracs == 2:
            return floor_psd.interp_p(fs, grid, [0.0], [0.0])[0].astype(np.float32).astype(np.float64)
        return floor_psd.interp_p(fs, grid, fx, fy).astype(np.float32).astype(np.float64)

    def finalize(self) -> EnsembleResult:
        if self.n_images == 0:
            raise RuntimeError("no pairs accumulated")
        L, d, st = self.L, self.dev, self._stream()
        h, w = self.corr
        P = h * w
        hw = self.Hp * self.Wp
        f32, f64 = torch.float32, torch.float64
        bg = None
        if self.bg_correlation:
            ma = torch.empty(hw, dtype=f32, device=d)
            mb = torch.empty(hw, dtype=f32, device=d)
            _native.check(L.pivamd_mean_image(self.sum_a.data_ptr(), float(self.n_images), hw, ma.data_ptr(), st),
                          "mean_image")
            _native.check(L.pivamd_mean_image(self.sum_b.data_ptr(), float(self.n_images), hw, mb.data_ptr(), st),
                          "mean_image")
            bg = [torch.zeros(self.nwin * P, dtype=f32, device=d) for _ in range(3)]
            self._accumulate(ma, mb, 1, bg, False)
        env = env_ab = None
        if self.envelope_divide:
            e = floor_psd.pair_envelope(self.w_host, self.w_host, self.corr).astype(np.float32)
            env = env_ab = torch.as_tensor(e, device=d)
            if self.single:
                e = floor_psd.pair_envelope(self.wa_host, self.w_host, self.corr).astype(np.float32)
                env_ab = torch.as_tensor(e, device=d)
        planes = [torch.empty(self.nwin * P, dtype=f64, device=d) for _ in range(3)]   # AA, BB, AB
        _native.check(L.pivamd_ensemble_finalize(
            self.sums[1].data_ptr(), self.sums[2].data_ptr(), self.sums[0].data_ptr(),
            None if bg is None else bg[1].data_ptr(), None if bg is None else bg[2].data_ptr(),
            None if bg is None else bg[0].data_ptr(), None if env is None else env.data_ptr(),
            None if env_ab is None else env_ab.data_ptr(), self.nwin, h, w, float(np.float32(self.n_images)),
            self.ab_scale, int(self.dc_zero), planes[0].data_ptr(), planes[1].data_ptr(), planes[2].data_ptr(), st),
            "ensemble_finalize")
        pw = None
        if self.floor == "coloured":
            pw = torch.as_tensor(self.floor_shape(), device=d)
        n = self.nwin
        gauss = torch.empty((n, 16), dtype=f64, device=d)
        init = torch.empty((n, 16), dtype=f64, device=d)
        status = torch.empty(n, dtype=torch.int32, device=d)
        diag = {k: torch.empty(n, dtype=f64, device=d) for k in ("gain", "N0", "b4x", "b4y", "cost_per_pt")}
        diag.update({k: torch.empty(n, dtype=torch.int32, device=d) for k in ("n_valid", "iter")})
        diag["conv"] = torch.empty(n, dtype=torch.uint8, device=d)
        wb = int(L.pivamd_kspace_workspace(n, h, w))
        work = torch.empty(max(wb, 256), dtype=torch.uint8, device=d)
        _native.check(L.pivamd_kspace_fit(
            planes[0].data_ptr(), planes[1].data_ptr(), planes[2].data_ptr(), self.mask_u8.data_ptr(),
            None if pw is None else pw.data_ptr(), n, h, w, self.kx4, self.ky4, gauss.data_ptr(),
            status.data_ptr(), init.data_ptr(), diag["gain"].data_ptr(), diag["N0"].data_ptr(),
            diag["b4x"].data_ptr(), diag["b4y"].data_ptr(), diag["cost_per_pt"].data_ptr(),
            diag["n_valid"].data_ptr(), diag["iter"].data_ptr(), diag["conv"].data_ptr(),
            1 if pw is None or pw.dim() == 1 else 0, work.data_ptr(), work.numel(), st), "kspace_fit")
        out = [torch.empty(n, dtype=f32, device=d) for _ in range(6)]
        reason = torch.empty(n, dtype=torch.int32, device=d)
        _native.check(L.pivamd_ensemble_extract(
            gauss.data_ptr(), status.data_ptr(), self.mask_u8.data_ptr(), n, h, w, self.peak_scale,
            *[o.data_ptr() for o in out],
            reason.data_ptr(), st), "ensemble_extract")
        if self.predictor is not None:
            # the fit measured the residual of the warped pairs
            out[0] += self.predictor[1].reshape(-1)
            out[1] += self.predictor[0].reshape(-1)
        shp = (self.ny, self.nx)
        return EnsembleResult(out[0].view(shp), out[1].view(shp), out[2].view(shp), out[3].view(shp),
                              out[4].view(shp), out[5].view(shp), reason.view(shp), gauss, status, init, diag,
                              tuple(p.view(n, h, w) for p in planes), self.n_images)


class EnsembleMultipass:
    """Multi-pass ensemble PIV over pairs resident on the device.

    ``window_sizes`` / ``overlaps`` give the passes and ``run_types`` (one
    per pass, default all 'standard') their run type; 'single' passes share
    ``sum_window`` / ``fit_window`` and keep their centres in original image
    coordinates, so the predictor chain is the same for both.  Pass p > 0
    warps every pair
    by the previous pass's field used as one shared predictor
    (``fused_symmetric_warp_batch(shared_predictor=1, round_shifts)``,
    ``cpu_ensemble.py:1445-1484``):

    * the field is edge-padded onto the previous pass's padded centres,
      Gaussian-smoothed when ``predictor_smoothing`` (the reference's
      ensemble default is off) and rounded to the nearest even integer when
      ``predictor_rounding`` (``cpu_ensemble.py:1534-1601``);
    * the cubic remap of it onto this pass's centres (rounded the same way,
      ``:1620-1640``) is the predictor recorded for the pass: it sets the
      coloured floor's per-window fractions and is added back to the fitted
      displacements;
    * failed fits (NaN) are infilled before they feed the next pass:
      ``mid_infill`` between passes (default nearest), ``final_infill`` on
      the last pass (default biharmonic, or 'none'),
      ``single_pass_accumulator.py:1330-1400``.
    """

    def __init__(self, image_shape, window_sizes, overlaps, *, window_type="square", bg_method="correlation",
                 kspace_floor="coloured", kspace_shape="gaussian", envelope_divide=False,
                 predictor_smoothing=False, predictor_rounding=False, warp_interpolation="cubic",
                 mid_infill="nearest", final_infill="biharmonic", run_types=None, sum_window=(32, 32),
                 fit_window=None, device=None):
        self.H, self.W = (int(image_shape[0]), int(image_shape[1]))
        self.window_sizes = [tuple(w) if np.ndim(w) else (w, w) for w in window_sizes]
        self.overlaps = list(overlaps)
        if len(self.overlaps) != len(self.window_sizes):
            raise ValueError("one overlap per pass")
        self.run_types = ["standard"] * len(self.window_sizes) if run_types is None else list(run_types)
        if len(self.run_types) != len(self.window_sizes):
            raise ValueError("one run type per pass")
        if mid_infill not in ("nearest", "harmonic", "biharmonic"):
            raise ValueError(f"mid-pass infill {mid_infill!r}: nearest, harmonic or biharmonic")
        if final_infill not in ("nearest", "harmonic", "biharmonic", "none"):
            raise ValueError(f"final-pass infill {final_infill!r}")
        if warp_interpolation not in ("cubic", "lanczos"):
            raise ValueError(f"warp_interpolation {warp_interpolation!r}: cubic or lanczos")
        self.grids = build_pass_grids((self.H, self.W), self.window_sizes, self.overlaps, self.run_types,
                                      sum_window)
        self.sum_window, self.fit_window = sum_window, fit_window
        self.opts = dict(window_type=window_type, bg_method=bg_method, kspace_floor=kspace_floor,
                         kspace_shape=kspace_shape, envelope_divide=envelope_divide,
                         warp_kernel=warp_interpolation)
        self.smoothing = bool(predictor_smoothing)
        self.rounding = bool(predictor_rounding)
        self.interp = 0 if warp_interpolation == "cubic" else 1
        self.mid_infill, self.final_infill = mid_infill, final_infill
        self.dev = torch.device(device) if device is not None else torch.device("cuda")
        self.L = _native.core()
        import pivtools_amd.multipass  # noqa: F401  (binds the glue entry points)

    def _st(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _predictor(self, p, field):
        """Padded (2, 1, nyp, nxp) shared predictor of pass p from the
        previous pass's field (2, 1, ny, nx), and its remap onto pass p's
        window centres (2, ny, nx)."""
        from .multipass import gaussian_weights_1d
        L, st, d = self.L, self._st(), self.dev
        gp, g = self.grids[p - 1], self.grids[p]
        ny, nx = gp.n_win
        nyp, nxp = len(gp.ctrs_y_all), len(gp.ctrs_x_all)
        pad = torch.empty((2, 1, nyp, nxp), dtype=torch.float32, device=d)
        _native.check(L.pivamd_pad_edge(field.data_ptr(), 1, ny, nx, gp.n_pre[0], gp.n_pre[1], nyp, nxp,
                                        pad.data_ptr(), st), "pad")
        if self.smoothing and g.sd > 0:
            wts, rad = gaussian_weights_1d(g.sd, (g.ksize[0] - 1) / (2 * g.sd))
            dw = torch.as_tensor(wts, device=d)
            tmp, sm = torch.empty_like(pad), torch.empty_like(pad)
            _native.check(L.pivamd_smooth_axis(pad.data_ptr(), 2, nyp, nxp, 0, dw.data_ptr(), rad, tmp.data_ptr(),
                                               st), "smooth")
            _native.check(L.pivamd_smooth_axis(tmp.data_ptr(), 2, nyp, nxp, 1, dw.data_ptr(), rad, sm.data_ptr(),
                                               st), "smooth")
            pad = sm
        if self.rounding:
            pad = 2.0 * torch.round(pad / 2.0)        # half to even, as np.round
        mx = torch.as_tensor(g.pred_map_x, device=d)
        my = torch.as_tensor(g.pred_map_y, device=d)
        nyn, nxn = g.n_win
        win = torch.empty((2, 1, nyn, nxn), dtype=torch.float32, device=d)
        _native.check(L.pivamd_remap_cubic(pad.data_ptr(), 1, nyp, nxp, mx.data_ptr(), my.data_ptr(), nyn, nxn,
                                           win.data_ptr(), st), "predictor remap")
        if self.rounding:
            win = 2.0 * torch.round(win / 2.0)
        return pad.contiguous(), win.reshape(2, nyn, nxn).contiguous()

    def _infill(self, p, res):
        method = self.final_infill if p == len(self.grids) - 1 else self.mid_infill
        g = self.grids[p]
        ny, nx = g.n_win
        u = torch.stack([res.uy, res.ux]).reshape(2, 1, ny, nx).contiguous()
        if method == "none" or not bool(torch.isnan(u).any()):
            return u
        L, st = self.L, self._st()
        out = torch.empty_like(u)
        err = torch.zeros(1, dtype=torch.int32, device=self.dev)
        if method == "nearest":
            _native.check(L.pivamd_infill_nearest(u.data_ptr(), None, 1, ny, nx, float(g.spacing[0]),
                                                  float(g.spacing[1]), out.data_ptr(), err.data_ptr(), st), "infill")
        else:
            wk = torch.empty(max(4, L.pivamd_infill_ls_workspace(1, ny, nx)), dtype=torch.uint8, device=self.dev)
            _native.check(L.pivamd_infill_ls(u.data_ptr(), None, 1, ny, nx, float(g.spacing[0]), float(g.spacing[1]),
                                             1 if method == "harmonic" else 2, out.data_ptr(), wk.data_ptr(),
                                             wk.numel(), err.data_ptr(), st), "infill")
        if int(err.item()) != 0:
            raise RuntimeError(f"ensemble pass {p + 1}: infill found no valid donor")
        res.uy.copy_(out[0, 0])
        res.ux.copy_(out[1, 0])
        return out

    def run(self, img_a: torch.Tensor, img_b: torch.Tensor, batch: int = 64) -> list:
        """All passes over the resident pairs (n, H, W) (float32 or 16-bit
        counts); returns one EnsembleResult per pass (displacements with the
        predictor added back, failed fits infilled)."""
        if img_a.shape != img_b.shape or img_a.dim() != 3 or tuple(img_a.shape[1:]) != (self.H, self.W):
            raise ValueError(f"expected two (n, {self.H}, {self.W}) stacks, got {tuple(img_a.shape)}")
        n = int(img_a.shape[0])
        u16 = img_a.dtype in (torch.int16, torch.uint16)
        L, d = self.L, self.dev
        warp_fn = L.pivamd_warp_u16 if u16 else L.pivamd_warp
        results, field = [], None
        wa = wb = None
        for p, g in enumerate(self.grids):
            single = self.run_types[p] == "single"
            ens = EnsemblePIV((self.H, self.W), g.win, self.overlaps[p], device=d, run_type=self.run_types[p],
                              sum_window=self.sum_window, fit_window=self.fit_window if single else None,
                              **self.opts)
            pad = None
            if p > 0:
                pad, win = self._predictor(p, field)
                ens.set_predictor(win)
                gp = self.grids[p - 1]
                cy = torch.as_tensor(gp.ctrs_y_all.astype(np.float32), device=d)
                cx = torch.as_tensor(gp.ctrs_x_all.astype(np.float32), device=d)
                if wa is None:
                    wa = torch.empty((min(batch, n), self.H, self.W), dtype=torch.float32, device=d)
                    wb = torch.empty_like(wa)
            for s in range(0, n, batch):
                a, b = img_a[s:s + batch].contiguous(), img_b[s:s + batch].contiguous()
                m = int(a.shape[0])
                if pad is not None:
                    _native.check(warp_fn(a.data_ptr(), b.data_ptr(), wa.data_ptr(), wb.data_ptr(),
                                          pad[0].data_ptr(), pad[1].data_ptr(), m, self.H, self.W, pad.shape[2],
                                          pad.shape[3], cy.data_ptr(), cx.data_ptr(), self.interp, 1,
                                          int(self.rounding), self._st()), "shared-predictor warp")
                    a, b = wa[:m], wb[:m]
                ens.accumulate(a, b)
            res = ens.finalize()
            field = self._infill(p, res)
            results.append(res)
        return results
