"""Per-pair results on the host and the reference's output files.

``MultipassPIV.run`` returns per-pass device tensors batched over the pairs.
This module turns them into the reference's per-pair result objects
(pivtools_cli/piv/piv_result.py:7-46 PIVPassResult / PIVResult, filled at
pivtools_cli/piv/piv_backend/cpu_instantaneous.py:876-891) and writes the
reference's files from them:

* ``save_piv_result``: the instantaneous ``.mat`` file
  (pivtools_cli/piv/save_results.py:23-86 and :790-950): one struct
  ``piv_result`` with one element per pass; fields ux, uy, b_mask, nan_mask,
  win_ctrs_x, win_ctrs_y, peak_mag, peak_ratio, peak_choice, n_windows,
  predictor_field, window_size, nan_reason ('full') or ux, uy, b_mask
  ('minimal'); float fields stored as float16 clamped to +-65504 with NaN
  kept, except peak_ratio (float32) and nan_reason (int8); passes outside
  ``runs_to_save`` hold empty arrays.
* ``save_corr_plane_dump``: the correlation-plane debug dump
  (save_results.py:88-172), ``debug_corr_planes/<stem>_corrplanes.npz`` with
  per-pass ``pass{i}_`` keys.

Only the layout is the reference's; the file I/O is plain scipy / numpy.
"""

from __future__ import annotations

import json
import logging
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

# nan_reason codes (pivtools_cli/piv/piv_backend/nan_reason_codes.py:18-36)
NAN_REASON_LABELS = {
    -1: "masked window",
    0: "valid",
    1: "peak-fit failure (no peak / LM did not converge)",
    6: "large-displacement rejection",
    9: "unclassified (in nan_mask, no stage claimed it)",
    10: "outlier detection (primary peak)",
    12: "outlier detection (substituted secondary peak)",
    13: "chosen peak magnitude NaN (peaks exhausted)",
}

FLOAT16_MAX = 65504.0


@dataclass
class PIVPassResult:
    n_windows: np.ndarray | None = None
    ux_mat: np.ndarray | None = None
    uy_mat: np.ndarray | None = None
    nan_mask: np.ndarray | None = None
    peak_mag: np.ndarray | None = None
    peak_ratio: np.ndarray | None = None
    peak_choice: np.ndarray | None = None
    predictor_field: np.ndarray | None = None
    primary_peak_mag: np.ndarray | None = None
    b_mask: np.ndarray | None = None
    window_size: tuple | None = None
    win_ctrs_x: np.ndarray | None = None
    win_ctrs_y: np.ndarray | None = None
    nan_reason: np.ndarray | None = None
    debug_dump: dict | None = None


@dataclass
class PIVResult:
    passes: list = field(default_factory=list)

    def add_pass(self, p: PIVPassResult) -> None:
        self.passes.append(p)


def _np(t):
    return None if t is None else t.detach().cpu().numpy()


def pair_result(passes, k: int) -> PIVResult:
    """PIVResult of pair k from ``MultipassPIV.run(..., keep_passes=True)``
    (or its final pass alone)."""
    res = PIVResult()
    for p in passes:
        ny, nx = p.n_windows
        pred = _np(p.predictor)
        if pred is None:
            pf = np.zeros((ny, nx, 2), np.float32)              # first pass: no predictor
        else:
            pf = np.stack([pred[0, k], pred[1, k]], axis=-1)     # [uy, ux]
        dd = None
        if p.debug_dump is not None:
            dd = {key: (val if key == "peak_finder" else _np(val[k])) for key, val in p.debug_dump.items()}
        res.add_pass(PIVPassResult(
            n_windows=np.array([ny, nx], dtype=np.int32),
            ux_mat=_np(p.ux[k]).copy(), uy_mat=_np(p.uy[k]).copy(),
            nan_mask=_np(p.nan_mask[k]).astype(bool),
            nan_reason=_np(p.nan_reason[k]).astype(np.int8),
            peak_mag=_np(p.peak_mag[k]).copy(), peak_ratio=_np(p.peak_ratio[k]).copy(),
            peak_choice=None if p.peak_choice is None else _np(p.peak_choice[k]).astype(np.int64),
            predictor_field=pf,
            primary_peak_mag=None if p.primary_peak_mag is None else _np(p.primary_peak_mag[k]).copy(),
            b_mask=np.zeros((ny, nx), bool) if p.b_mask is None else np.asarray(p.b_mask, bool),
            window_size=np.asarray(p.window_size, dtype=np.int32),
            win_ctrs_x=np.asarray(p.win_ctrs_x), win_ctrs_y=np.asarray(p.win_ctrs_y),
            debug_dump=dd))
    return res


def to_half(arr: np.ndarray) -> np.ndarray:
    """float -> float16 clamped to +-65504, NaN kept (save_results.py:1066-1097)."""
    if arr is None or arr.size == 0 or arr.dtype.kind != "f":
        return arr
    fin = arr[np.isfinite(arr)]
    if fin.size and np.abs(fin).max() > FLOAT16_MAX:
        logging.warning("values beyond the float16 range are clamped to +-%g", FLOAT16_MAX)
    c = np.clip(arr, -FLOAT16_MAX, FLOAT16_MAX)
    c = np.where(np.isnan(arr), np.nan, c)
    return c.astype(np.float16)


FULL_FIELDS = ("ux", "uy", "b_mask", "nan_mask", "win_ctrs_x", "win_ctrs_y", "peak_mag", "peak_ratio",
               "peak_choice", "n_windows", "predictor_field", "window_size", "nan_reason")
MINIMAL_FIELDS = ("ux", "uy", "b_mask")


def piv_struct(result: PIVResult, runs_to_save=None, save_mode: str = "full") -> np.ndarray:
    """The struct array the reference saves as ``piv_result``."""
    names = MINIMAL_FIELDS if save_mode == "minimal" else FULL_FIELDS
    n = len(result.passes)
    st = np.empty((n,), dtype=[(f, object) for f in names])
    first = result.passes[0]
    dt = first.ux_mat.dtype if first.ux_mat is not None and first.ux_mat.size else np.float64
    empty = np.empty((0, 0), dtype=dt)
    for i in range(n):
        for f in names:
            st[f][i] = empty
    runs = list(range(n)) if runs_to_save is None else list(runs_to_save)
    for i, p in enumerate(result.passes):
        if i not in runs:
            continue
        st["ux"][i] = to_half(p.ux_mat)
        st["uy"][i] = to_half(p.uy_mat)
        st["b_mask"][i] = p.b_mask if p.b_mask is not None else p.nan_mask
        if save_mode == "minimal":
            continue
        st["nan_mask"][i] = p.nan_mask
        st["nan_reason"][i] = p.nan_reason
        st["win_ctrs_x"][i] = to_half(p.win_ctrs_x)
        st["win_ctrs_y"][i] = to_half(p.win_ctrs_y)
        st["peak_mag"][i] = to_half(p.peak_mag)
        st["peak_ratio"][i] = p.peak_ratio.astype(np.float32, copy=False)
        if p.peak_choice is not None:
            st["peak_choice"][i] = p.peak_choice
        st["n_windows"][i] = p.n_windows
        st["predictor_field"][i] = to_half(p.predictor_field)
        st["window_size"][i] = p.window_size
    return st


def save_piv_result(result: PIVResult, output_path, frame_number: int, runs_to_save=None,
                    vector_fmt: str = "B%05d.mat", save_mode: str = "full", do_compression: bool = True) -> str:
    """Write ``<output_path>/<vector_fmt % frame_number>`` as the reference does."""
    import scipy.io

    out = Path(output_path)
    out.mkdir(parents=True, exist_ok=True)
    fn = out / (vector_fmt % frame_number)
    if not result.passes:
        logging.warning("result of frame %d has no passes; nothing written", frame_number)
        return str(fn)
    scipy.io.savemat(fn, {"piv_result": piv_struct(result, runs_to_save, save_mode)}, oned_as="row",
                     do_compression=do_compression)
    return str(fn)


DUMP_KEYS = ("planes", "pk_loc_x_raw", "pk_loc_y_raw", "pk_height_raw", "pk_ratio_raw", "sx", "sy", "sxy",
             "peak_finder")


def save_corr_plane_dump(result: PIVResult, output_path, frame_number: int,
                         vector_fmt: str = "B%05d.mat") -> str | None:
    """Write the correlation-plane dump npz, or nothing (None) when no pass
    carries one."""
    if not any(p.debug_dump is not None for p in result.passes):
        return None
    d = Path(output_path) / "debug_corr_planes"
    d.mkdir(parents=True, exist_ok=True)
    fn = d / f"{Path(vector_fmt % frame_number).stem}_corrplanes.npz"
    payload = {
        "frame_number": np.int64(frame_number),
        "n_passes": np.int64(len(result.passes)),
        "nan_reason_labels": json.dumps({str(k): v for k, v in NAN_REASON_LABELS.items()}),
        "pk_origin_note": ("plane peak pixel = (pk_loc_y_raw + H//2, pk_loc_x_raw + W//2); planes are RAW "
                           "(what the peak search saw); all-zero plane = window skipped (masked/out of bounds)"),
    }
    for i, p in enumerate(result.passes):
        dd = p.debug_dump
        if dd is None:
            continue
        pre = f"pass{i}_"
        for key in DUMP_KEYS:
            payload[pre + key] = dd[key]
        payload[pre + "ux"] = p.ux_mat
        payload[pre + "uy"] = p.uy_mat
        payload[pre + "nan_mask"] = p.nan_mask
        payload[pre + "nan_reason"] = p.nan_reason
        payload[pre + "peak_choice"] = p.peak_choice
        payload[pre + "b_mask"] = p.b_mask
        payload[pre + "predictor"] = p.predictor_field
        payload[pre + "win_ctrs_x"] = p.win_ctrs_x
        payload[pre + "win_ctrs_y"] = p.win_ctrs_y
        payload[pre + "window_size"] = np.asarray(p.window_size)
    np.savez_compressed(fn, **payload)
    return str(fn)
