"""Autotuner for the correlation kernel's launch variants.

Every square window size has several builds of the correlation kernel in
the library (``pivamd_xcorr_variants``):

* ``R`` = 1, 2 or 4: the FP32 FFT form with each row/column transform
  shared by R lanes.  Larger R means more lanes (and waves) per window, so
  more latency hiding, at the cost of cross-lane exchange and twiddle work;
  it also changes the LDS tile layout and the windows per workgroup.
* ``16``: the split-precision MFMA DFT form (16, 32, 48 and 64 points),
  staged through LDS between its four products.
* ``96``, ``97``, ``98`` (16 points): the split-precision DFT as
  register-chained MFMA products, one wave per window, with fp16 hi/lo,
  bf16 hi/lo (about 16 significant bits: selectable, never tuned in) and
  bf16 hi/mid/lo operands.
* ``36``: the 4-lane FFT as a persistent kernel (64, 96 and 128 points, one
  window per workgroup) that loads the next window's samples while the
  current one is transformed.
* ``64 + R``: the R-lane FFT with one real LDS plane per window instead of
  the complex tile (``68``: 48, 64, 96 and 128 points; ``65``: sizes up to 32
  points, inverse stages packed over the workgroup's windows): the stage
  exchanges move real then imaginary parts, so more windows fit a CU.

Which one is fastest depends on the device and the size, so the library
lets the caller pick per size (``pivamd_xcorr_set_variant``, process-wide,
like the reference's peak-fit selector) and this module measures: for each
size it runs every variant on a synthetic batch that fills the GPU, checks
that the variant reproduces the default variant's displacements, and keeps
the fastest.  Results can be cached in a JSON file keyed by the device and
the library build, so a server pays the measurement once.

The reference has no counterpart (its C path has one build per size); the
variants all compute the same correlation, so a choice only moves rounding
at the float32 level (checked here to 1e-3 px, and by the GPU tests to the
parity tolerances).
"""

from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _native
from .window import compute_window_centers, window_weight

MFMA = 16
PERSIST = 32
HALF = 64
GEMM = 96                      # + 0 fp16, + 1 bf16, + 2 bf16 x3 (16 points)
MFMA_FORMS = (MFMA, GEMM, GEMM + 1, GEMM + 2)
_TARGET_WINDOWS = {8: 262144, 12: 262144, 16: 262144, 24: 196608, 32: 131072, 48: 65536, 64: 49152,
                   96: 20480, 128: 16384}


def variants(n: int) -> list[int]:
    """Variant codes built for window size n."""
    L = _native.core()
    buf = (ctypes.c_int * 16)()
    k = L.pivamd_xcorr_variants(n, buf, 16)
    return [int(buf[i]) for i in range(min(k, 16))]


def default_variant(n: int) -> int:
    return int(_native.core().pivamd_xcorr_default_variant(n))


def get_variant(n: int) -> int:
    return int(_native.core().pivamd_xcorr_get_variant(n))


def set_variant(n: int, v: int) -> int:
    """Select variant v for size n (0 = default); returns the previous one."""
    prev = int(_native.core().pivamd_xcorr_set_variant(n, v))
    if prev < 0:
        raise ValueError(f"variant {v} is not built for {n}x{n} windows (have {variants(n)})")
    return prev


def apply(table: dict) -> None:
    """Select the variants of a tuning table ({size: {"variant": v, ...}})."""
    for n, rec in table.items():
        set_variant(int(n), int(rec["variant"]))


class _Case:
    """A synthetic correlation call of one window size, on the device."""

    def __init__(self, n: int, device, shape=(1024, 1024), seed: int = 5):
        import torch

        self.torch = torch
        self.n = n
        self.dev = device
        wc = compute_window_centers(shape, (n, n), 50)
        self.ny, self.nx = wc.n_win_y, wc.n_win_x
        per_img = self.ny * self.nx
        self.n_img = int(max(1, min(64, -(-_TARGET_WINDOWS.get(n, 65536) // per_img))))
        g = torch.Generator(device="cpu")
        g.manual_seed(seed + n)
        # speckle-like frames: smoothed noise, B = A moved by (1.3, -0.7) px
        a = torch.rand((self.n_img, shape[0] + 8, shape[1] + 8), generator=g)
        a = torch.nn.functional.avg_pool2d(a[:, None], 3, 1, 1)[:, 0]
        self.A = a[:, 4:-4, 4:-4].contiguous().to(device)
        b = torch.roll(a, shifts=(1, -1), dims=(1, 2))
        self.B = (0.7 * b[:, 4:-4, 4:-4] + 0.3 * a[:, 4:-4, 4:-4]).contiguous().to(device)
        self.H, self.W = shape
        self.cx = torch.as_tensor(wc.win_ctrs_x, device=device)
        self.cy = torch.as_tensor(wc.win_ctrs_y, device=device)
        self.w = torch.as_tensor(window_weight((n, n), "gaussian"), device=device)
        self.pw = torch.empty((n, n), device=device)
        L = _native.core()
        st = torch.cuda.current_stream(device).cuda_stream
        _native.check(L.pivamd_plane_weight(self.w.data_ptr(), n, n, self.pw.data_ptr(), st), "plane weight")
        shp = (self.n_img, self.ny, self.nx)
        self.out = {k: torch.empty(shp, device=device) for k in ("pkx", "pky", "pkh", "sx", "sy", "sxy", "ratio")}
        ws = L.pivamd_xcorr_workspace(self.n_img, self.ny, self.nx, n, n, 1, 0)
        self.work = torch.empty(max(ws, 4), dtype=torch.uint8, device=device)

    @property
    def windows(self) -> int:
        return self.n_img * self.ny * self.nx

    def run(self) -> None:
        L = _native.core()
        o = self.out
        st = self.torch.cuda.current_stream(self.dev).cuda_stream
        rc = L.pivamd_xcorr(self.A.data_ptr(), self.B.data_ptr(), None, self.H, self.W, self.n_img,
                            self.cx.data_ptr(), self.cy.data_ptr(), self.ny, self.nx, self.w.data_ptr(),
                            self.w.data_ptr(), self.pw.data_ptr(), 0, self.n, self.n, 1, 3, 0,
                            o["pkx"].data_ptr(), o["pky"].data_ptr(), o["pkh"].data_ptr(), o["sx"].data_ptr(),
                            o["sy"].data_ptr(), o["sxy"].data_ptr(), o["ratio"].data_ptr(), None,
                            self.work.data_ptr(), self.work.numel(), st)
        _native.check(rc, "bulkxcorr2d")

    def result(self):
        return self.out["pkx"].cpu().numpy().copy(), self.out["pky"].cpu().numpy().copy()

    def time_ms(self, reps: int) -> float:
        torch = self.torch
        self.run()                                  # warm (first launch of this variant)
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            self.run()
            e1.record()
            ts.append((e0, e1))
        torch.cuda.synchronize(self.dev)
        return float(np.median([a.elapsed_time(b) for a, b in ts]))


def agree(ref, got, tol_px: float = 1e-3, max_flip: float = 1e-3) -> bool:
    """Displacements of two variants agree: NaN masks nearly identical and
    the common valid vectors within tol_px."""
    (rx, ry), (gx, gy) = ref, got
    rn, gn = np.isnan(rx), np.isnan(gx)
    if np.mean(rn != gn) > max_flip:
        return False
    both = ~rn & ~gn
    if not both.any():
        return True
    d = np.hypot(rx[both] - gx[both], ry[both] - gy[both])
    return bool(np.mean(d > tol_px) <= max_flip)


def _device_key(device) -> str:
    import torch

    p = torch.cuda.get_device_properties(device)
    lib = os.path.join(_native.LIB_DIR, "libpivtools_amd.so")
    st = os.stat(lib)
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}|{st.st_size}|{int(st.st_mtime)}"


def tune_xcorr(sizes, device=None, reps: int = 10, cache: str | None = None, include_mfma: bool = True,
               log=None, margin: float = 0.03) -> dict:
    """Measure every variant of each window size, select the fastest that
    agrees with the default (the default unless one is faster by more than
    ``margin``), and return {size: {"variant", "default", "ms",
    "windows"}} ("ms": median milliseconds per call, per variant).  With
    ``cache`` (a JSON path), a table stored for this device and library
    build is applied without measuring, and new results are written back."""
    import torch

    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    key = _device_key(dev)
    stored = {}
    if cache and os.path.isfile(cache):
        try:
            with open(cache) as f:
                stored = json.load(f)
        except (OSError, ValueError):
            stored = {}
    known = stored.get(key, {})
    table = {}
    for n in sorted({int(s) for s in sizes}):
        if str(n) in known:
            table[n] = known[str(n)]
            set_variant(n, int(table[n]["variant"]))
            continue
        cands = [v for v in variants(n) if v != GEMM + 1 and (include_mfma or v not in MFMA_FORMS)]
        dflt = default_variant(n)
        case = _Case(n, dev)
        prev = set_variant(n, dflt)
        try:
            case.run()
            torch.cuda.synchronize(dev)
            ref = case.result()
            ms = {}
            ok = {}
            for v in cands:
                set_variant(n, v)
                ms[v] = case.time_ms(reps)
                ok[v] = agree(ref, case.result())
        finally:
            set_variant(n, prev)
        good = [v for v in cands if ok[v]] or [dflt]
        best = min(good, key=lambda v: ms.get(v, float("inf")))
        # leave the default unless a variant is clearly faster: differences
        # within the timing noise would make the choice flip between runs
        if dflt in good and ms.get(best, 0.0) > (1.0 - margin) * ms.get(dflt, 0.0):
            best = dflt
        set_variant(n, best)
        table[n] = {"variant": best, "default": dflt, "windows": case.windows,
                    "ms": {str(v): round(t, 4) for v, t in ms.items()},
                    "rejected": [v for v in cands if not ok[v]]}
        if log:
            log(f"[autotune] {n}x{n}: {table[n]['ms']} ms -> variant {best}")
        del case
    if cache:
        stored[key] = {**known, **{str(k): v for k, v in table.items()}}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(stored, f, indent=1)
        os.replace(tmp, cache)
    return table


def get_warp_variant() -> int:
    return int(_native.core().pivamd_warp_get_variant())


def set_warp_variant(v: int) -> None:
    """Select the FMA warp's form: 1 stages each 64 x 32 tile's source
    rectangles in LDS (512-lane workgroups), 0 reads every tap from global
    memory (256-lane workgroups on 64 x 16 tiles)."""
    _native.check(_native.core().pivamd_warp_set_variant(int(v)), "warp variant")


def tune_warp(shape=(2160, 2560), n_pairs: int = 16, device=None, reps: int = 10, log=None,
              margin: float = 0.03) -> dict:
    """Time the FMA warp's two tile shapes (64 x 16 tiles reading global
    memory, 64 x 32 tiles staged in LDS) on a synthetic 16-bit batch with a
    smooth per-image predictor on a 64-pixel grid, check that they give the
    same frames, select the faster (the default unless the other is faster
    by more than ``margin``) and return {"variant", "default", "ms"}."""
    import torch

    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    L = _native.core()
    H, W = shape
    g = torch.Generator(device="cpu")
    g.manual_seed(11)
    frames = torch.randint(0, 4096, (2, n_pairs, H, W), generator=g, dtype=torch.int16).to(dev)
    cy = torch.arange(-64.0, H + 64.0, 64.0, device=dev)
    cx = torch.arange(-64.0, W + 64.0, 64.0, device=dev)
    yy, xx = torch.meshgrid(cy / H, cx / W, indexing="ij")
    k = torch.arange(n_pairs, device=dev, dtype=torch.float32)[:, None, None]
    pdy = (3.0 * torch.sin(6.0 * yy + k) + 1.0).contiguous()
    pdx = (4.0 * torch.cos(5.0 * xx - k) - 2.0).contiguous()
    outs = {v: (torch.empty((n_pairs, H, W), device=dev), torch.empty((n_pairs, H, W), device=dev)) for v in (0, 1)}
    st = torch.cuda.current_stream(dev).cuda_stream
    prev = get_warp_variant()
    ms = {}
    try:
        for v in (0, 1):
            set_warp_variant(v)
            oa, ob = outs[v]

            def run():
                _native.check(L.pivamd_warp_u16(frames[0].data_ptr(), frames[1].data_ptr(), oa.data_ptr(),
                                                ob.data_ptr(), pdy.data_ptr(), pdx.data_ptr(), n_pairs, H, W,
                                                len(cy), len(cx), cy.data_ptr(), cx.data_ptr(), 2, 0, 0, st), "warp")
            run()
            ts = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                run()
                e1.record()
                ts.append((e0, e1))
            torch.cuda.synchronize(dev)
            ms[v] = float(np.median([a.elapsed_time(b) for a, b in ts]))
    finally:
        set_warp_variant(prev)
    same = all(torch.equal(a, b) for a, b in zip(outs[0], outs[1]))
    dflt = 1
    best = min(ms, key=ms.get) if same else dflt
    if best != dflt and ms[best] > (1.0 - margin) * ms[dflt]:
        best = dflt
    set_warp_variant(best)
    rec = {"variant": best, "default": dflt, "ms": {str(v): round(t, 4) for v, t in ms.items()}, "equal": same}
    if log:
        log(f"[autotune] warp {H}x{W} x {n_pairs}: {rec['ms']} ms -> variant {best}")
    return rec
