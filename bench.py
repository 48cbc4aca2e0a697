#!/usr/bin/env python
"""Throughput of the multipass instantaneous PIV hot path on one MI355X.

Workload (one step = one multipass over one batch): 64 seeded pairs of
2160 x 2560 12-bit images (uint16, Poisson shot noise, Lamb-Oseen vortex
rc 200 px / 8 px peak plus uniform 2 px), windows 128 -> 64 -> 32 -> 16 at
50 % overlap, gaussian window weight, bicubic image deformation between
passes, normalised-median outlier test, nearest infill, FP32 throughout,
3-point Gaussian sub-pixel fit.  Images are resident in HBM before timing.

Prints one JSON line: pairs/s, per-pair latency p50/p99 (every pair of a
batch completes with its batch), the roofline of the correlation kernel
measured with HIP events on its stream, and the CPU oracle timed on a
bounded sample of the same workload on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "python-pivtools_amd"))

SHAPE = (2160, 2560)
SCHEDULE = [(128, 128), (64, 64), (32, 32), (16, 16)]
OVERLAP = [50, 50, 50, 50]
PEAK_FP32_TFLOPS = 157.3
HBM_TBS = 8.0
# HBM bytes per correlation launch (xcorr_kernel + fit_kernel) of the 64-pair
# 2160x2560 workload, from rocprofv3 FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0)
# counters calibrated on a copy kernel of known size; recorded in
# profiles/hbm_counters_multipass_2160x2560_u16.md.
MEASURED_TRAFFIC_64_PAIRS = {128: 1.430e9 + 0.449e9 + 0.003e9 + 0.002e9,   # reads the 16-bit frames; writes are register spills
                             64: 2.811e9 + 0.012e9 + 0.011e9 + 0.008e9,
                             32: 2.832e9 + 0.049e9 + 0.044e9 + 0.033e9,
                             16: 2.832e9 + 0.198e9 + 0.176e9 + 0.132e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def f_corr(n):
    """Algorithmic flops per window: 2 r2c + 1 c2r (5 N^2 log2 N each),
    apodise / mean / energy (10 N^2), conj-multiply 6 N (N/2+1), scale N^2."""
    import math
    return 15 * n * n * math.log2(n) + 14 * n * n + 6 * n


def make_batch(n_pairs, device):
    import torch
    from pivtools_amd import synthetic as S
    A = torch.empty((n_pairs,) + SHAPE, dtype=torch.int16, device=device)
    B = torch.empty_like(A)
    for k in range(n_pairs):
        a, b = S.multipass_pair(seed=1000 + k, shape=SHAPE, device=device)
        A[k].copy_(a)
        B[k].copy_(b)
    return A, B


def cpu_baseline(A_dev, B_dev, gpu_final, finder, budget_s, loop_kw=None):
    """Oracle (C restatement, OpenMP) timed on a bounded sample of the same
    workload -- one full-size pair -- plus the accuracy of the device result
    for that pair against the FP64 NumPy oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import multipass as M
    import oracle as O
    from pivtools_amd.window import build_pass_grids

    O.build() if not os.path.isfile(os.path.join(ROOT, "oracle", "liboracle.so")) else None
    grids = build_pass_grids(SHAPE, SCHEDULE, OVERLAP)
    a = A_dev[0:1].cpu().numpy().astype(np.float32)
    b = B_dev[0:1].cpu().numpy().astype(np.float32)
    t0 = time.perf_counter()
    loop_kw = loop_kw or {}
    ref = M.run_multipass(a, b, grids, finder=finder, **loop_kw)[0]
    t_cpu = time.perf_counter() - t0
    out = {"value": round(1.0 / t_cpu, 4), "unit": "pairs/s", "cores": int(O.lib().orc_max_threads()),
           "kind": "port", "sample": "1 pair of the same workload (2160x2560, 4 passes), C oracle + NumPy glue",
           "seconds": round(t_cpu, 2)}
    gx, gy, gm = gpu_final
    rx, ry, rm = ref["ux"][0], ref["uy"][0], ref["nan_mask"][0]
    flips = float(np.mean(gm != rm))
    same = ~gm & ~rm
    rms_oracle = float(np.sqrt(np.mean((gx[same] - rx[same]) ** 2 + (gy[same] - ry[same]) ** 2)))
    acc = {"rms_px_vs_oracle": rms_oracle, "mask_flip_frac_vs_oracle": flips}
    if finder == 3 and budget_s > 0:
        t0 = time.perf_counter()
        r64 = M.run_multipass(a, b, grids, finder=3, flavour="fp64", **loop_kw)[0]
        fx, fy, fm = r64["ux"][0], r64["uy"][0], r64["nan_mask"][0]
        same = ~gm & ~fm
        acc["rms_px_vs_fp64"] = float(np.sqrt(np.mean((gx[same] - fx[same]) ** 2 + (gy[same] - fy[same]) ** 2)))
        acc["mask_flip_frac_vs_fp64"] = float(np.mean(gm != fm))
        acc["fp64_oracle_s"] = round(time.perf_counter() - t0, 2)
    cal = warp_calibration(a, b, grids)
    if cal is not None:
        out["calibration"] = cal
    return out, acc


def warp_calibration(a, b, grids, reps=3):
    """R = t(oracle) / t(reference C) for the one reference kernel of the
    path that compiles from its own sources: fused_symmetric_warp_batch
    (bicubic), timed on the same pair, predictor grid and thread count as
    the oracle's warp (median of `reps`).  The reference build (AVX2
    interior sampler, and its scalar form) is the checker's, made from the
    reference sources in the build container; None where it is absent.  The
    correlation (libbulkxcorr2d) is not built from the reference: its FFT
    includes generated codelet headers."""
    import ctypes
    import numpy as np
    import oracle as O

    here = os.path.join(ROOT, "oracle", "_ref")
    simd, scal = os.path.join(here, "libref_warp_simd.so"), os.path.join(here, "libref_warp.so")
    if not (os.path.isfile(simd) and os.path.isfile(scal)):
        return None
    g = grids[1]
    cy = np.ascontiguousarray(g.ctrs_y_all, np.float32)
    cx = np.ascontiguousarray(g.ctrs_x_all, np.float32)
    rng = np.random.default_rng(5)
    dy = rng.uniform(-2, 2, (1, len(cy), len(cx))).astype(np.float32)
    dx = rng.uniform(-2, 2, (1, len(cy), len(cx))).astype(np.float32)
    A = np.ascontiguousarray(a, np.float32)
    B = np.ascontiguousarray(b, np.float32)
    N, H, W = A.shape
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")

    def ref_ms(path, impl):
        lib = ctypes.CDLL(path)
        lib.fused_symmetric_warp_batch.argtypes = [f32p] * 6 + [ctypes.c_int] * 5 + [f32p, f32p] + \
            [ctypes.c_int] * 3
        lib.fused_warp_set_impl.argtypes = [ctypes.c_int]
        lib.fused_warp_set_impl(impl)
        oa, ob = np.zeros_like(A), np.zeros_like(B)
        ts = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            r = lib.fused_symmetric_warp_batch(A, B, oa, ob, dy, dx, N, H, W, len(cy), len(cx), cy, cx, 0, 0, 0)
            ts.append(time.perf_counter() - t0)
            if r != 0:
                raise RuntimeError(f"reference warp failed ({r})")
        return 1e3 * float(np.median(ts[1:])), oa

    def orc_ms():
        ts = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            oa, _ = O.fused_warp_batch(A, B, dy, dx, cy, cx, interp_mode=0)
            ts.append(time.perf_counter() - t0)
        return 1e3 * float(np.median(ts[1:])), oa

    t_simd, w_simd = ref_ms(simd, 1)
    t_scal, w_scal = ref_ms(scal, 0)
    t_orc, w_orc = orc_ms()
    return {"kernel": "fused_symmetric_warp_batch, bicubic, 1 pair of 2160x2560, pass-1 predictor grid",
            "threads": int(O.lib().orc_max_threads()),
            "reference_simd_ms": round(t_simd, 2), "reference_scalar_ms": round(t_scal, 2),
            "oracle_ms": round(t_orc, 2), "R_vs_simd": round(t_orc / t_simd, 3),
            "R_vs_scalar": round(t_orc / t_scal, 3),
            "oracle_equals_reference_scalar": bool(np.array_equal(w_orc, w_scal)),
            "max_abs_simd_vs_scalar": float(np.max(np.abs(w_simd - w_scal))),
            "correlation": "reference libbulkxcorr2d not built (generated codelet headers); no R"}


ENS_SHAPE = (1024, 1024)
# HBM bytes per 64-pair accumulate launch (1024x1024 f32, 15625 windows of
# 32x32): rocprofv3 FETCH_SIZE 539 MiB (x2.0) + WRITE_SIZE 183 MiB (x1.0),
# the copy-kernel calibration of the multipass counters; recorded in
# profiles/kernel_stats_ensemble_1024_w32_o75.md.
ENS_TRAFFIC_64_PAIRS = 1.130e9 + 0.192e9
ENS_WIN = (32, 32)


def f_ens(n):
    """Algorithmic flops per window and pair of the ensemble accumulation:
    2 r2c (5 N^2 log2 N each), weighting/mean (12 N^2) and the three spectral
    products AB, AA, BB (18 N (N/2+1))."""
    import math
    return 10 * n * n * math.log2(n) + 12 * n * n + 18 * n * (n // 2 + 1)


def ensemble_measure(args):
    """One timed ens_pairs-pair ensemble pass (1024 x 1024, 32 x 32 windows
    at 75 % overlap, correlation background, coloured floor, k-space fit)
    after a warmup pass: (ens, result, wall s, accumulate ms, finalize ms,
    distinct pairs, device pools)."""
    import torch
    from pivtools_amd import synthetic as S
    from pivtools_amd.ensemble import EnsemblePIV

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nd = args.ens_distinct
    log(f"[bench] generating {nd} distinct ensemble pairs of {ENS_SHAPE} f32 on the device")
    pool_a = torch.empty((nd,) + ENS_SHAPE, dtype=torch.float32, device=dev)
    pool_b = torch.empty_like(pool_a)
    for k in range(nd):
        a, b = S.ensemble_pair(seed=2000 + k, shape=ENS_SHAPE, device=dev)
        pool_a[k].copy_(a)
        pool_b[k].copy_(b)
    torch.cuda.synchronize()
    batch = args.pairs

    def one_pass(n_pairs, timed):
        ens = EnsemblePIV(ENS_SHAPE, ENS_WIN, 75, "square", bg_method="correlation", kspace_floor="coloured")
        ev = []
        done = 0
        while done < n_pairs:
            n = min(batch, n_pairs - done)
            i0 = done % nd
            n = min(n, nd - i0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            ens.accumulate(pool_a[i0:i0 + n], pool_b[i0:i0 + n])
            e1.record()
            ev.append((e0, e1, n))
            done += n
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        f0.record()
        res = ens.finalize()
        f1.record()
        return ens, res, ev, (f0, f1)

    log("[bench] ensemble warmup")
    one_pass(min(2 * batch, args.ens_pairs), False)
    torch.cuda.synchronize()
    log(f"[bench] timing one {args.ens_pairs}-pair ensemble pass")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens, res, ev, fin = one_pass(args.ens_pairs, True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    acc_ms = sum(a.elapsed_time(b) for a, b, _ in ev)
    fin_ms = fin[0].elapsed_time(fin[1])
    return ens, res, wall, acc_ms, fin_ms, nd, (pool_a, pool_b)


def ensemble_fit_stats(res):
    import numpy as np
    ok = res.nan_reason.cpu().numpy() == 0
    stats = {"valid_frac": round(float(ok.mean()), 4)}
    if ok.any():
        for k in ("UU", "VV", "UV"):
            stats[f"{k}_median_px2"] = round(float(np.median(getattr(res, k).cpu().numpy()[ok])), 4)
    return stats


def ensemble_leg(args):
    """The ensemble workload's pass, reported inside the default line."""
    ens, res, wall, acc_ms, fin_ms, nd, _ = ensemble_measure(args)
    achieved = f_ens(ENS_WIN[0]) * ens.nwin * args.ens_pairs / (acc_ms * 1e-3) / 1e12
    return {"workload": "ensemble_1024_w32_o75", "pairs": args.ens_pairs, "pairs_per_s": round(args.ens_pairs / wall, 3),
            "ms": round(wall * 1e3, 3), "accumulate_ms": round(acc_ms, 3), "finalize_ms": round(fin_ms, 3),
            "accumulate_tflops": round(achieved, 3), "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "fit": ensemble_fit_stats(res)}


def run_ensemble(args):
    """Ensemble workload: one 2000-pair pass at 1024 x 1024, 32 x 32 windows
    at 75 % overlap, correlation background, coloured floor, k-space fit."""
    ens, res, wall, acc_ms, fin_ms, nd, (pool_a, pool_b) = ensemble_measure(args)
    batch = args.pairs
    value = args.ens_pairs / wall
    nwin = ens.nwin
    flops = f_ens(ENS_WIN[0]) * nwin * args.ens_pairs
    achieved = flops / (acc_ms * 1e-3) / 1e12
    stats = ensemble_fit_stats(res)
    cpu = None
    if args.cpu_budget > 0:
        cpu = ensemble_cpu_baseline(pool_a, pool_b, ens, res, args.ens_pairs)
    line = {
        "metric": "ensemble pairs/sec (one 2000-pair pass incl. finalize and k-space fit), 1 GPU",
        "value": round(value, 3), "unit": "pairs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 accumulate in FP64 spectra, FP64 fit",
        "data": f"synthetic ({nd} distinct seeded pairs cycled to {args.ens_pairs})",
        "config": {"workload": "ensemble_1024_w32_o75", "pairs": args.ens_pairs, "window": list(ENS_WIN),
                   "overlap": 75, "windows": nwin, "bg_method": "correlation", "kspace_floor": "coloured",
                   "batch": batch},
        "accumulate_ms": round(acc_ms, 3), "finalize_ms": round(fin_ms, 3),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": ENS_TRAFFIC_64_PAIRS if batch == 64 else None,
                     "traffic_note": "HBM bytes of ONE 64-pair accumulate launch, a profile-derived "
                                     "estimate (FETCH_SIZE x2.0 + WRITE_SIZE of a separate rocprofv3 run), "
                                     "not measured in this run",
                     "kernel": "ens_kernel_reg<32,true> (accumulate)",
                     "note": "algorithmic flops 10N^2log2N + 12N^2 + 18N(N/2+1) per window and pair; "
                             "FP32 VALU work (register FFT), no MFMA"},
        "cpu_baseline": cpu, "fit": stats,
    }
    print(json.dumps(line), flush=True)


def ensemble_cpu_baseline(pool_a, pool_b, ens, res, n_pairs):
    """Oracle accumulate_triple timed on 4 pairs and the oracle finalize +
    k-space fit timed once, extrapolated to the n_pairs pass."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    a = pool_a[:4].cpu().numpy()
    b = pool_b[:4].cpu().numpy()
    g = ens.grid
    w = ens.w_host
    mask = np.zeros(ens.nwin, np.float32)
    t0 = time.perf_counter()
    O.accumulate_triple(a, b, mask, g.win_ctrs_x, g.win_ctrs_y, ENS_WIN, w, w, w, w)
    t_pair = (time.perf_counter() - t0) / 4
    planes = [p.cpu().numpy() for p in res.planes]
    pc = ens.floor_shape()
    t0 = time.perf_counter()
    O.kspace_fit(planes[0], planes[1], planes[2], np.zeros(ens.nwin, bool), ENS_WIN,
                 p_win=np.tile(pc.ravel(), ens.nwin))
    t_fit = time.perf_counter() - t0
    total = n_pairs * t_pair + t_fit
    return {"value": round(n_pairs / total, 4), "unit": "pairs/s", "cores": int(O.lib().orc_max_threads()),
            "kind": "port", "sample": f"accumulate_triple on 4 pairs ({t_pair:.3f} s/pair) + one k-space fit "
            f"of {ens.nwin} windows ({t_fit:.2f} s), extrapolated to {n_pairs} pairs"}


def run_serving(args):
    """Serving workload: Poisson-arriving pairs from 6 camera streams
    (1024^2, 2048^2, 2160x2560 uint16; 3- or 4-pass schedules; two streams
    also feed ensemble accumulators) through the continuous-batching server.
    One line per offered rate in the sweep."""
    import numpy as np
    import torch
    from pivtools_amd import synthetic as S
    from pivtools_amd.serving import PIVServer, StreamSpec

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    s3, s4 = [(64, 64), (32, 32), (16, 16)], [(128, 128), (64, 64), (32, 32), (16, 16)]
    specs = [StreamSpec("cam0", (1024, 1024), s3, [50] * 3, ensemble=True),
             StreamSpec("cam1", (1024, 1024), s4, [50] * 4, ensemble=True),
             StreamSpec("cam2", (2048, 2048), s3, [50] * 3),
             StreamSpec("cam3", (2048, 2048), s4, [50] * 4),
             StreamSpec("cam4", (2160, 2560), s3, [50] * 3),
             StreamSpec("cam5", (2160, 2560), s4, [50] * 4)]
    log("[bench] generating 4 pairs per camera stream")
    pool = []
    for i, sp in enumerate(specs):
        pool.append([S.multipass_pair(seed=4000 + 10 * i + k, shape=sp.shape, device=dev) for k in range(4)])
    torch.cuda.synchronize()
    srv = PIVServer(specs, max_batch=args.serve_batch, max_slots_per_stream=2,
                    hbm_budget_bytes=int(args.serve_hbm_gb * (1 << 30)), autotune=args.autotune)
    if srv.tuned:
        log(f"[bench] autotuned variants {({n: r['variant'] for n, r in srv.tuned.items()})}")
    rng = np.random.default_rng(12345)

    def drive(rate, duration):
        t0 = time.perf_counter()
        arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * duration * 1.5) + 10))
        arrivals = arrivals[arrivals < duration]
        which = rng.integers(0, len(specs), len(arrivals))
        reqs, k = [], 0
        srv.completed.clear()
        while k < len(arrivals) or srv.busy():
            now = time.perf_counter() - t0
            while k < len(arrivals) and arrivals[k] <= now:
                i = int(which[k])
                a, b = pool[i][k % 4]
                reqs.append(srv.submit(i, a, b, t_arrival=t0 + arrivals[k]))
                k += 1
            srv.poll()
        torch.cuda.synchronize()
        lat = np.array([r.latency for r in reqs]) * 1e3
        span = max(r.t_done for r in reqs) - t0
        return {"offered_pairs_s": rate, "pairs": len(reqs), "achieved_pairs_s": round(len(reqs) / span, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "max_ms": round(float(lat.max()), 2)}

    log("[bench] serving warmup (graph capture per stream, slot and batch bucket)")
    for rate in (200, 2000, 2000):
        drive(rate, 1.0)
    sweep = []
    for rate in args.serve_rates:
        r = drive(rate, args.serve_seconds)
        log(f"[bench] serving {r}")
        sweep.append(r)
    ens = srv.finalize_ensembles()
    top = sweep[-1]
    line = {
        "metric": "serving pairs/sec + p50/p99 pair latency, 6 Poisson camera streams, 1 GPU",
        "value": top["achieved_pairs_s"], "unit": "pairs/s", "n_gpus": 1, "steps": len(sweep), "warmup": 3,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (4 seeded uint16 pairs per stream, reused)",
        "config": {"workload": "serving_6streams", "streams": [[sp.name, list(sp.shape), [w[0] for w in sp.window_sizes],
                                                              sp.ensemble] for sp in specs],
                   "max_batch": args.serve_batch, "slots_per_stream": 2, "graphs": True,
                   "seconds_per_rate": args.serve_seconds},
        "latency_ms_p50": top["p50_ms"], "latency_ms_p99": top["p99_ms"], "sweep": sweep,
        "ensemble_pairs": {k: v.n_images for k, v in ens.items()},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


MIXED_SHAPES = ((1024, 1024), (2048, 2048))
MIXED_SCHEDULE = [(64, 64), (48, 48), (32, 32), (16, 16)]


def planes_fp64(img_a, img_b, ctrs_x, ctrs_y, n, weight):
    """FP64 correlation planes of every in-image window of one pair, for the
    accuracy report of the mixed workload (torch float64, any device).

    Windows follow the instantaneous rule of PIV_2d_cross_correlate.c:284-354:
    origin floor(c - (n-1)/2 + 0.5), weighted, minus the weighted window's
    mean; plane = fftshift(real(ifft2(FFT(b) conj(FFT(a))))), zero lag at
    [n/2, n/2].  Returns (planes (ny, nx, n, n), in-image mask (ny, nx))."""
    import numpy as np
    import torch

    H, W = img_a.shape
    dev = img_a.device
    r0 = np.floor(np.asarray(ctrs_y, np.float32).astype(np.float64) - (n - 1) / 2.0 + 0.5).astype(np.int64)
    c0 = np.floor(np.asarray(ctrs_x, np.float32).astype(np.float64) - (n - 1) / 2.0 + 0.5).astype(np.int64)
    ok = ((r0 >= 0) & (r0 + n <= H))[:, None] & ((c0 >= 0) & (c0 + n <= W))[None, :]
    rows = torch.as_tensor(np.clip(r0[:, None] + np.arange(n), 0, H - 1), device=dev)
    cols = torch.as_tensor(np.clip(c0[:, None] + np.arange(n), 0, W - 1), device=dev)
    w = torch.as_tensor(np.asarray(weight, np.float64), device=dev)
    ri, ci = rows[:, None, :, None], cols[None, :, None, :]

    def windows(img):
        x = img.to(torch.float64)[ri, ci] * w
        return x - x.mean(dim=(-2, -1), keepdim=True)

    fa = torch.fft.fft2(windows(img_a))
    fb = torch.fft.fft2(windows(img_b))
    planes = torch.fft.fftshift(torch.fft.ifft2(fb * fa.conj()).real, dim=(-2, -1))
    return planes, torch.as_tensor(ok, device=dev)


def _peak_locations(lib, planes, pw, finder):
    """(row, col) of the fitted peak of each plane through the library's own
    peak search and fit (pivamd_peakfit), NaN where the fit fails."""
    import torch

    n, h, w = planes.shape
    loc = torch.empty(3 * n, dtype=torch.float32, device=planes.device)
    std = torch.empty_like(loc)
    ratio = torch.empty(n, dtype=torch.float32, device=planes.device)
    st = torch.cuda.current_stream().cuda_stream
    rc = lib.pivamd_peakfit(planes.data_ptr(), n, h, w, 1, finder, 0, pw.data_ptr(), loc.data_ptr(),
                            std.data_ptr(), ratio.data_ptr(), st)
    if rc != 0:
        raise RuntimeError(f"pivamd_peakfit failed ({rc})")
    return loc[:n], loc[n:2 * n]


def mixed_accuracy(pairs, finder, log, forms=("fft", "split")):
    """Correlation-plane relative L2 error and sub-pixel peak RMS error of the
    FP32 FFT and the split-fp16 MFMA DFT against FP64 planes, per window size
    and image size, on one pair of each size (first pass: unwarped frames)."""
    import numpy as np
    import torch
    from pivtools_amd import _native
    from pivtools_amd.multipass import MultipassPIV
    from pivtools_amd.window import window_weight

    lib = _native.core()
    out = []
    for shape, (a, b) in pairs:
        for win in MIXED_SCHEDULE:
            n = win[0]
            got = {}
            for dft in forms:
                mp = MultipassPIV(shape, [win], [50], peak_finder=finder, dft=dft, dump_correlation_planes=True)
                res = mp.run(a[None], b[None], keep_passes=True)
                got[dft] = res[0].debug_dump["planes"][0]
                g, fid = mp.grids[0], mp.finder
            ref, ok = planes_fp64(a, b, g.ctrs_x, g.ctrs_y, n, window_weight(win, "gaussian"))
            okf = ok.reshape(-1)
            ref = ref.reshape(-1, n, n)[okf]
            wt = torch.as_tensor(window_weight(win, "gaussian"), device=a.device)
            pw = torch.empty(win, dtype=torch.float32, device=a.device)
            _native.check(lib.pivamd_plane_weight(wt.data_ptr(), n, n, pw.data_ptr(),
                                                  torch.cuda.current_stream().cuda_stream), "plane weight")
            ry, rx = _peak_locations(lib, ref.to(torch.float32).contiguous(), pw, fid)
            row = {"image": f"{shape[0]}x{shape[1]}", "window": n, "windows": int(okf.sum())}
            for dft, pl in got.items():
                pl = pl.reshape(-1, n, n)[okf]
                l2 = float(torch.linalg.norm(pl.double() - ref) / torch.linalg.norm(ref))
                py, px = _peak_locations(lib, pl.contiguous(), pw, fid)
                both = torch.isfinite(py) & torch.isfinite(ry)
                d2 = (py - ry)[both].double() ** 2 + (px - rx)[both].double() ** 2
                row[dft] = {"plane_rel_l2": float(f"{l2:.3e}"),
                            "peak_rms_px": float(f"{float(torch.sqrt(d2.mean())):.3e}"),
                            "peak_max_px": float(f"{float(torch.sqrt(d2.max())):.3e}"),
                            "peaks_off_by_1e-3px": int((d2 > 1e-6).sum()),
                            "fit_nan_mismatch": int((torch.isfinite(py) != torch.isfinite(ry)).sum())}
            log(f"[bench] accuracy {row}")
            out.append(row)
    return out


def mixed_groups(args):
    """The mixed workload's frames: mixed_pairs uint16 pairs split over the
    image sizes (the multipass physics, seeds 3000+)."""
    import torch
    from pivtools_amd import synthetic as S

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    per = args.mixed_pairs // len(MIXED_SHAPES)
    groups = []
    for gi, shape in enumerate(MIXED_SHAPES):
        log(f"[bench] generating {per} pairs of {shape} uint16 on the device")
        A = torch.empty((per,) + shape, dtype=torch.int16, device=dev)
        B = torch.empty_like(A)
        for k in range(per):
            a, b = S.multipass_pair(seed=3000 + gi * per + k, shape=shape, device=dev)
            A[k].copy_(a)
            B[k].copy_(b)
        groups.append((shape, A, B))
    torch.cuda.synchronize()
    return groups


def mixed_time(groups, args, dft):
    """warmup + steps multipasses of every group with one correlation form:
    (leg dict, last results per group)."""
    import numpy as np
    import torch
    from pivtools_amd.multipass import MultipassPIV

    mps = [MultipassPIV(shape, MIXED_SCHEDULE, [50] * 4, peak_finder=args.finder, dft=dft,
                        mid_infill=args.infill, final_infill=args.infill, warp_arithmetic=args.warp)
           for shape, _, _ in groups]
    log(f"[bench] {dft}: warmup")
    for _ in range(args.warmup):
        for mp, (_, A, B) in zip(mps, groups):
            mp.run(A, B)
    torch.cuda.synchronize()
    for mp in mps:
        mp.check_infill()
    log(f"[bench] {dft}: timing {args.steps} steps")
    batch_ev, pass_ev = [], [[] for _ in mps]
    finals = [None] * len(mps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for gi, (mp, (_, A, B)) in enumerate(zip(mps, groups)):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            res = mp.run(A, B, events=pass_ev[gi])
            e1.record()
            batch_ev.append((gi, e0, e1))
            finals[gi] = res[-1]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = []
    for gi, e0, e1 in batch_ev:
        lat += [e0.elapsed_time(e1)] * groups[gi][1].shape[0]
    passes = []
    for gi, mp in enumerate(mps):
        per_pass = {}
        for p, (a, b) in pass_ev[gi]:
            per_pass.setdefault(p, []).append(a.elapsed_time(b))
        for p, g in enumerate(mp.grids):
            ms = float(np.mean(per_pass[p]))
            nwin = groups[gi][1].shape[0] * g.n_win[0] * g.n_win[1]
            passes.append({"image": f"{groups[gi][0][0]}x{groups[gi][0][1]}", "window": g.win[0],
                           "windows": nwin, "ms": round(ms, 4),
                           "tflops": round(f_corr(g.win[0]) * nwin / (ms * 1e-3) / 1e12, 3)})
    total = sum(int(A.shape[0]) for _, A, _ in groups) * args.steps
    leg = {"pairs_per_s": round(total / wall, 3), "ms_per_step": round(wall * 1e3 / args.steps, 3),
           "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
           "latency_ms_p99": round(float(np.percentile(lat, 99)), 3), "passes": passes}
    log(f"[bench] {dft}: {leg['pairs_per_s']} pairs/s, {leg['ms_per_step']} ms/step")
    del mps
    torch.cuda.empty_cache()
    return leg, finals


def run_mixed(args):
    """Mixed workload: pairs of 1024 x 1024 and 2048 x 2048 uint16 frames
    (the multipass physics), ragged windows 64 -> 48 -> 32 -> 16 at 50 %
    overlap, timed with the FP32 FFT (autotuned variants) and with the
    split-fp16 MFMA DFT, plus FP64 accuracy of both forms."""
    import numpy as np
    import torch
    from pivtools_amd.multipass import MultipassPIV

    groups = mixed_groups(args)
    tuned = None
    if args.autotune:
        from pivtools_amd import autotune
        tuned = autotune.tune_xcorr([w[0] for w in MIXED_SCHEDULE], reps=10, log=log)

    legs = {}
    finals = {}
    if "fft" not in args.mixed_dft:
        args.mixed_dft = ["fft"] + list(args.mixed_dft)
    for dft in args.mixed_dft:
        legs[dft], finals[dft] = mixed_time(groups, args, dft)

    # both image sizes through one ragged correlation launch per pass
    from pivtools_amd.multipass import run_ragged
    mps = [MultipassPIV(shape, MIXED_SCHEDULE, [50] * 4, peak_finder=args.finder, mid_infill=args.infill,
                        final_infill=args.infill, warp_arithmetic=args.warp) for shape, _, _ in groups]
    items = [(mp, A, B) for mp, (_, A, B) in zip(mps, groups)]
    log("[bench] ragged: warmup")
    for _ in range(args.warmup):
        run_ragged(items)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rres = run_ragged(items)
    torch.cuda.synchronize()
    rwall = time.perf_counter() - t0
    exact = all(bool(torch.equal(r[-1].ux.nan_to_num(7.0), f.ux.nan_to_num(7.0))) and
                bool(torch.equal(r[-1].nan_mask, f.nan_mask)) for r, f in zip(rres, finals["fft"]))
    legs["ragged"] = {"pairs_per_s": round(args.mixed_pairs * args.steps / rwall, 3),
                      "ms_per_step": round(rwall * 1e3 / args.steps, 3), "equals_separate_runs": exact}
    # small batches (two pairs per image size): where one launch per pass matters
    small = [(mp, A[:2], B[:2]) for mp, (_, A, B) in zip(mps, groups)]
    for _ in range(2):
        run_ragged(small)
        for mp, a, b in small:
            mp.run(a, b)
    torch.cuda.synchronize()
    tt = {}
    for name in ("separate", "ragged"):
        t0 = time.perf_counter()
        for _ in range(20):
            if name == "ragged":
                run_ragged(small)
            else:
                for mp, a, b in small:
                    mp.run(a, b)
        torch.cuda.synchronize()
        tt[name] = (time.perf_counter() - t0) / 20 * 1e3
    legs["ragged"]["two_pairs_per_size_ms"] = {k: round(v, 3) for k, v in tt.items()}
    log(f"[bench] ragged: {legs['ragged']}")
    del mps, items
    torch.cuda.empty_cache()

    # the whole loop: final-pass vectors of each MFMA form against the FFT form
    loop = []
    for gi, (shape, _, _) in enumerate(groups):
        row = {"image": f"{shape[0]}x{shape[1]}"}
        for form in args.mixed_dft:
            if form == "fft":
                continue
            f, s = finals["fft"][gi], finals[form][gi]
            ok = (f.nan_mask == 0) & (s.nan_mask == 0)
            d2 = (f.ux - s.ux)[ok].double() ** 2 + (f.uy - s.uy)[ok].double() ** 2
            row["vectors"] = int(ok.sum())
            row[f"{form}_vs_fft_rms_px"] = float(f"{float(torch.sqrt(d2.mean())):.3e}")
        loop.append(row)
    log(f"[bench] loop forms vs fft: {loop}")
    acc = mixed_accuracy([(shape, (A[0], B[0])) for shape, A, B in groups], args.finder, log, args.mixed_dft)

    fft = legs["fft"]
    dom = max(fft["passes"], key=lambda d: d["ms"])
    achieved = f_corr(dom["window"]) * dom["windows"] / (dom["ms"] * 1e-3) / 1e12
    line = {
        "metric": "image pairs/sec + p50/p99 pair latency at <=0.01 px RMS error, 1 GPU",
        "value": fft["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fft["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "mixed_1024_2048_u16", "pairs_per_step": args.mixed_pairs,
                   "images": [f"{s[0]}x{s[1]}" for s in MIXED_SHAPES], "schedule": [w[0] for w in MIXED_SCHEDULE],
                   "overlap": 50, "peak_finder": args.finder, "infill": args.infill,
                   "warp": "bicubic" + (" (separable FMA)" if args.warp == "fma" else " (reference arithmetic)"),
                   "global_batch": args.mixed_pairs, "parallelism": "single-gpu",
                   "xcorr_variants": ({str(n): r["variant"] for n, r in tuned.items()} if tuned else "default")},
        "latency_ms_p50": fft["latency_ms_p50"],
        "latency_ms_p99": fft["latency_ms_p99"],
        "fft": fft,
        **{form: legs[form] for form in args.mixed_dft if form != "fft"},
        "ragged": legs["ragged"],
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "kernel": f"{_xcorr_kernel_name(dom['window'], 'fft')} + fit_kernel ({dom['image']})",
                     "note": "FP32 VALU FFT; peak = MI355X FP32 157.3 TFLOP/s"},
        "cpu_baseline": None,
        "accuracy_vs_fp64": acc,
        "loop_split_vs_fft": loop,
        "autotune": ({str(n): r for n, r in tuned.items()} if tuned else None),
    }
    print(json.dumps(line), flush=True)


def serving_leg(A, B, piv_options, rate, seconds, max_batch, one_pair_rate):
    """The same workload through the continuous-batching server: two camera
    streams of 2160 x 2560 pairs with the 128 -> 16 schedule (captured
    graphs, several slots per stream).  Reports per-request latency at one
    Poisson offered rate and the saturated rate with every request queued at
    once; the latter over the one-pair-per-step rate is the serving form's
    gain over feeding the pass loop one pair at a time."""
    import numpy as np
    import torch
    from pivtools_amd.serving import PIVServer, StreamSpec

    spec = dict(shape=SHAPE, window_sizes=SCHEDULE, overlaps=OVERLAP, piv_options=dict(piv_options))
    specs = [StreamSpec("cam_a", **spec), StreamSpec("cam_b", **spec)]
    srv = PIVServer(specs, max_batch=max_batch, max_slots_per_stream=3)
    rng = np.random.default_rng(777)
    n_pool = A.shape[0]

    def burst(n):
        srv.completed.clear()
        t0 = time.perf_counter()
        reqs = [srv.submit(k % 2, A[k % n_pool], B[k % n_pool], t_arrival=t0) for k in range(n)]
        srv.drain()
        torch.cuda.synchronize()
        return reqs, max(r.t_done for r in reqs) - t0

    def poisson(r, dur):
        srv.completed.clear()
        arr = np.cumsum(rng.exponential(1.0 / r, int(r * dur * 1.5) + 10))
        arr = arr[arr < dur]
        reqs, k = [], 0
        t0 = time.perf_counter()
        while k < len(arr) or srv.busy():
            now = time.perf_counter() - t0
            while k < len(arr) and arr[k] <= now:
                reqs.append(srv.submit(k % 2, A[k % n_pool], B[k % n_pool], t_arrival=t0 + arr[k]))
                k += 1
            srv.poll()
        torch.cuda.synchronize()
        lat = np.array([q.latency for q in reqs]) * 1e3
        return {"offered_pairs_s": r, "pairs": len(reqs),
                "achieved_pairs_s": round(len(reqs) / (max(q.t_done for q in reqs) - t0), 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3)}

    log("[bench] serving leg: warmup (graph capture per slot and batch size)")
    for _ in range(3):
        burst(4 * max_batch)
    poisson(rate, 0.5)
    reqs, span = burst(8 * n_pool)
    sat = len(reqs) / span
    lat = poisson(rate, seconds)
    out = {"streams": 2, "shape": list(SHAPE), "schedule": [w[0] for w in SCHEDULE], "max_batch": max_batch,
           "slots_per_stream": 3, "graphs": True, "saturated_pairs_s": round(sat, 2),
           "saturated_pairs": len(reqs), "at_rate": lat, "stats": dict(srv.stats)}
    if one_pair_rate:
        out["serving_over_one_pair"] = round(sat / one_pair_rate, 3)
    log(f"[bench] serving leg: {out}")
    return out


def host_fed(mp, A, B, steps: int, fmt: str) -> dict:
    """Pairs/s when the batch starts in pinned host memory: per step, the
    frames (16-bit) or Mono12p payloads cross PCIe on a copy stream, double
    buffered so the next batch's copy overlaps this batch's pass loop; Mono12p
    payloads are expanded on the device before the pass loop."""
    import numpy as np
    import torch
    from pivtools_amd import ingest

    n, H, W = A.shape
    dev = A.device
    if fmt == "mono12p":
        def pack(x):     # 12-bit counts -> LSB-first pairs of pixels in three bytes
            v = x.reshape(n, -1).to(torch.int32)
            p0, p1 = v[:, 0::2], v[:, 1::2]
            out = torch.stack([p0 & 0xFF, ((p0 >> 8) & 0xF) | ((p1 & 0xF) << 4), p1 >> 4], dim=-1)
            return out.to(torch.uint8).reshape(n, -1)
        host = [pack(A).cpu().pin_memory(), pack(B).cpu().pin_memory()]
        frames = [torch.empty_like(A), torch.empty_like(B)]
    else:
        host = [A.cpu().pin_memory(), B.cpu().pin_memory()]
    raw = [[torch.empty_like(h, device=dev) for h in host] for _ in range(2)]
    # one copy stream per frame of the pair, so the two copies can use two DMA engines
    copy_st = [torch.cuda.Stream(device=dev) for _ in host]
    comp = torch.cuda.current_stream(dev)
    copied, consumed = [[None, None], [None, None]], [None, None]

    def step(s):
        k = s % 2
        for j, (d, h) in enumerate(zip(raw[k], host)):
            with torch.cuda.stream(copy_st[j]):
                if consumed[k] is not None:
                    copy_st[j].wait_event(consumed[k])
                d.copy_(h, non_blocking=True)
                copied[k][j] = torch.cuda.Event()
                copied[k][j].record(copy_st[j])
        for j in range(len(host)):
            comp.wait_event(copied[k][j])
        if fmt == "mono12p":
            for d, f in zip(raw[k], frames):
                ingest.decode("mono12p", d, H, W, n_frames=n, out=f)
            ev = torch.cuda.Event()
            ev.record(comp)
            consumed[k] = ev
            mp.run(frames[0], frames[1])
        else:
            mp.run(raw[k][0], raw[k][1])
            ev = torch.cuda.Event()
            ev.record(comp)
            consumed[k] = ev

    for s in range(2):
        step(s)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for s in range(steps):
        step(s)
    torch.cuda.synchronize(dev)
    dt = time.perf_counter() - t0
    per_pair = sum(h.numel() * h.element_size() for h in host) / n
    # the parts alone: the batch's copies (both streams), and the decode
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    for j, (d, h) in enumerate(zip(raw[0], host)):
        with torch.cuda.stream(copy_st[j]):
            d.copy_(h, non_blocking=True)
    torch.cuda.synchronize(dev)
    copy_ms = (time.perf_counter() - t1) * 1e3
    dec_ms = None
    if fmt == "mono12p":
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for d, f in zip(raw[0], frames):
            ingest.decode("mono12p", d, H, W, n_frames=n, out=f)
        e1.record()
        torch.cuda.synchronize(dev)
        dec_ms = round(e0.elapsed_time(e1), 3)
    return {"pairs_per_s": round(n * steps / dt, 3), "h2d_bytes_per_pair": int(per_pair),
            "h2d_GB_per_s": round(per_pair * n * steps / dt / 1e9, 2), "steps": steps,
            "batch_copy_alone_ms": round(copy_ms, 3),
            "copy_alone_GB_per_s": round(per_pair * n / (copy_ms * 1e-3) / 1e9, 2), "decode_ms": dec_ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--finder", default="gauss3")
    ap.add_argument("--impl", default="scalar")
    ap.add_argument("--dft", choices=["fft", "split", "split-bf16", "split-bf16x3"], default="fft",
                    help="split: split-fp16 MFMA DFT for the 16/32/48/64 passes; split-bf16(x3): split-bf16 "
                         "MFMA DFT for the 16-point pass")
    ap.add_argument("--cpu-budget", type=float, default=60.0, help="0 disables the CPU baseline leg")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default correlation-kernel variants instead of timing them first")
    ap.add_argument("--one-pair-steps", type=int, default=10,
                    help="after the timed region, time the same loop fed one pair per step (0 disables)")
    ap.add_argument("--warp", choices=["exact", "fma"], default="fma",
                    help="warp arithmetic: the reference's bit for bit, or separable fused multiply-adds")
    ap.add_argument("--infill", choices=["nearest", "harmonic", "biharmonic"], default="nearest",
                    help="infill of invalid vectors, mid-pass and final")
    ap.add_argument("--peak-mag", type=float, default=None,
                    help="add the peak-magnitude outlier test with this threshold")
    ap.add_argument("--host-fed-steps", type=int, default=4,
                    help="after the timed region, time the loop fed from pinned host memory, 16-bit frames "
                         "and Mono12p payloads (0 disables); reported beside value, never as value")
    ap.add_argument("--workload", choices=["multipass", "ensemble", "serving", "mixed"], default="multipass")
    ap.add_argument("--mixed-dft", nargs="+", default=["fft", "split"],
                    choices=["fft", "split", "split-bf16", "split-bf16x3"],
                    help="correlation forms the mixed workload times (the first is the reported one)")
    ap.add_argument("--mixed-pairs", type=int, default=256,
                    help="mixed workload: pairs per step, half 1024x1024 and half 2048x2048")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[50, 200, 500, 1000, 2000])
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-batch", type=int, default=8)
    ap.add_argument("--serve-leg-seconds", type=float, default=2.0,
                    help="multipass workload: seconds of the Poisson serving leg (0 skips the serving leg)")
    ap.add_argument("--serve-leg-rate", type=float, default=1000.0,
                    help="multipass workload: offered pairs/s of the serving leg")
    ap.add_argument("--serve-hbm-gb", type=float, default=64.0)
    ap.add_argument("--ens-pairs", type=int, default=2000)
    ap.add_argument("--mixed-leg-steps", type=int, default=3,
                    help="multipass workload: timed steps of the mixed-size leg reported beside value (0 skips it)")
    ap.add_argument("--ensemble-leg-pairs", type=int, default=2000,
                    help="multipass workload: pairs of the ensemble pass reported beside value (0 skips it)")
    ap.add_argument("--ens-distinct", type=int, default=64)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (byte-counter calibration for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent: N GPUs would each run this batch on their own stream of pairs)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        run_ensemble(args)
        return
    if args.workload == "serving":
        run_serving(args)
        return
    if args.workload == "mixed":
        run_mixed(args)
        return

    import numpy as np
    import torch
    from pivtools_amd.multipass import MultipassPIV

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log(f"[bench] generating {args.pairs} pairs of {SHAPE} uint16 on the device")
    A, B = make_batch(args.pairs, dev)
    torch.cuda.synchronize()
    methods = None
    if args.peak_mag is not None:
        methods = [{"type": "peak_mag", "threshold": args.peak_mag},
                   {"type": "median_2d", "epsilon": 0.1, "threshold": 2, "size": 5}]
    mp = MultipassPIV(SHAPE, SCHEDULE, OVERLAP, peak_finder=args.finder, peak_fit_impl=args.impl, dft=args.dft,
                      outlier_methods=methods, mid_infill=args.infill, final_infill=args.infill,
                      warp_arithmetic=args.warp)
    tuned = None
    if args.autotune and args.dft == "fft":
        from pivtools_amd import autotune
        tuned = autotune.tune_xcorr([w[0] for w in SCHEDULE], reps=10, log=log)
        tuned["warp"] = autotune.tune_warp(SHAPE, n_pairs=16, reps=10, log=log)
    log("[bench] warmup")
    for _ in range(args.warmup):
        mp.run(A, B)
    torch.cuda.synchronize()
    mp.check_infill()

    if args.calibrate_copy:
        from pivtools_amd import _native
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        _native.core().pivamd_copy_f32(src.data_ptr(), dst.data_ptr(), src.numel(),
                                       torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        del src, dst
    log(f"[bench] timing {args.steps} steps")
    step_ms = []
    events = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        res = mp.run(A, B, events=events)
        e1.record()
        step_ms.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    steps_ms = [a.elapsed_time(b) for a, b in step_ms]
    total_pairs = args.pairs * args.steps
    value = total_pairs / wall

    # correlation-kernel roofline per pass (algorithmic flops / event time)
    per_pass = {}
    for p, (a, b) in events:
        per_pass.setdefault(p, []).append(a.elapsed_time(b))
    passes = []
    for p, g in enumerate(mp.grids):
        ms = float(np.mean(per_pass[p]))
        nwin = args.pairs * g.n_win[0] * g.n_win[1]
        fl = f_corr(g.win[0]) * nwin
        passes.append({"pass": p, "window": g.win[0], "windows": nwin, "ms": round(ms, 4),
                       "tflops": round(fl / (ms * 1e-3) / 1e12, 3)})
    dom = max(passes, key=lambda d: d["ms"])
    dom_flops = f_corr(dom["window"]) * dom["windows"]
    achieved = dom_flops / (dom["ms"] * 1e-3) / 1e12
    traffic = None
    if args.pairs == 64 and dom["window"] in MEASURED_TRAFFIC_64_PAIRS:
        traffic = MEASURED_TRAFFIC_64_PAIRS[dom["window"]]

    log(f"[bench] {value:.2f} pairs/s, {wall * 1e3 / args.steps:.1f} ms/step, passes {passes}")
    res_final = res[-1]
    gx = res_final.ux[0].cpu().numpy()
    gy = res_final.uy[0].cpu().numpy()
    gm = res_final.nan_mask[0].cpu().numpy().astype(bool)

    # the same pass loop fed one pair per step (no batching), after the timed region
    one = None
    if args.one_pair_steps > 0:
        a1, b1 = A[:1].contiguous(), B[:1].contiguous()
        for _ in range(2):
            mp.run(a1, b1)
        torch.cuda.synchronize()
        t1 = []
        for _ in range(args.one_pair_steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mp.run(a1, b1)
            torch.cuda.synchronize()
            t1.append((time.perf_counter() - t0) * 1e3)
        rate = 1e3 / float(np.mean(t1))
        one = {"pairs_per_s": round(rate, 3), "latency_ms_p50": round(float(np.percentile(t1, 50)), 3),
               "latency_ms_p99": round(float(np.percentile(t1, 99)), 3), "steps": args.one_pair_steps,
               "batched_over_one_pair": round(value / rate, 3)}
        log(f"[bench] one pair per step: {rate:.1f} pairs/s (batched / one-pair = {value / rate:.2f})")
    serve = None
    if args.serve_leg_seconds > 0:
        serve = serving_leg(A, B, {"outlier_methods": methods, "mid_infill": args.infill,
                                   "final_infill": args.infill,
                                   "warp_arithmetic": args.warp}, args.serve_leg_rate, args.serve_leg_seconds,
                            args.serve_batch, one["pairs_per_s"] if one else None)
    fed = None
    if args.host_fed_steps > 0:
        fed = {f: host_fed(mp, A, B, args.host_fed_steps, f) for f in ("u16", "mono12p")}
        log(f"[bench] host-fed (PCIe-inclusive): {fed}")
    ens_leg = None
    if args.ensemble_leg_pairs > 0:
        log(f"[bench] ensemble leg: one {args.ensemble_leg_pairs}-pair pass")
        ens_args = argparse.Namespace(**vars(args))
        ens_args.ens_pairs = args.ensemble_leg_pairs
        ens_leg = ensemble_leg(ens_args)
        log(f"[bench] ensemble leg: {ens_leg}")
    mixed_leg = None
    if args.mixed_leg_steps > 0:
        log(f"[bench] mixed leg: {args.mixed_pairs} pairs over {MIXED_SHAPES}, windows "
            f"{[w[0] for w in MIXED_SCHEDULE]}, FP32 FFT")
        mx_args = argparse.Namespace(**vars(args))
        mx_args.steps, mx_args.warmup = args.mixed_leg_steps, 1
        groups = mixed_groups(mx_args)
        mixed_leg, _ = mixed_time(groups, mx_args, "fft")
        mixed_leg = {"workload": "mixed_1024_2048_u16_w64-16", "pairs_per_step": args.mixed_pairs,
                     "steps": args.mixed_leg_steps, **mixed_leg}
        del groups
        torch.cuda.empty_cache()
    cpu = None
    acc = {}
    if args.cpu_budget > 0:
        log("[bench] CPU oracle leg on one pair")
        cpu, acc = cpu_baseline(A, B, (gx, gy, gm), mp.finder, args.cpu_budget,
                                {"outlier_methods": methods, "mid_infill": args.infill, "final_infill": args.infill})

    line = {
        "metric": "image pairs/sec + p50/p99 pair latency at <=0.01 px RMS error, 1 GPU",
        "value": round(value, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "multipass_2160x2560_u16", "pairs_per_step": args.pairs,
                   "schedule": [w[0] for w in SCHEDULE], "overlap": 50, "peak_finder": args.finder,
                   "peak_fit_impl": args.impl, "dft": args.dft, "warp": "bicubic" + (" (separable FMA)" if args.warp == "fma" else " (reference arithmetic)"),
                   "outlier": "median 5x5 eps 0.1 thr 2" + (f" + peak_mag {args.peak_mag}" if args.peak_mag is not None else ""),
                   "infill": args.infill,
                   "global_batch": args.pairs, "parallelism": "single-gpu",
                   "xcorr_variants": ({str(n): r["variant"] for n, r in tuned.items()} if tuned else "default")},
        "latency_ms_p50": round(float(np.percentile(steps_ms, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(steps_ms, 99)), 3),
        "vectors_per_s": round(value * mp.grids[-1].n_win[0] * mp.grids[-1].n_win[1], 1),
        "passes": passes,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "kernel": f"{_xcorr_kernel_name(dom['window'], args.dft)} + fit_kernel",
                     "note": "FP32 VALU FFT; peak = MI355X FP32 157.3 TFLOP/s"},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "one_pair_fp32": one,
        "serving": serve,
        "host_fed": fed,
        "ensemble": ens_leg,
        "mixed": mixed_leg,
        "autotune": ({str(n): r for n, r in tuned.items()} if tuned else None),
    }
    print(json.dumps(line), flush=True)


def _xcorr_kernel_name(n: int, dft: str) -> str:
    """Name of the correlation kernel that runs for n x n windows."""
    from pivtools_amd import autotune

    v = autotune.get_variant(n) or autotune.default_variant(n)
    if dft != "fft" and n == 16:
        return f"xcorr_gemm16_kernel<{ {'split': 0, 'split-bf16': 1, 'split-bf16x3': 2}[dft]}>"
    if (dft == "split" and n in (32, 48, 64)) or v == autotune.MFMA:
        return f"xcorr_mfma_kernel<{n}>"
    if autotune.GEMM <= v <= autotune.GEMM + 2:
        return f"xcorr_gemm16_kernel<{v - autotune.GEMM}>"
    if v == autotune.HALF + 1:
        return f"xcorr_kernel_half1<{n}>"
    if v & autotune.HALF:
        return f"xcorr_kernel_half<{n},{v & ~autotune.HALF}>"
    if v & autotune.PERSIST:
        return f"xcorr_kernel_pf<{n},{v & ~autotune.PERSIST}>"
    return f"xcorr_kernel<{n},{v}>"


if __name__ == "__main__":
    main()
