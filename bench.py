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
MEASURED_TRAFFIC_64_PAIRS = {128: 2.796e9 + 0.011e9 + 0.010e9 + 0.002e9,
                             64: 2.811e9 + 0.044e9 + 0.043e9 + 0.008e9,
                             32: 2.832e9 + 0.180e9 + 0.175e9 + 0.033e9,
                             16: 2.832e9 + 0.725e9 + 0.703e9 + 0.132e9}


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


def cpu_baseline(A_dev, B_dev, gpu_final, finder, budget_s):
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
    ref = M.run_multipass(a, b, grids, finder=finder)[0]
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
        r64 = M.run_multipass(a, b, grids, finder=3, flavour="fp64")[0]
        fx, fy, fm = r64["ux"][0], r64["uy"][0], r64["nan_mask"][0]
        same = ~gm & ~fm
        acc["rms_px_vs_fp64"] = float(np.sqrt(np.mean((gx[same] - fx[same]) ** 2 + (gy[same] - fy[same]) ** 2)))
        acc["mask_flip_frac_vs_fp64"] = float(np.mean(gm != fm))
        acc["fp64_oracle_s"] = round(time.perf_counter() - t0, 2)
    return out, acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--finder", default="gauss3")
    ap.add_argument("--impl", default="scalar")
    ap.add_argument("--cpu-budget", type=float, default=60.0, help="0 disables the CPU baseline leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one plain 256 MiB device copy (byte-counter calibration for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent: N GPUs would each run this batch on their own stream of pairs)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pivtools_amd.multipass import MultipassPIV

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log(f"[bench] generating {args.pairs} pairs of {SHAPE} uint16 on the device")
    A, B = make_batch(args.pairs, dev)
    torch.cuda.synchronize()
    mp = MultipassPIV(SHAPE, SCHEDULE, OVERLAP, peak_finder=args.finder, peak_fit_impl=args.impl)
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
    cpu = None
    acc = {}
    if args.cpu_budget > 0:
        log("[bench] CPU oracle leg on one pair")
        cpu, acc = cpu_baseline(A, B, (gx, gy, gm), mp.finder, args.cpu_budget)

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
                   "peak_fit_impl": args.impl, "warp": "bicubic", "outlier": "median 5x5 eps 0.1 thr 2",
                   "global_batch": args.pairs, "parallelism": "single-gpu"},
        "latency_ms_p50": round(float(np.percentile(steps_ms, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(steps_ms, 99)), 3),
        "vectors_per_s": round(value * mp.grids[-1].n_win[0] * mp.grids[-1].n_win[1], 1),
        "passes": passes,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "kernel": f"xcorr_kernel<{dom['window']},{dom['window']}> + fit_kernel",
                     "note": "FP32 VALU FFT; peak = MI355X FP32 157.3 TFLOP/s"},
        "cpu_baseline": cpu,
        "accuracy": acc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
