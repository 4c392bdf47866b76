"""gPb throughput on one MI355X at 481x321 (synthetic Voronoi/grating image).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--check]

A step is one full gPb pass (RGB8 resident in HBM -> oriented gPb) over one
481x321 image.  `value` is images/s of the mixed-precision path; the plain
FP32 path (the reference algorithm) is timed beside it.  The roofline object
covers the stencil SpMV (the kernel inside every Lanczos step), timed with
HIP events on the library's stream; the CPU baseline is the oracle's FP32
reference-order pipeline on one image on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H, CELLS, SEED = 481, 321, 40, 20261202
HBM_PEAK_GBS = 8000.0
SPMV_BYTES_PER_ROW = 81 * 4 + 4 + 4  # 81 matrix entries + x + y (algorithmic)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--spmv-reps", type=int, default=200)
    p.add_argument("--check", action="store_true", help="also report errors vs the FP64 oracle")
    return p.parse_args()


def time_pipeline(ctx, d_rgb, precision, steps, warmup):
    import torch

    for _ in range(warmup):
        ctx.gpb(d_rgb, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.gpb(d_rgb, precision=precision)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, out


def spmv_roofline(ctx, out, reps):
    """Average duration of the stencil SpMV on this image's L_sym, HIP events
    on the library stream (the stream every dmsc call is enqueued on)."""
    import torch

    import damascene_amd as dm

    n = W * H
    m = ctx.intervene(out["nmax"])
    ctx.convert_matrix(W, H, m)
    x = torch.rand(n, device=ctx.device)
    y = torch.empty(n, device=ctx.device)
    for _ in range(5):
        y = ctx.spmv(W, H, m, x)
    stream = ctx.stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        ctx.spmv(W, H, m, x)
    e1.record(stream)
    e1.synchronize()
    avg_s = e0.elapsed_time(e1) / 1e3 / reps
    algo_bytes = SPMV_BYTES_PER_ROW * n
    achieved = algo_bytes / avg_s / 1e9
    return {
        "kernel": "k_spmv (stencil SpMV, 81 diagonals)",
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "avg_us": round(avg_s * 1e6, 3), "bytes_per_launch": algo_bytes,
    }


def cpu_baseline(rgb):
    from oracle import pyoracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    r = O.pipeline(rgb)
    dt = time.perf_counter() - t0
    return {
        "value": round(1.0 / dt, 4), "unit": "images/s", "cores": threads, "kind": "port",
        "sample": f"one {W}x{H} image, full plain-FP32 gPb in reference order (oracle), {dt:.2f} s",
        "stage_ms": [round(v, 1) for v in r["stage_ms"]],
    }, r


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch

    import damascene_amd as dm

    rgb = dm.generate_image(W, H, CELLS, SEED)
    ctx = dm.Context(0)
    d_rgb = ctx.upload(rgb)
    t_plain, out_plain = time_pipeline(ctx, d_rgb, dm.PREC_FP32, a.steps, a.warmup)
    t_mixed, out_mixed = time_pipeline(ctx, d_rgb, dm.PREC_MIXED, a.steps, a.warmup)
    stage_plain = ctx.gpb(d_rgb, precision=dm.PREC_FP32, timing=True)
    stage_mixed = ctx.gpb(d_rgb, precision=dm.PREC_MIXED, timing=True)
    roof = spmv_roofline(ctx, out_plain, a.spmv_reps)
    res = {
        "metric": "gPb images/s @481x321 (mixed vs plain FP32, 1 GPU); SpMV GB/s; err vs FP64",
        "value": round(1.0 / t_mixed, 3),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (mixed: f64 Lanczos inner products)",
        "data": "synthetic (Voronoi cells + gratings + noise, seed 20261202)",
        "config": {"workload": f"gpb-{W}x{H}-single-image", "width": W, "height": H, "cells": CELLS,
                   "n_eig": 9, "tol": 1e-3, "textons": 32},
        "plain_fp32": {"images_per_s": round(1.0 / t_plain, 3), "ms_per_image": round(t_plain * 1e3, 3),
                       "stage_ms": [round(v, 3) for v in stage_plain["stage_ms"]],
                       "lanczos_iterations": stage_plain["lanczos_iterations"]},
        "mixed": {"images_per_s": round(1.0 / t_mixed, 3), "ms_per_image": round(t_mixed * 1e3, 3),
                  "stage_ms": [round(v, 3) for v in stage_mixed["stage_ms"]],
                  "lanczos_iterations": stage_mixed["lanczos_iterations"]},
        "speedup_mixed_vs_plain": round(t_plain / t_mixed, 3),
        "stage_names": ["textons", "localcues", "combine+nonmax", "intervene", "eigensolve", "spectralPb+gPb"],
        "roofline": roof,
    }
    ref = None
    if not a.no_cpu:
        res["cpu_baseline"], ref = cpu_baseline(rgb)
    if ref is not None:
        n = W * H
        res["parity_vs_oracle"] = {
            "textons_mismatch": int(np.count_nonzero(out_plain["textons"].cpu().numpy() != ref["textons"])),
            "nonmax_mismatch": int(np.count_nonzero(
                out_plain["nmax"].cpu().numpy().view(np.uint32) != ref["nmax"].view(np.uint32))),
            "gpb_maxabs_plain": float(np.abs(out_plain["gpb"].cpu().numpy() - ref["gpb"]).max()),
            "gpb_maxabs_mixed": float(np.abs(out_mixed["gpb"].cpu().numpy() - ref["gpb"]).max()),
        }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
