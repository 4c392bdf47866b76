"""gPb throughput on one MI355X at 481x321 (synthetic Voronoi/grating image).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]

A step is one full gPb pass (RGB8 resident in HBM -> oriented gPb) over one
481x321 image.  `value` is images/s of the mixed-precision path: BF16 MFMA
k-means distances with an exact FP32 re-check, and the block eigensolver
(Chebyshev-filtered subspace iteration on 32 FP32 vectors, FP64 Gram
matrices, Rayleigh-Ritz and residuals).  The plain FP32 path (the reference
algorithm) and the mixed path on its Lanczos solver are timed beside it, and
all report their errors against the FP64 ground truth in
tests/golden/truth_481x321.npz (eigenvalue relative error, subspace angle of
eigenvectors 1..8, gPb max-abs error, nonmax mismatches against the oracle).

roofline: the dominant kernel of the timed region, the block solver's filter
step at the tile the library picks at 481x321 (reported as `tile`: 2337,
k_cheb_ldp<3, 7>, 64 x 21 pixels per block, whose 512 blocks fill one round
of the chip's block slots).  One launch per filter step
forms the 81-diagonal stencil product of a block of 32 vectors, y = a (A x -
c x) - b z, at 708 algorithmic bytes per row (the matrix row 324 B; x, z and
y 128 B each).  Its average duration is measured with HIP events on the
library's stream over
FILTER_TIMED launches; `traffic` is the HBM bytes per launch from the
rocprofv3 FETCH_SIZE/WRITE_SIZE runs committed under profiles/.  Beside it:
`valu` (the same launch's FP32 multiply-adds against the FP32 peak),
`lanczos_step` (the fused mixed Lanczos step, 348 B per row) and `spmv` (the
standalone SpMV, 332 B per row, over its launch shapes).

batched: dmsc_gpb_batch over 16 images of 481x321 (seeds 20261202 + i) and
over 64 images of uneven size (sides uniform in [200, 1024]), 4 concurrent
lanes, both precisions; image 0 of the 16 is the single-image workload and
is checked bit for bit against the single-image call.

large: the 1024x768 configuration (120 cells): images/s of both precisions,
the step and SpMV rooflines at that size, and FP64 a posteriori residuals of
the eigenpairs (no stored truth at that size).

differentiable: 200 Adam steps (lr 1e-2) refitting the combine weights on
16 images of 481x321 through the hand-written spectral Pb + combine backward
(loss before and after, time per forward + backward over the 16 images).

cpu_baseline: the oracle's reference-order pipeline on one image on the host
cores (OpenMP).
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H, CELLS, SEED, NVEC, TOL = 481, 321, 40, 20261202, 9, 1e-3
HBM_PEAK_GBS = 8000.0
BATCH, BATCH_LANES, UNEVEN = 16, 4, 64  # 481x321 batch, concurrent lanes, uneven-size batch
SPMV_BYTES_PER_ROW = 81 * 4 + 4 + 4  # 81 matrix entries + x + y (algorithmic)
SPMV_SHAPES = {0: "k_spmv_full", 1: "k_spmv<1>", 2: "k_spmv<2>", 4: "k_spmv<4>"}
# fused mixed Lanczos step, per row: 81 matrix entries (324 B), the staged
# z'_{k-1} (8 B) and q_{k-1} (4 B) read once, q_k (4 B) and z'_k (8 B) written
STEP_KERNEL = "k_lzmt<32, 8, 0>"
STEP_BYTES_PER_ROW = 81 * 4 + 8 + 4 + 4 + 8
STEP_TIMED = 500
# block eigensolver filter step (the default tile), per row: the matrix row
# (324 B) and, for 32 vectors, x, z read and y written (3 x 128 B)
# the filter kernel of each pipelined LDS tile (chebyshev_shape value); the
# library picks 2337 at 481x321 (one round of the chip's block slots)
FILTER_KERNELS = {2328: "k_cheb_ldp<2, 8>", 2337: "k_cheb_ldp<3, 7>", 2324: "k_cheb_ldp<2, 4>",
                  1328: "k_cheb_lds<2, 8, false, true>", 1337: "k_cheb_lds<3, 7, false, true>",
                  1324: "k_cheb_lds<2, 4, false, true>", 1336: "k_cheb_lds<3, 6, false, true>",
                  1344: "k_cheb_lds<4, 4, false, true>", 1348: "k_cheb_lds<4, 8, false, true>"}
FILTER_BYTES_PER_ROW = 81 * 4 + 3 * 32 * 4
FILTER_FLOPS_PER_ROW = 2 * 81 * 32
FILTER_TIMED = 300
FP32_PEAK_TFLOPS = 157.3
TRUTH = os.path.join(ROOT, "tests", "golden", "truth_481x321.npz")
# kernel groups for --pmc-summarize: (name regex, algorithmic bytes per image);
# the histogram stage's bytes: u16 integral images written once and read
# once, label frames, L/a/b/texton inputs and the 96 chi-squared planes
PMC_GROUPS = {
    "histogram": (r"k_fill_u8|k_cue_labels|k_integral_\w+|k_band_carry|k_gradients\w*", 890.3e6),
}
PMC_SUMMARIES = [os.path.join(ROOT, "profiles", f) for f in
                 ("chebyshev_filter_481x321_pmc.txt", "lanczos_step_481x321_pmc.txt",
                  "stencil_spmv_481x321_pmc.txt")]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--no-batch", action="store_true", help="skip the batched and differentiable measurements")
    p.add_argument("--spmv-reps", type=int, default=200)
    p.add_argument("--block-profile", action="store_true",
                   help="only run the mixed pipeline with the block eigensolver (eigensolver = 1) three times, "
                        "for a kernel trace")
    p.add_argument("--pmc-workload", action="store_true",
                   help="run only the workload of the rocprofv3 counter passes: one 256 MiB device copy "
                        "(calibration), 20 launches of the selected SpMV shape and 100 mixed Lanczos steps, "
                        "then exit")
    p.add_argument("--spmv-shape", type=int, default=1, choices=sorted(SPMV_SHAPES),
                   help="SpMV launch shape for --pmc-workload (0 = k_spmv_full, else rows per lane)")
    p.add_argument("--pmc-summarize", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                   help="summarise two rocprofv3 --pmc output folders (FETCH_SIZE, WRITE_SIZE passes of "
                        "--pmc-workload) into HBM bytes per launch of --kernel, calibrated on the 256 MiB copy")
    p.add_argument("--pmc-mfma", metavar="DIR",
                   help="summarise a rocprofv3 --pmc SQ_INSTS_VALU_MFMA_MOPS_BF16 SQ_VALU_MFMA_BUSY_CYCLES "
                        "GRBM_GUI_ACTIVE pass of --pmc-workload into MFMA FLOPs and utilisation of --kernel")
    p.add_argument("--pmc-valu", metavar="DIR",
                   help="summarise a rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_VMEM_RD SQ_WAVES "
                        "GRBM_GUI_ACTIVE pass of --pmc-workload into VALU issue utilisation of --kernel (a kernel "
                        "name or the 'histogram' group)")
    p.add_argument("--pmc-sq", metavar="DIR",
                   help="summarise a rocprofv3 --pmc pass of SQ wave-state counters (SQ_WAVES SQ_WAVE_CYCLES "
                        "SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_VMEM "
                        "SQ_ACTIVE_INST_LDS) of --pmc-workload for --kernel")
    p.add_argument("--kernel", default=STEP_KERNEL,
                   help="kernel for --pmc-summarize: the Lanczos step kernel or an SpMV shape name")
    return p.parse_args()


def pmc_mfma(d, kernel):
    """MFMA work and utilisation of `kernel` from one rocprofv3 pass with
    SQ_INSTS_VALU_MFMA_MOPS_BF16, SQ_VALU_MFMA_BUSY_CYCLES and GRBM_GUI_ACTIVE:
    BF16 MFMA FLOPs = 512 per MOP; utilisation = busy cycles summed over the
    SIMDs / (GPU-active cycles per XCD x 1024 SIMDs), GRBM_GUI_ACTIVE summing
    the 8 XCDs (and including the profiler's dispatch window, so the figure
    is a lower bound)."""
    import csv
    import glob

    rows = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if not re.search(re.escape(kernel) + r"\(", r["Kernel_Name"]):
                continue
            e = rows.setdefault(r["Dispatch_Id"], {"ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"])})
            c, v = r["Counter_Name"], float(r["Counter_Value"])
            # SQ counters are summed over their instances, GRBM_GUI_ACTIVE is the
            # largest instance (rocprofv3's MfmaUtil: reduce(GRBM_GUI_ACTIVE, max))
            e[c] = max(e.get(c, 0.0), v) if c == "GRBM_GUI_ACTIVE" else e.get(c, 0.0) + v
    if not rows:
        raise SystemExit(f"no counter rows for {kernel} under {d}")
    n = len(rows)
    flops = sum(e["SQ_INSTS_VALU_MFMA_MOPS_BF16"] for e in rows.values()) * 512 / n
    busy = sum(e["SQ_VALU_MFMA_BUSY_CYCLES"] for e in rows.values()) / n
    active = sum(e["GRBM_GUI_ACTIVE"] for e in rows.values()) / n
    ns = sum(e["ns"] for e in rows.values()) / n
    print(f"{kernel} ({n} dispatches): BF16 MFMA {flops / 1e6:.1f} MFLOP per dispatch, "
          f"MFMA busy {busy:.0f} SIMD-cycles, GPU active {active:.0f} cycles, "
          f"MFMA utilisation >= {100 * busy / (active / 8 * 1024):.3f} %, {ns / 1e3:.1f} us under the profiler "
          f"({flops / ns / 1e3:.2f} TFLOP/s)")


def pmc_valu(d, kernel):
    """VALU issue of `kernel` (a name, or a PMC_GROUPS group summed over its
    kernels) from one rocprofv3 pass with SQ_INSTS_VALU, SQ_INSTS_LDS,
    SQ_INSTS_VMEM_RD, SQ_WAVES and GRBM_GUI_ACTIVE: VALU instructions per
    wave, and the VALU issue utilisation = SQ_INSTS_VALU x 4 cycles (a wave64
    FP32 instruction on a 16-lane SIMD) / (cycles x 1024 SIMDs).  Two cycle
    counts are used, each an upper bound on the kernel's own cycles, so both
    figures are lower bounds on how busy the vector ALUs were:
    GRBM_GUI_ACTIVE / 8 (the counter sums the 8 XCDs and includes the
    profiler's dispatch window) and the dispatch duration at the 2.4 GHz peak
    engine clock."""
    import csv
    import glob

    group = PMC_GROUPS.get(kernel)
    pattern = (r"(?:^|::|\s)(?:" + group[0] + r")[<(]") if group else (re.escape(kernel) + r"\(")
    rows = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if not re.search(pattern, r["Kernel_Name"]):
                continue
            e = rows.setdefault((r["Kernel_Name"], r["Dispatch_Id"]),
                                {"ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"])})
            c, v = r["Counter_Name"], float(r["Counter_Value"])
            e[c] = max(e.get(c, 0.0), v) if c == "GRBM_GUI_ACTIVE" else e.get(c, 0.0) + v
    if not rows:
        raise SystemExit(f"no counter rows for {kernel} under {d}")
    by_kernel = {}
    for (k, _), e in rows.items():
        name = re.sub(r"\(.*", "", k.replace("dmsc::(anonymous namespace)::", "").replace("void ", ""))
        acc = by_kernel.setdefault(name, {"n": 0})
        acc["n"] += 1
        for c, v in e.items():
            acc[c] = acc.get(c, 0.0) + v
    def util(valu, grbm, ns):
        return (100 * valu * 4 / max(grbm / 8 * 1024, 1), 100 * valu * 4 / max(ns * 2.4 * 1024, 1))

    tot = {"SQ_INSTS_VALU": 0.0, "GRBM_GUI_ACTIVE": 0.0, "ns": 0.0}
    for name, a in sorted(by_kernel.items(), key=lambda kv: -kv[1].get("ns", 0.0)):
        n = a["n"]
        waves = a.get("SQ_WAVES", 0.0) / n
        valu, act, ns = a.get("SQ_INSTS_VALU", 0.0) / n, a.get("GRBM_GUI_ACTIVE", 0.0) / n, a["ns"] / n
        u1, u2 = util(valu, act, ns)
        print(f"{name} ({n} dispatches, {ns / 1e3:.1f} us each under the profiler): {waves:.0f} waves, VALU "
              f"{valu / max(waves, 1):.0f} / LDS {a.get('SQ_INSTS_LDS', 0.0) / n / max(waves, 1):.0f} / VMEM read "
              f"{a.get('SQ_INSTS_VMEM_RD', 0.0) / n / max(waves, 1):.0f} instructions per wave; VALU issue >= "
              f"{u1:.1f} % (GRBM_GUI_ACTIVE / 8 = {act / 8:.0f} cycles), >= {u2:.1f} % (duration at 2.4 GHz)")
        for k in tot:
            tot[k] += a.get(k, 0.0)
    if group:
        u1, u2 = util(tot["SQ_INSTS_VALU"], tot["GRBM_GUI_ACTIVE"], tot["ns"])
        print(f"{kernel} (all kernels, summed, {tot['ns'] / 1e3:.1f} us): VALU issue >= {u1:.1f} % (GRBM_GUI_ACTIVE "
              f"/ 8), >= {u2:.1f} % (duration at 2.4 GHz)")


def pmc_sq(d, kernel):
    """Wave-state counters of `kernel` from one rocprofv3 pass with SQ_WAVES,
    SQ_WAVE_CYCLES, SQ_BUSY_CYCLES, SQ_WAIT_ANY, SQ_WAIT_INST_ANY,
    SQ_ACTIVE_INST_VALU, SQ_ACTIVE_INST_VMEM and SQ_ACTIVE_INST_LDS: the
    share of the waves' lifetime spent waiting (on memory counters) and
    issuing each instruction kind."""
    import csv
    import glob

    rows = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            if re.search(re.escape(kernel) + r"\(", r["Kernel_Name"]):
                e = rows.setdefault(r["Dispatch_Id"], {})
                e[r["Counter_Name"]] = e.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    if not rows:
        raise SystemExit(f"no counter rows for {kernel} under {d}")
    n = len(rows)
    avg = {k: sum(e.get(k, 0.0) for e in rows.values()) / n for k in next(iter(rows.values()))}
    wc = avg["SQ_WAVE_CYCLES"]
    print(f"{kernel} ({n} dispatches): {avg['SQ_WAVES']:.0f} waves, wave-cycles {wc:.0f}; share of wave "
          f"lifetime: waiting {avg['SQ_WAIT_ANY'] / wc:.3f}, waiting to issue {avg['SQ_WAIT_INST_ANY'] / wc:.3f}, "
          f"VALU {avg['SQ_ACTIVE_INST_VALU'] / wc:.3f}, LDS {avg['SQ_ACTIVE_INST_LDS'] / wc:.3f}, "
          f"VMEM {avg['SQ_ACTIVE_INST_VMEM'] / wc:.3f}")


def pmc_summarize(fetch_dir, write_dir, kernel):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the launches of `kernel`,
    each scaled by what the same counter reads for the 256 MiB copy (expected
    262,144 KiB read and 262,144 KiB written)."""
    import csv
    import glob

    def per_kernel(d, counter):
        """{(kernel, dispatch): value} from rocprofv3's CSV or SQLite output."""
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] == counter:
                    key = (r["Kernel_Name"], r.get("Dispatch_Id", ""))
                    vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
        for f in glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True):
            import sqlite3

            con = sqlite3.connect(f)
            q = "select kernel_name, dispatch_id, value from counters_collection where counter_name = ?"
            for k, disp, v in con.execute(q, (counter,)):
                vals[(k, disp)] = vals.get((k, disp), 0.0) + float(v)
        if not vals:
            raise SystemExit(f"no {counter} values under {d}")
        return vals

    group = PMC_GROUPS.get(kernel)
    # the name as rocprofv3 prints it, up to the argument list; a group
    # sums every dispatch of its kernels (the workload's one gPb image)
    pattern = (r"(?:^|::|\s)(?:" + group[0] + r")[<(]") if group else (re.escape(kernel) + r"\(")
    out = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        v = per_kernel(d, counter)
        copies = [x for (k, _), x in v.items() if "copy" in k.lower()]
        hits = [x for (k, _), x in v.items() if re.search(pattern, k)]
        if not copies or not hits:
            raise SystemExit(f"{counter}: copy kernel or {kernel} not found in {d}")
        cal = max(copies)
        per = sum(hits) if group else sum(hits) / len(hits)
        out[counter] = dict(copy_kib=cal, scale=262144.0 / cal, kib=per, launches=len(hits))
    f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
    traffic = (f["kib"] * f["scale"] + w["kib"] * w["scale"]) * 1024
    if group:
        algo = group[1]
    elif kernel in FILTER_KERNELS.values():
        algo = FILTER_BYTES_PER_ROW * W * H
    else:
        algo = (STEP_BYTES_PER_ROW if kernel == STEP_KERNEL else SPMV_BYTES_PER_ROW) * W * H
    lines = [
        f"calibration copy (256 MiB): FETCH_SIZE {f['copy_kib']:.0f} KiB (scale {f['scale']:.4f}), "
        f"WRITE_SIZE {w['copy_kib']:.0f} KiB (scale {w['scale']:.4f})",
        f"{kernel} ({f['launches']} launches): FETCH_SIZE {f['kib']:.1f} KiB, WRITE_SIZE {w['kib']:.1f} KiB "
        f"{'summed over the image' if group else 'per launch'}",
        f"algorithmic bytes {'per image' if group else 'per launch'} {algo:.0f}",
        f"traffic_bytes_per_launch {kernel} {traffic:.0f}",
        f"traffic / algorithmic = {traffic / algo:.3f}",
    ]
    print("\n".join(lines))


def pmc_workload(shape):
    """Counter-pass workload: a 256 MiB copy whose FETCH_SIZE/WRITE_SIZE
    calibrate the counters (256 MiB read + 256 MiB written), then 20 stencil
    SpMVs on the 481x321 L_sym with the given launch shape, then 125 fused
    mixed Lanczos steps (25 warm-up + 100 timed) on the same matrix, then 56
    block-solver filter steps (default tile) on it, then one mixed-precision
    k-means (BF16 MFMA distances) on the image's responses."""
    import torch

    import damascene_amd as dm
    from damascene_amd._lib import check, lib

    ctx = dm.Context(0)
    out = ctx.gpb(ctx.upload(dm.generate_image(W, H, CELLS, SEED)), precision=dm.PREC_FP32)
    m = ctx.intervene(out["nmax"])
    ctx.convert_matrix(W, H, m)
    a = torch.rand(64 << 20, device=ctx.device)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    x = torch.rand(W * H, device=ctx.device)
    y = torch.empty_like(x)
    check(lib().dmsc_tune_set(b"spmv_rows_per_lane", shape))
    for _ in range(20):
        check(lib().dmsc_stencil_spmv(ctx.handle, W, H, m.data_ptr(), m.shape[1], x.data_ptr(), y.data_ptr()))
    torch.cuda.synchronize()
    m2 = ctx.intervene(out["nmax"])
    ms = ctypes.c_float()
    check(lib().dmsc_lanczos_time_steps(ctx.handle, W, H, m2.data_ptr(), m2.shape[1], dm.PREC_MIXED, 100,
                                        ctypes.byref(ms)))
    torch.cuda.synchronize()
    m3 = ctx.intervene(out["nmax"])
    ctx.convert_matrix(W, H, m3)
    check(lib().dmsc_chebyshev_time_products(ctx.handle, W, H, m3.data_ptr(), m3.shape[1], 50, ctypes.byref(ms)))
    torch.cuda.synchronize()
    d_rgb = ctx.upload(dm.generate_image(W, H, CELLS, SEED))
    resp = ctx.texton_responses(ctx.rgb_to_grey(d_rgb))
    ctx.kmeans(resp, precision=dm.PREC_MIXED)
    torch.cuda.synchronize()
    print(f"pmc workload done: copy 256 MiB, 20 x {SPMV_SHAPES[shape]}, 125 mixed Lanczos steps, "
          "56 block-solver filter steps, one mixed k-means", flush=True)


def step_roofline(ctx, nmax):
    """Average duration of the fused mixed Lanczos step (HIP events on the
    library's stream over STEP_TIMED steps, no checks) and its roofline."""
    from damascene_amd._lib import check, lib
    import damascene_amd as dm

    m = ctx.intervene(nmax)
    ms = ctypes.c_float()
    check(lib().dmsc_lanczos_time_steps(ctx.handle, W, H, m.data_ptr(), m.shape[1], dm.PREC_MIXED, STEP_TIMED,
                                        ctypes.byref(ms)))
    avg_s = ms.value / 1e3
    algo = STEP_BYTES_PER_ROW * W * H
    achieved = algo / avg_s / 1e9
    return {
        "kernel": STEP_KERNEL + " (fused mixed Lanczos step: FP64-accumulated 81-diagonal SpMV + update)",
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(STEP_KERNEL),
        "avg_us": round(avg_s * 1e6, 3), "algorithmic_bytes_per_launch": algo, "launches_timed": STEP_TIMED,
    }


def filter_roofline(ctx, nmax):
    """Average duration of the block eigensolver's filter step (the default
    tile, HIP events on the library's stream over FILTER_TIMED launches on
    the image's L_sym) and its roofline: HBM on 708 algorithmic bytes per
    row, with its FP32 multiply-adds against the FP32 peak beside it."""
    from damascene_amd._lib import check, lib

    m = ctx.intervene(nmax)
    ctx.convert_matrix(W, H, m)
    ms = ctypes.c_float()
    check(lib().dmsc_chebyshev_time_products(ctx.handle, W, H, m.data_ptr(), m.shape[1], FILTER_TIMED,
                                             ctypes.byref(ms)))
    tile = ctx.chebyshev_tile(W, H)
    kernel = FILTER_KERNELS.get(tile, f"filter tile {tile}")
    avg_s = ms.value / 1e3
    algo = FILTER_BYTES_PER_ROW * W * H
    achieved = algo / avg_s / 1e9
    tflops = FILTER_FLOPS_PER_ROW * W * H / avg_s / 1e12
    return {
        "kernel": kernel + " (block eigensolver filter step: 81-diagonal stencil product of 32 vectors, "
                           "LDS-staged input, pipelined matrix loads)",
        "tile": tile,
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(kernel),
        "avg_us": round(avg_s * 1e6, 3), "algorithmic_bytes_per_launch": algo, "launches_timed": FILTER_TIMED,
        "valu": {"achieved_tflops": round(tflops, 2), "peak_tflops": FP32_PEAK_TFLOPS,
                 "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "flops_per_launch": FILTER_FLOPS_PER_ROW * W * H},
    }


def time_pipeline(ctx, d_rgb, precision, steps, warmup):
    import torch

    for _ in range(warmup):
        ctx.gpb(d_rgb, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.gpb(d_rgb, precision=precision)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def batch_images(ctx):
    """The 16-image batch at 481x321 (seeds SEED + i; image 0 is the
    single-image workload) and the uneven 64-image batch: width and height
    each uniform in [200, 1024] (sizes drawn with seed SEED), cells = area /
    4000, image seeds SEED + 100 + i."""
    import damascene_amd as dm

    same = [ctx.upload(dm.generate_image(W, H, CELLS, SEED + i)) for i in range(BATCH)]
    specs = uneven_specs()
    uneven = [ctx.upload(dm.generate_image(*s)) for s in specs]
    return same, uneven, [s[:2] for s in specs]


def uneven_specs():
    """(w, h, cells, seed) of the uneven 64-image batch: width and height
    each uniform in [200, 1024] (drawn with seed SEED), cells = area / 4000
    (at least 4), image seeds SEED + 100 + i."""
    rng = np.random.default_rng(SEED)
    sizes = [(int(rng.integers(200, 1025)), int(rng.integers(200, 1025))) for _ in range(UNEVEN)]
    return [(w, h, max(4, w * h // 4000), SEED + 100 + i) for i, (w, h) in enumerate(sizes)]


def time_batch(ctx, rgbs, precision, steps, warmup):
    """Seconds per batch call (dmsc_gpb_batch, BATCH_LANES lanes) and the
    outputs of the last call; two or more warm-up calls size every lane's
    arenas (a lane may take other images on the second call)."""
    import torch

    bufs = ctx.gpb_batch_buffers(rgbs)
    for _ in range(max(2, warmup)):
        ctx.gpb_batch(rgbs, precision=precision, lanes=BATCH_LANES, buffers=bufs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.gpb_batch(rgbs, precision=precision, lanes=BATCH_LANES, buffers=bufs)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def same_bits(a, b):
    fields = ("textons", "nmax", "eigvecs", "gpb_all", "gpb")
    return all(np.array_equal(a[f].cpu().numpy().view(np.uint32), b[f].cpu().numpy().view(np.uint32))
               for f in fields) and np.array_equal(a["eigvals"].view(np.uint32), b["eigvals"].view(np.uint32))


def batched(ctx, out_plain, out_mixed, steps, warmup):
    same, uneven, sizes = batch_images(ctx)
    res = {"lanes": BATCH_LANES, "images": BATCH, "size": f"{W}x{H}",
           "what": "dmsc_gpb_batch: images spread over concurrent lanes (child contexts with their own "
                   "streams and host threads), largest first; per-image results bit-identical to single calls; "
                   "eigensolver_iterations: Lanczos steps (plain), block filter products (mixed)"}
    for name, prec, single in (("plain_fp32", 0, out_plain), ("mixed", 1, out_mixed)):
        t, out = time_batch(ctx, same, prec, steps, warmup)
        res[name] = {"images_per_s": round(BATCH / t, 3), "ms_per_batch": round(t * 1e3, 3),
                     "eigensolver_iterations": [o["lanczos_iterations"] for o in out],
                     "image0_bit_identical_to_single": bool(same_bits(out[0], single))}
    tu = {}
    for name, prec in (("plain_fp32", 0), ("mixed", 1)):
        t, out = time_batch(ctx, uneven, prec, 1, 2)
        tu[name] = {"images_per_s": round(UNEVEN / t, 3), "s_per_batch": round(t, 3),
                    "max_eigensolver_iterations": max(o["lanczos_iterations"] for o in out)}
    res["uneven"] = {"images": UNEVEN, "sides": "uniform in [200, 1024]", "megapixels": round(
        sum(w * h for w, h in sizes) / 1e6, 2), **tu}
    return res


def differentiable(ctx, steps=200, images=16, lr=1e-2):
    """Refit of the combine weights through the hand-written spectral Pb +
    combine backward: 16 images of 481x321 (seeds SEED + 200 + i) with their
    Voronoi cell boundaries dilated by 1 px as the masks, eigenvectors from
    the mixed eigensolver, class-balanced BCE before non-max; Adam on the
    12 cue x scale weights and gamma.  Reports the loss before and after and
    the time of one forward + backward over the 16 images."""
    import torch

    import damascene_amd as dm
    from damascene_amd import autograd as AG

    ops, vecs, masks, etas = [], [], [], []
    for i in range(images):
        rgb, cells = dm.generate_image(W, H, CELLS, SEED + 200 + i, with_cells=True)
        d_rgb = ctx.upload(rgb)
        out = ctx.gpb(d_rgb, precision=dm.PREC_MIXED)
        L, a, b = ctx.rgb_to_lab(d_rgb)
        tex, _ = ctx.find_textons(ctx.rgb_to_grey(d_rgb))
        ops.append(AG.SpbCombine(ctx, W, H, ctx.local_cues(L, a, b, tex)))
        vecs.append(out["eigvecs"])
        masks.append(torch.from_numpy(dm.boundary_mask(cells).astype(bool)).to(ctx.device))
        etas.append(AG.reference_params(out["eigvals"], device=ctx.device)[13:])
    w0 = AG.reference_params(np.ones(NVEC, np.float32), device=ctx.device)[:13]
    wg = w0.clone().requires_grad_(True)
    opt = torch.optim.Adam([wg], lr=lr)

    def step():
        opt.zero_grad()
        loss = sum(AG.balanced_bce(op(torch.cat([wg, e]), v), m) for op, v, m, e in zip(ops, vecs, masks, etas))
        loss = loss / images
        loss.backward()
        return loss

    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        losses.append(step().item())
        opt.step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"images": images, "size": f"{W}x{H}", "adam_steps": steps, "lr": lr,
            "loss_initial": round(losses[0], 6), "loss_final": round(losses[-1], 6),
            "loss_drop": round(1 - losses[-1] / losses[0], 4),
            "ms_per_step_fwd_bwd_16_images": round(dt * 1e3, 3),
            "what": "Adam on the 12 cue x scale weights and the sPb weight through dmsc_spb_combine_forward/"
                    "backward (torch.autograd.Function); class-balanced BCE on the PostProcess logit of the "
                    "pre-non-max gPb against cell boundaries dilated 1 px"}


def stage_roofline(stage, iterations, kmeans_iterations, mixed):
    """Per-stage roofline fractions of one 481x321 image from a timed run with
    every stage in order on one stream (tune key overlap_stages = 0), so each
    window holds exactly its own kernels: [6] colour + texton filter bank,
    [0] - [6] k-means, [7] rotated labels + integral images + chi-squared
    gradients (the histogram stage), [1] - [7] parabola fits, [4] eigensolve,
    [5] spectral Pb + gPb.  Each stage's algorithmic work over its window,
    against the bound it is written for: FP32 VALU 157.3 TFLOP/s, BF16 MFMA
    2.5 PFLOP/s (the mixed k-means distances), HBM 8 TB/s."""
    n = W * H
    valu, bf16 = 157.3e12, 2.5e15
    t = [v * 1e-3 for v in stage]  # seconds
    bank = 2 * 9010 * n
    km = kmeans_iterations * 2 * n * 32 * 34
    parabola = 2 * 58720 * n
    hist = 890.3e6
    eig = iterations * (FILTER_BYTES_PER_ROW if mixed else 352) * n
    spb = 6272 * n
    t_bank, t_km = t[6], t[0] - t[6]
    t_hist, t_par = t[7], t[1] - t[7]
    km_peak = bf16 if mixed else valu
    return {
        "timing": "stages in order on one stream (overlap_stages 0); HIP events between stages",
        "texton_bank": {"bound": "fp32 valu", "work": "2*9010 flop/px", "ms": round(t_bank * 1e3, 3),
                        "achieved_tflops": round(bank / t_bank / 1e12, 2), "frac": round(bank / t_bank / valu, 4)},
        "kmeans": {"bound": "bf16 mfma" if mixed else "fp32 valu", "ms": round(t_km * 1e3, 3),
                   "work": f"{kmeans_iterations} iterations x 2*N*32*34 flop",
                   "achieved_tflops": round(km / t_km / 1e12, 2), "frac": round(km / t_km / km_peak, 5)},
        "histogram": {"bound": "hbm", "work": "890.3 MB (u16 integral images written once and read once, labels, "
                                              "chi-squared planes)", "ms": round(t_hist * 1e3, 3),
                      "achieved_gbs": round(hist / t_hist / 1e9, 1), "frac": round(hist / t_hist / 8e12, 4)},
        "parabola": {"bound": "fp32 valu", "work": "2*58720 flop/px", "ms": round(t_par * 1e3, 3),
                     "achieved_tflops": round(parabola / t_par / 1e12, 2),
                     "frac": round(parabola / t_par / valu, 4)},
        "eigensolve": {"bound": "hbm", "ms": round(t[4] * 1e3, 3),
                       "work": (f"{iterations} block filter products x {FILTER_BYTES_PER_ROW} B/row" if mixed else
                                f"{iterations} Lanczos steps x 352 B/row"),
                       "achieved_gbs": round(eig / t[4] / 1e9, 1), "frac": round(eig / t[4] / 8e12, 4)},
        "spectralpb_gpb": {"bound": "fp32 valu", "work": "6272 flop/px", "ms": round(t[5] * 1e3, 3),
                           "achieved_tflops": round(spb / t[5] / 1e12, 2), "frac": round(spb / t[5] / valu, 4)},
    }


def sequential_stages(ctx, d_rgb, precision):
    """One timed gPb with the stages in order on one stream."""
    from damascene_amd._lib import check, lib

    check(lib().dmsc_tune_set(b"overlap_stages", 0))
    try:
        ctx.gpb(d_rgb, precision=precision, timing=True)
        return ctx.gpb(d_rgb, precision=precision, timing=True)
    finally:
        check(lib().dmsc_tune_set(b"overlap_stages", 1))


def fp64_residuals(ctx, out, w, h):
    """A posteriori FP64 check of the returned eigenpairs, needing no stored
    truth: u_j = D^1/2 v_j with the D^1/2 direction projected out (undoing
    the per-vector min-max map), normalised; the Rayleigh quotient
    lam_j = u_j' L u_j and the residual |L u_j - lam_j u_j| in FP64, with L
    the library's L_sym of the same image (FP32 entries, exact in FP64) and
    a plain DIA product over the 81 stencil offsets; and the Davis-Kahan bound
    sin(angle_j) <= |r_j| / gap_j against the computed neighbours (the gap
    of the last vector to the unknown next eigenvalue is not bounded).
    rounding_floor_j bounds what the FP32 storage of the normalised output
    vector alone contributes to residual_j."""
    import torch

    n = w * h
    m = ctx.intervene(out["nmax"])
    _, rs = ctx.convert_matrix(w, h, m)
    # diagonal k of the DIA matrix: the k-th point of the radius-5 disc in scan order
    offs = [dy * w + dx for dy in range(-5, 6) for dx in range(-5, 6) if dy * dy + dx * dx <= 25]
    pad = 5 * w + 5
    A = m[:, :n].double()
    sq = 1.0 / rs.reshape(-1).double()
    u0 = sq / sq.norm()
    V = out["eigvecs"].reshape(NVEC, n).double()[1:] * sq[None, :]
    V -= (V @ u0)[:, None] * u0[None, :]
    # what FP32 storage of the min-max mapped vectors alone leaves: entries in
    # [0, 1] carry up to 2^-25, which D^1/2 scales and L (|L| <= 2) passes on,
    # against the norm of the part that survives the projection
    unorm = V.norm(dim=1)
    floor = (2.0 * 2.0 ** -25 * sq.norm() / unorm).cpu().numpy()
    V /= unorm[:, None]
    xp = torch.zeros(V.shape[0], n + 2 * pad, dtype=torch.float64, device=V.device)
    xp[:, pad:pad + n] = V
    LV = torch.zeros_like(V)
    for k in range(81):
        LV += A[k][None, :] * xp[:, pad + int(offs[k]):pad + int(offs[k]) + n]
    lam = (V * LV).sum(dim=1)
    res = (LV - lam[:, None] * V).norm(dim=1)
    lam, res = lam.cpu().numpy(), res.cpu().numpy()
    got = np.asarray(out["eigvals"], np.float64)[1:]
    gaps = np.array([np.min(np.abs(lam[j] - np.delete(lam, j))) for j in range(len(lam))])
    return {"residual_max": float(res.max()), "residuals": [float(f"{r:.3e}") for r in res],
            "rounding_floor": [float(f"{f:.3e}") for f in floor],
            "eigval_vs_rayleigh_max_rel": float(np.max(np.abs(got - lam) / lam)),
            "davis_kahan_sin_bound_max": float(np.max(res[:-1] / gaps[:-1]))}


def large(ctx, steps):
    """The 1024x768 configuration (120 cells, seed SEED): images/s of both
    precisions, the filter, Lanczos step and SpMV rooflines at that size, and
    the FP64 a posteriori residuals of both paths' eigenpairs."""
    import ctypes

    import torch

    import damascene_amd as dm
    from damascene_amd._lib import check, lib

    w, h = 1024, 768
    d_rgb = ctx.upload(dm.generate_image(w, h, 120, SEED))
    res = {"size": f"{w}x{h}", "cells": 120}
    outs = {}
    for name, prec in (("plain_fp32", dm.PREC_FP32), ("mixed", dm.PREC_MIXED)):
        t, out = time_pipeline(ctx, d_rgb, prec, steps, 1)
        outs[name] = out
        res[name] = {"images_per_s": round(1.0 / t, 3), "ms_per_image": round(t * 1e3, 3),
                     "lanczos_iterations": out["lanczos_iterations"]}
    res["mixed"]["fp64_check"] = fp64_residuals(ctx, outs["mixed"], w, h)
    res["plain_fp32"]["fp64_check"] = fp64_residuals(ctx, outs["plain_fp32"], w, h)
    n = w * h
    mf = ctx.intervene(outs["mixed"]["nmax"])
    ctx.convert_matrix(w, h, mf)
    ms = ctypes.c_float()
    check(lib().dmsc_chebyshev_time_products(ctx.handle, w, h, mf.data_ptr(), mf.shape[1], 100, ctypes.byref(ms)))
    tile = ctx.chebyshev_tile(w, h)
    res["filter"] = {"kernel": FILTER_KERNELS.get(tile, f"filter tile {tile}"), "avg_us": round(ms.value * 1e3, 3),
                     "achieved_gbs": round(FILTER_BYTES_PER_ROW * n / (ms.value * 1e-3) / 1e9, 1)}
    res["filter"]["frac"] = round(res["filter"]["achieved_gbs"] / HBM_PEAK_GBS, 4)
    m = ctx.intervene(outs["mixed"]["nmax"])
    check(lib().dmsc_lanczos_time_steps(ctx.handle, w, h, m.data_ptr(), m.shape[1], dm.PREC_MIXED, STEP_TIMED,
                                        ctypes.byref(ms)))
    res["step"] = {"kernel": STEP_KERNEL, "avg_us": round(ms.value * 1e3, 3),
                   "achieved_gbs": round(STEP_BYTES_PER_ROW * n / (ms.value * 1e-3) / 1e9, 1)}
    res["step"]["frac"] = round(res["step"]["achieved_gbs"] / HBM_PEAK_GBS, 4)
    best, us = ctx.autotune_spmv(w, h, m, reps=100)
    res["spmv"] = {"best_rows_per_lane": best, "avg_us": round(us[best], 3),
                   "achieved_gbs": round(332 * n / (us[best] * 1e-6) / 1e9, 1)}
    res["spmv"]["frac"] = round(res["spmv"]["achieved_gbs"] / HBM_PEAK_GBS, 4)
    check(lib().dmsc_tune_set(b"spmv_rows_per_lane", 1))
    torch.cuda.synchronize()
    return res


def spmv_sweep(ctx, nmax, reps):
    """Average duration of each SpMV launch shape (HIP events on the library
    stream, the stream every dmsc call is enqueued on); leaves the fastest
    shape selected and returns the roofline object for it."""
    import torch

    from damascene_amd._lib import check, lib

    n = W * H
    m = ctx.intervene(nmax)
    ctx.convert_matrix(W, H, m)
    x = torch.rand(n, device=ctx.device)
    y = torch.empty(n, device=ctx.device)
    spmv = lib().dmsc_stencil_spmv
    args = (ctx.handle, W, H, m.data_ptr(), m.shape[1], x.data_ptr(), y.data_ptr())
    times = {}
    for shape in SPMV_SHAPES:
        check(lib().dmsc_tune_set(b"spmv_rows_per_lane", shape))
        for _ in range(5):
            check(spmv(*args))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(ctx.stream)
        for _ in range(reps):  # launches are queued faster than they run
            spmv(*args)
        e1.record(ctx.stream)
        e1.synchronize()
        times[shape] = e0.elapsed_time(e1) / 1e3 / reps
    best = min(times, key=times.get)
    check(lib().dmsc_tune_set(b"spmv_rows_per_lane", best))
    avg_s = times[best]
    algo = SPMV_BYTES_PER_ROW * n
    achieved = algo / avg_s / 1e9
    roof = {
        "kernel": SPMV_SHAPES[best] + " (stencil SpMV, 81 diagonals)",
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(SPMV_SHAPES[best]),
        "avg_us": round(avg_s * 1e6, 3), "algorithmic_bytes_per_launch": algo,
        "shapes_us": {SPMV_SHAPES[k]: round(v * 1e6, 3) for k, v in times.items()},
    }
    return roof


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summaries (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel)."""
    for path in PMC_SUMMARIES:
        if not os.path.exists(path):
            continue
        m = re.search(r"^traffic_bytes_per_launch\s+" + re.escape(kernel) + r"\s+(\d+)", open(path).read(), re.M)
        if m:
            return int(m.group(1))
    return None


def errors_vs_truth(out, truth):
    """Eigenvalue relative error, largest principal angle between the spans of
    u_1..u_8 (u = D^1/2 v, with the D^1/2 direction projected out, so the
    per-vector min-max map does not matter) and gPb max-abs error."""
    vals = np.asarray(out["eigvals"], np.float64)
    tv = truth["eigvals"][:NVEC]
    sq = truth["sqrt_d"].astype(np.float64)
    u0 = sq / np.linalg.norm(sq)
    G = out["eigvecs"].cpu().numpy().reshape(NVEC, -1)[1:].astype(np.float64) * sq[None, :]
    G -= np.outer(G @ u0, u0)
    R = truth["u"].astype(np.float64)
    R -= np.outer(R @ u0, u0)
    qa, _ = np.linalg.qr(G.T)
    qb, _ = np.linalg.qr(R.T)
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return {
        "eigval_max_rel": float(np.max(np.abs(vals[1:] - tv[1:]) / tv[1:])),
        "subspace_angle": float(np.arccos(np.clip(s.min(), -1, 1))),
        "gpb_max_abs": float(np.abs(out["gpb"].cpu().numpy() - truth["gpb"]).max()),
    }


def cpu_baseline(rgb):
    """The oracle's plain-FP32 pipeline in reference order on the host cores
    (OpenMP), in its timing mode: the texton bank, parabola and spectral Pb
    correlations and the interior rows of the SpMV run as register-tiled
    SIMD loops (AVX-512 where the host has it, else AVX2; the technique of
    the reference's cpu_convolve2d.cc) with every output's operation
    sequence unchanged, so its results equal the scalar oracle's bit for
    bit.  One warm-up, then the median of five runs of the 481x321 image
    with every thread; beside it one run of the scalar loops, and one thread
    on the 160x120 image (the reference's CPU-runnable configuration, 12
    cells) after a warm-up."""
    import ctypes
    import platform
    import statistics

    import damascene_amd as dm
    from oracle import pyoracle as O

    lib = O.lib()
    lib.orc_max_threads.restype = ctypes.c_int
    threads = lib.orc_max_threads()
    lib.orc_set_simd(0)
    t0 = time.perf_counter()
    scalar = O.pipeline(rgb, nvec=NVEC, tol=TOL)
    dt_scalar = time.perf_counter() - t0
    lib.orc_set_simd(1)
    try:
        r = O.pipeline(rgb, nvec=NVEC, tol=TOL)
        runs = []
        for _ in range(5):
            t0 = time.perf_counter()
            r = O.pipeline(rgb, nvec=NVEC, tol=TOL)
            runs.append(time.perf_counter() - t0)
        dt = statistics.median(runs)
        same = all(np.array_equal(np.asarray(r[k]).view(np.uint32), np.asarray(scalar[k]).view(np.uint32))
                   for k in ("textons", "nmax", "gpb"))
        small = dm.generate_image(160, 120, 12, SEED)
        lib.orc_set_threads(1)
        try:
            O.pipeline(small, nvec=NVEC, tol=TOL)
            t0 = time.perf_counter()
            O.pipeline(small, nvec=NVEC, tol=TOL)
            dt1 = time.perf_counter() - t0
        finally:
            lib.orc_set_threads(threads)
    finally:
        lib.orc_set_simd(0)
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), cpu)
    except OSError:
        pass
    return {
        "value": round(1.0 / dt, 4), "unit": "images/s", "cores": threads, "kind": "port",
        "sample": f"one {W}x{H} image, whole plain-FP32 gPb in reference order (oracle, OpenMP, {threads} threads, "
                  f"register-tiled SIMD correlations and SpMV, {lib.orc_simd_width()} columns per tile): median of "
                  f"5 runs after a warm-up, {dt:.2f} s",
        "stage_ms": [round(v, 1) for v in r["stage_ms"]],
        "simd_results_identical_to_scalar": bool(same),
        "scalar_loops_s_per_image": round(dt_scalar, 3),
        "one_thread_160x120": {"images_per_s": round(1.0 / dt1, 4), "s_per_image": round(dt1, 3),
                               "nproc": os.cpu_count()},
        "host_cpu": cpu,
    }, r


def lanczos_solver(ctx, d_rgb, steps, warmup):
    """The mixed pipeline on its Lanczos solver (tune key "eigensolver" =
    0): seconds per image, the outputs, and one timed run's stage times."""
    import damascene_amd as dm

    with dm.tuned(eigensolver=0):
        t, out = time_pipeline(ctx, d_rgb, dm.PREC_MIXED, steps, warmup)
        stage = ctx.gpb(d_rgb, precision=dm.PREC_MIXED, timing=True)
    return t, out, stage


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    if a.pmc_workload:
        pmc_workload(a.spmv_shape)
        return
    if a.pmc_summarize:
        pmc_summarize(*a.pmc_summarize, a.kernel)
        return
    if a.pmc_mfma:
        pmc_mfma(a.pmc_mfma, a.kernel)
        return
    if a.pmc_valu:
        pmc_valu(a.pmc_valu, a.kernel)
        return
    if a.pmc_sq:
        pmc_sq(a.pmc_sq, a.kernel)
        return
    import torch

    import damascene_amd as dm

    rgb = dm.generate_image(W, H, CELLS, SEED)
    ctx = dm.Context(0)
    d_rgb = ctx.upload(rgb)
    if a.block_profile:
        t, _ = time_pipeline(ctx, d_rgb, dm.PREC_MIXED, 3, 1)
        st = ctx.gpb(d_rgb, precision=dm.PREC_MIXED, timing=True)
        print(json.dumps({"block_solver_ms_per_image": round(t * 1e3, 3), "block_products": st["lanczos_iterations"],
                          "stage_ms": [round(v, 3) for v in st["stage_ms"]]}))
        return
    t_plain, out_plain = time_pipeline(ctx, d_rgb, dm.PREC_FP32, a.steps, a.warmup)
    t_mixed, out_mixed = time_pipeline(ctx, d_rgb, dm.PREC_MIXED, a.steps, a.warmup)
    t_lz, out_lz, stage_lz = lanczos_solver(ctx, d_rgb, a.steps, a.warmup)
    stage_plain = ctx.gpb(d_rgb, precision=dm.PREC_FP32, timing=True)
    stage_mixed = ctx.gpb(d_rgb, precision=dm.PREC_MIXED, timing=True)
    seq_plain = sequential_stages(ctx, d_rgb, dm.PREC_FP32)
    seq_mixed = sequential_stages(ctx, d_rgb, dm.PREC_MIXED)
    batch = None if a.no_batch else batched(ctx, out_plain, out_mixed, max(3, a.steps), a.warmup)
    diff = None if a.no_batch else differentiable(ctx)
    big = None if a.no_batch else large(ctx, 2)
    spmv = spmv_sweep(ctx, out_plain["nmax"], a.spmv_reps)
    roof = filter_roofline(ctx, out_plain["nmax"])
    roof["lanczos_step"] = step_roofline(ctx, out_plain["nmax"])
    torch.cuda.synchronize()
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
        "dtype": "f32 (mixed: bf16 MFMA k-means distances with f32 re-check; f32 filter, f64 Gram/Rayleigh-Ritz)",
        "data": "synthetic (Voronoi cells + gratings + noise, seed 20261202)",
        "config": {"workload": f"gpb-{W}x{H}-single-image", "width": W, "height": H, "cells": CELLS,
                   "n_eig": NVEC, "tol": TOL, "textons": 32},
        "plain_fp32": {"images_per_s": round(1.0 / t_plain, 3), "ms_per_image": round(t_plain * 1e3, 3),
                       "stage_ms": [round(v, 3) for v in stage_plain["stage_ms"][:6]],
                       "lanczos_iterations": stage_plain["lanczos_iterations"],
                       "stage_roofline": stage_roofline(seq_plain["stage_ms"], seq_plain["lanczos_iterations"],
                                                        seq_plain["kmeans_iterations"], False)},
        "mixed": {"images_per_s": round(1.0 / t_mixed, 3), "ms_per_image": round(t_mixed * 1e3, 3),
                  "stage_ms": [round(v, 3) for v in stage_mixed["stage_ms"][:6]],
                  "block_products": stage_mixed["lanczos_iterations"],
                  "stage_roofline": stage_roofline(seq_mixed["stage_ms"], seq_mixed["lanczos_iterations"],
                                                   seq_mixed["kmeans_iterations"], True)},
        "speedup_mixed_vs_plain": round(t_plain / t_mixed, 3),
        "stage_names": ["textons (+ overlapped L/a/b local cues)", "local cues (rest)", "combine+nonmax",
                        "intervene", "eigensolve", "spectralPb+gPb"],
        "roofline": roof,
        "spmv": spmv,
    }
    res["mixed_lanczos"] = {
        "what": "mixed pipeline on its Lanczos solver (tune key eigensolver = 0: FP64-accumulated fused steps "
                "from the reference's uniform start, adapted Cullum screening, stop at residual tol*1e-4)",
        "images_per_s": round(1.0 / t_lz, 3), "ms_per_image": round(t_lz * 1e3, 3),
        "stage_ms": [round(v, 3) for v in stage_lz["stage_ms"][:6]],
        "lanczos_iterations": stage_lz["lanczos_iterations"]}
    if diff:
        res["differentiable"] = diff
    if big:
        res["large"] = big
    res["mixed"]["fp64_check"] = fp64_residuals(ctx, out_mixed, W, H)
    if batch:
        res["batched"] = batch
        res["batched"]["mixed_batched_vs_plain_single"] = round(batch["mixed"]["images_per_s"] * t_plain, 3)
    if os.path.exists(TRUTH):
        truth = np.load(TRUTH)
        res["plain_fp32"]["err_vs_fp64"] = errors_vs_truth(out_plain, truth)
        res["mixed"]["err_vs_fp64"] = errors_vs_truth(out_mixed, truth)
        res["mixed_lanczos"]["err_vs_fp64"] = errors_vs_truth(out_lz, truth)
    if not a.no_cpu:
        res["cpu_baseline"], ref = cpu_baseline(rgb)
        for name, out in (("plain_fp32", out_plain), ("mixed", out_mixed)):
            res[name]["textons_mismatch_vs_oracle"] = int(np.count_nonzero(out["textons"].cpu().numpy()
                                                                           != ref["textons"]))
            res[name]["nonmax_mismatch_vs_oracle"] = int(np.count_nonzero(
                out["nmax"].cpu().numpy().view(np.uint32) != ref["nmax"].view(np.uint32)))
        res["plain_fp32"]["gpb_max_abs_vs_oracle_plain"] = float(
            np.abs(out_plain["gpb"].cpu().numpy() - ref["gpb"]).max())
    print(json.dumps(res))


if __name__ == "__main__":
    main()
