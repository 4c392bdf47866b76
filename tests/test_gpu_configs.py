"""Parity at the full workloads of the benchmark, through the C ABI.

* 481x321, plain FP32 path (the reference algorithm, damascene.cu:269-470)
  against the CPU oracle's plain pipeline on the same image: texton labels,
  mPb and nonmax bit-exact, eigenvalues within 1e-4 relative, gPb within
  1e-3 (the oracle runs the same FP32 recurrence, so the two agree far more
  closely).
* 481x321, mixed path: labels, mPb and nonmax bit-exact against the oracle
  (BF16 MFMA k-means distances with the FP32 re-check must give the FP32
  labels); eigenvalues, eigenvector subspace (u = D^1/2 v, vectors 1..8)
  and gPb against the committed FP64 truth of the image
  (tests/golden/truth_481x321.npz): relative 1e-4, angle 1e-4, 1e-3.
* The uneven batch: 64 images with width and height uniform in [200, 1024]
  (the benchmark's sizes and seeds) through dmsc_gpb_batch in both
  precisions; every output of every image bit-identical to a dmsc_gpb call
  on that image, and image 0 against the oracle (exact stages).
* convert.h helpers (mirrorImage float/int, unMirrorImage, quantizeImage)
  against the oracle's restatements (convert.cu:122-246) and numpy.
"""
import os

import numpy as np
import pytest
import torch

import bench as B
import damascene_amd as dm
from damascene_amd._lib import check, lib
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

W, H, CELLS, SEED, NVEC, TOL = 481, 321, 40, 20261202, 9, 1e-3
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def host(t):
    return t.detach().cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def subspace_angle(A, B_):
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B_)
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1, 1)))


@pytest.fixture(scope="module")
def full():
    print(f"case: oracle pipeline {W}x{H} cells={CELLS} seed={SEED}", flush=True)
    rgb = dm.generate_image(W, H, CELLS, SEED)
    ref = O.pipeline(rgb, nvec=NVEC, tol=TOL)
    assert ref["rc"] == 0
    return rgb, ref


def exact_stages(out, ref, w, h):
    n = w * h
    tex = host(out["textons"])
    bad_tex = int(np.count_nonzero(tex != ref["textons"]))
    bad_mpb = int(np.count_nonzero(bits(host(out["mpb"])[:, :n].reshape(8, h, w)) != bits(ref["mpb"])))
    bad_nm = int(np.count_nonzero(bits(host(out["nmax"])) != bits(ref["nmax"])))
    return bad_tex, bad_mpb, bad_nm


def test_pipeline_plain_481x321(ctx, full):
    rgb, ref = full
    print(f"case: plain pipeline {W}x{H}", flush=True)
    out = ctx.gpb(ctx.upload(rgb), precision=dm.PREC_FP32, timing=True)
    bad = exact_stages(out, ref, W, H)
    ev = out["eigvals"].astype(np.float64)
    rel = np.abs(ev[1:] - ref["eigvals"][1:]) / np.abs(ref["eigvals"][1:])
    gerr = float(np.abs(host(out["gpb"]) - ref["gpb"]).max())
    print(f"plain {W}x{H}: texton/mPb/nonmax mismatches {bad}, Lanczos {out['lanczos_iterations']}, eigval max "
          f"rel vs oracle {rel.max():.3e}, gPb max abs vs oracle {gerr:.3e}", flush=True)
    assert bad == (0, 0, 0)
    assert out["lanczos_iterations"] % 500 == 0
    assert rel.max() < 1e-4
    assert gerr <= 1e-3


def test_pipeline_mixed_481x321(ctx, full):
    rgb, ref = full
    print(f"case: mixed pipeline {W}x{H}", flush=True)
    truth = np.load(os.path.join(GOLDEN, "truth_481x321.npz"))
    assert list(truth["meta"]) == [W, H, CELLS, SEED, NVEC]
    out = ctx.gpb(ctx.upload(rgb), precision=dm.PREC_MIXED, timing=True)
    bad = exact_stages(out, ref, W, H)
    n = W * H
    tv = truth["eigvals"][:NVEC]
    rel = np.abs(out["eigvals"][1:].astype(np.float64) - tv[1:]) / tv[1:]
    sq = truth["sqrt_d"].astype(np.float64)
    u0 = sq / np.linalg.norm(sq)
    G = host(out["eigvecs"]).reshape(NVEC, n)[1:].astype(np.float64) * sq[None, :]
    G -= np.outer(G @ u0, u0)
    R = truth["u"].astype(np.float64)
    R -= np.outer(R @ u0, u0)
    ang = subspace_angle(G.T, R.T)
    gerr = float(np.abs(host(out["gpb"]) - truth["gpb"]).max())
    print(f"mixed {W}x{H}: texton/mPb/nonmax mismatches {bad}, Lanczos {out['lanczos_iterations']}, eigval max "
          f"rel vs FP64 {rel.max():.3e}, subspace angle {ang:.3e}, gPb max abs vs FP64 truth {gerr:.3e}", flush=True)
    assert bad == (0, 0, 0)
    assert rel.max() < 1e-4
    assert ang <= 1e-4
    assert gerr <= 1e-3


FIELDS = ["textons", "mpb", "combined_g", "nmax", "eigvecs", "spb", "gpb_all", "gpb"]


def same_outputs(a, b):
    for f in FIELDS:
        x, y = host(a[f]), host(b[f])
        if x.shape != y.shape or not np.array_equal(x.view(np.uint32), y.view(np.uint32)):
            return f
    if not np.array_equal(a["eigvals"].view(np.uint32), b["eigvals"].view(np.uint32)):
        return "eigvals"
    if a["lanczos_iterations"] != b["lanczos_iterations"] or a["kmeans_iterations"] != b["kmeans_iterations"]:
        return "iterations"
    return None


@pytest.mark.parametrize("precision", [dm.PREC_FP32, dm.PREC_MIXED])
def test_batch_uneven_64(ctx, precision):
    specs = B.uneven_specs()
    assert len(specs) == 64
    print(f"case: uneven batch of {len(specs)} images ({sum(w * h for w, h, _, _ in specs) / 1e6:.1f} MP), "
          f"precision {precision}", flush=True)
    hosts = [dm.generate_image(*s) for s in specs]
    rgbs = [ctx.upload(r) for r in hosts]
    batch = ctx.gpb_batch(rgbs, precision=precision)
    torch.cuda.synchronize()
    diffs = []
    for i, r in enumerate(rgbs):
        single = ctx.gpb(r, precision=precision)
        d = same_outputs(single, batch[i])
        if d:
            diffs.append((i, specs[i][:2], d))
        del single
    print(f"batch vs single calls: {len(diffs)} images differ {diffs[:4]}; Lanczos iterations "
          f"{min(o['lanczos_iterations'] for o in batch)}..{max(o['lanczos_iterations'] for o in batch)}",
          flush=True)
    assert not diffs
    if precision == dm.PREC_FP32:
        w, h = specs[0][:2]
        print(f"case: oracle on batch image 0 ({w}x{h})", flush=True)
        ref = O.pipeline(hosts[0], nvec=NVEC, tol=TOL)
        assert ref["rc"] == 0
        bad = exact_stages(batch[0], ref, w, h)
        gerr = float(np.abs(host(batch[0]["gpb"]) - ref["gpb"]).max())
        print(f"image 0 vs oracle: texton/mPb/nonmax mismatches {bad}, gPb max abs {gerr:.3e}", flush=True)
        assert bad == (0, 0, 0)
        assert gerr <= 1e-3


@pytest.mark.parametrize("w,h,border", [(160, 120, 30), (61, 37, 3), (31, 31, 31)])
def test_convert_helpers(ctx, w, h, border):
    print(f"case: mirror/unmirror/quantize {w}x{h} border {border}", flush=True)
    rng = np.random.default_rng(w * 1000 + h)
    img = rng.random((h, w), dtype=np.float32)
    img[0, :3] = [0.0, 1.0, np.float32(0.99999994)]  # the top clamp and the exact ends
    lab = rng.integers(-5, 100, (h, w), dtype=np.int32)
    bw, bh = w + 2 * border, h + 2 * border
    mf = host(ctx.mirror_image(ctx.upload(img), border))
    mi = host(ctx.mirror_image(ctx.upload(lab), border))
    ref_f = np.zeros((bh, bw), np.float32)
    ref_i = np.zeros((bh, bw), np.int32)
    O.lib().orc_mirror_f(w, h, border, O.P(img), O.P(ref_f))
    O.lib().orc_mirror_i(w, h, border, O.P(lab), O.P(ref_i))
    np.testing.assert_array_equal(bits(mf), bits(ref_f))
    np.testing.assert_array_equal(mi, ref_i)
    np.testing.assert_array_equal(mf, np.pad(img, border, mode="symmetric"))
    un = host(ctx.unmirror_image(ctx.upload(mf), border))
    np.testing.assert_array_equal(bits(un), bits(img))
    for nbins in (25, 32, 64):
        q = host(ctx.quantize_image(ctx.upload(img), nbins))
        ref_q = np.zeros((h, w), np.int32)
        O.lib().orc_quantize(w, h, nbins, O.P(img), O.P(ref_q))
        np.testing.assert_array_equal(q, ref_q)
        assert q.min() >= 0 and q.max() == nbins - 1


def test_sequential_stages_bit_identical(ctx):
    """Tuning key overlap_stages = 0 runs every stage in order on one stream:
    the same kernels, so every output equals the overlapped run's bit for bit,
    and the timing splits the textons (filter bank) and the local cues
    (histogram stage) for the per-stage roofline figures."""
    from damascene_amd._lib import check, lib

    print("case: overlap_stages 0 vs 1, 160x120", flush=True)
    rgb = ctx.upload(dm.generate_image(160, 120, 12, 20261202))
    a = ctx.gpb(rgb, precision=dm.PREC_MIXED, timing=True)
    check(lib().dmsc_tune_set(b"overlap_stages", 0))
    try:
        b = ctx.gpb(rgb, precision=dm.PREC_MIXED, timing=True)
    finally:
        check(lib().dmsc_tune_set(b"overlap_stages", 1))
    assert same_outputs(a, b) is None
    assert a["stage_ms"][6] == 0 and a["stage_ms"][7] == 0
    st = b["stage_ms"]
    print("sequential stage ms", [round(v, 3) for v in st], flush=True)
    assert 0 < st[6] < st[0] and 0 < st[7] < st[1]


@pytest.mark.parametrize("w,h,texton_choice", [(481, 321, dm.TEXTON32), (1024, 768, dm.TEXTON32),
                                               (160, 120, dm.TEXTON64), (61, 37, dm.TEXTON32)])
def test_integral_passes_bit_identical(ctx, w, h, texton_choice):
    """The single-write banded integral images (tune key integral_pass 1,
    the default) against the row + column passes (0): the chi-squared
    planes of all four channels, which read every corner of every
    integral image, bit for bit."""
    print(f"case: integral passes {w}x{h} texton_choice {texton_choice}", flush=True)
    d_rgb = ctx.upload(dm.generate_image(w, h, max(4, w * h // 4000), SEED + 7))
    L, a, b = ctx.rgb_to_lab(d_rgb)
    tex, _ = ctx.find_textons(ctx.rgb_to_grey(d_rgb), texton_choice=texton_choice)
    raws = {}
    try:
        for mode in (0, 1):
            check(lib().dmsc_tune_set(b"integral_pass", mode))
            _, raws[mode] = ctx.local_cues(L, a, b, tex, raw=True, texton_choice=texton_choice)
            torch.cuda.synchronize()
    finally:
        check(lib().dmsc_tune_set(b"integral_pass", 1))
    g0, g1 = bits(host(raws[0])), bits(host(raws[1]))
    assert np.count_nonzero(g0 != g1) == 0
    assert np.isfinite(host(raws[1])).all() and float(np.abs(host(raws[1])).max()) > 0
