"""Batched gPb (dmsc_gpb_batch) against single-image calls.

A batch spreads its images over concurrent lanes (child contexts with their
own streams and host threads), largest first; every image runs the kernels
and launch shapes of dmsc_gpb, so each output must be bit-identical to a
dmsc_gpb call on the same image.  Sizes are uneven and some repeat, the
lane count is smaller than the image count (lanes take several images and
reuse their arenas), and one batch is run twice on the same context.
The first image is also checked against the CPU oracle (exact textons and
nonmax), tying the batch to the reference path.
"""
import numpy as np
import pytest
import torch

import damascene_amd as dm
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

SIZES = [(160, 120), (96, 128), (233, 97), (64, 200), (160, 120), (128, 96), (201, 149)]
FIELDS = ["textons", "mpb", "combined_g", "nmax", "eigvecs", "spb", "gpb_all", "gpb"]


def images(ctx):
    out = []
    for i, (w, h) in enumerate(SIZES):
        cells = max(4, (w * h) // 1600)
        out.append(dm.generate_image(w, h, cells, 20261202 + i))
    return out, [ctx.upload(r) for r in out]


def assert_same(a, b, label):
    for f in FIELDS:
        x, y = a[f].detach().cpu().numpy(), b[f].detach().cpu().numpy()
        assert x.shape == y.shape, f"{label}: {f} shape"
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{label}: {f} differs"
    assert np.array_equal(a["eigvals"].view(np.uint32), b["eigvals"].view(np.uint32)), f"{label}: eigvals"
    assert a["lanczos_iterations"] == b["lanczos_iterations"], f"{label}: Lanczos iterations"
    assert a["kmeans_iterations"] == b["kmeans_iterations"], f"{label}: k-means iterations"


@pytest.mark.parametrize("precision", [dm.PREC_FP32, dm.PREC_MIXED])
def test_batch_matches_single(precision):
    print(f"case: batch of {len(SIZES)} uneven images, precision {precision}, 3 lanes", flush=True)
    ctx = dm.Context(0)
    host_rgbs, rgbs = images(ctx)
    single = [ctx.gpb(r, precision=precision) for r in rgbs]
    torch.cuda.synchronize()
    bufs = ctx.gpb_batch_buffers(rgbs)
    for rep in range(2):
        batch = ctx.gpb_batch(rgbs, precision=precision, lanes=3, buffers=bufs)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(single, batch)):
            assert_same(a, b, f"rep {rep} image {i} {SIZES[i]}")
    # the first image against the oracle (exact stages)
    w, h = SIZES[0]
    ref = O.pipeline(host_rgbs[0], nvec=9, tol=1e-3)
    np.testing.assert_array_equal(batch[0]["textons"].cpu().numpy(), ref["textons"])
    np.testing.assert_array_equal(batch[0]["nmax"].cpu().numpy().reshape(h, w).view(np.uint32),
                                  ref["nmax"].reshape(h, w).view(np.uint32))
    ctx.close()


def test_batch_one_lane_and_errors():
    print("case: batch with one lane, then an invalid image", flush=True)
    ctx = dm.Context(0)
    _, rgbs = images(ctx)
    rgbs = rgbs[:3]
    single = [ctx.gpb(r) for r in rgbs]
    batch = ctx.gpb_batch(rgbs, lanes=1)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(single, batch)):
        assert_same(a, b, f"one lane image {i}")
    # an image below the minimum side is rejected before any lane starts
    small = ctx.upload(dm.generate_image(16, 16, 2, 1))
    with pytest.raises(dm.DmscError):
        ctx.gpb_batch([rgbs[0], small])
    assert ctx.gpb_batch([]) == []
    ctx.close()
