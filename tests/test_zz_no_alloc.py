"""dmsc_kmeans_noalloc (k-means on the fixed per-GPU scratch array, no device
allocation) against dmsc_kmeans (the same kernels on the context's arenas):
labels bit-identical, same iteration and change counts, in both precisions,
on two calls in a row with the second one smaller (the array keeps the
first call's contents, so the second call proves the state is re-cleared),
plus the enqueue-only form and the size limit."""
import ctypes

import numpy as np
import pytest
import torch

import damascene_amd as dm
from damascene_amd._lib import DmscError, check, lib

pytestmark = pytest.mark.gpu

SIZES = [(481, 321, 40, 20261202), (160, 120, 12, 20261203)]  # second call smaller


def responses(ctx, w, h, cells, seed):
    grey = ctx.rgb_to_grey(ctx.upload(dm.generate_image(w, h, cells, seed)))
    return ctx.texton_responses(grey)


def kmeans_noalloc(ctx, resp, precision, k, counts=True):
    _, h, w = resp.shape
    labels = torch.full((h, w), -7, dtype=torch.int32, device="cuda")
    ch, it = ctypes.c_int(-1), ctypes.c_int(-1)
    args = (ctypes.byref(ch), ctypes.byref(it)) if counts else (None, None)
    check(lib().dmsc_kmeans_noalloc(ctx.handle, w, h, k, 34, resp.data_ptr(), labels.data_ptr(), 15, 0, precision,
                                    *args))
    if not counts:
        ctx.synchronize()
    return labels, it.value, ch.value


def kmeans_arena(ctx, resp, precision, k):
    _, h, w = resp.shape
    labels = torch.empty((h, w), dtype=torch.int32, device="cuda")
    ch, it = ctypes.c_int(), ctypes.c_int()
    check(lib().dmsc_kmeans(ctx.handle, w, h, k, 34, resp.data_ptr(), labels.data_ptr(), 15, 0, precision,
                            ctypes.byref(ch), ctypes.byref(it)))
    return labels, it.value, ch.value


@pytest.mark.parametrize("precision,k", [(dm.PREC_FP32, 32), (dm.PREC_MIXED, 32), (dm.PREC_FP32, 64)])
def test_noalloc_matches_kmeans_twice(ctx, precision, k):
    for w, h, cells, seed in SIZES:
        print(f"case: dmsc_kmeans_noalloc {w}x{h} seed={seed} precision={precision} k={k}", flush=True)
        resp = responses(ctx, w, h, cells, seed)
        ref, it_ref, ch_ref = kmeans_arena(ctx, resp, precision, k)
        got, it, ch = kmeans_noalloc(ctx, resp, precision, k)
        torch.cuda.synchronize()
        assert it == it_ref and ch == ch_ref, (it, it_ref, ch, ch_ref)
        bad = int((got != ref).sum().item())
        assert bad == 0, f"{bad} of {w * h} labels differ"
        assert int(got.min().item()) >= 0 and int(got.max().item()) < k


def test_noalloc_enqueue_only(ctx):
    w, h, cells, seed = SIZES[1]
    print(f"case: dmsc_kmeans_noalloc enqueue-only {w}x{h}", flush=True)
    resp = responses(ctx, w, h, cells, seed)
    ref, _, _ = kmeans_arena(ctx, resp, dm.PREC_MIXED, 32)
    got, _, _ = kmeans_noalloc(ctx, resp, dm.PREC_MIXED, 32, counts=False)
    assert bool(torch.equal(got, ref))


def test_noalloc_size_limit(ctx):
    lim = lib().dmsc_kmeans_noalloc_max_pixels()
    assert lim >= 1920 * 1080
    # rejected on the host before any launch: the buffers are never touched
    small = torch.zeros(16, dtype=torch.float32, device="cuda")
    lab = torch.zeros(16, dtype=torch.int32, device="cuda")
    w = 2048
    h = lim // w + 1
    for precision in (dm.PREC_FP32, dm.PREC_MIXED):
        rc = lib().dmsc_kmeans_noalloc(ctx.handle, w, h, 32, 34, small.data_ptr(), lab.data_ptr(), 15, 0,
                                       precision, None, None)
        assert rc == 1
        assert b"fixed scratch" in lib().dmsc_last_error_string()
    with pytest.raises(DmscError):
        check(lib().dmsc_kmeans_noalloc(ctx.handle, w, h, 32, 34, small.data_ptr(), lab.data_ptr(), 15, 0,
                                        dm.PREC_FP32, None, None))
    torch.cuda.synchronize()
    assert int(lab.abs().sum().item()) == 0
