"""Step-kernel shapes and the autotuner.

Every Lanczos step shape the autotuner may pick (mixed: tiled fused with
eight tiles, linear fused with 512 or 256 rows per block, two-kernel with
256 or 64; plain: two-kernel with 256 or 64) must give the default shape's
results up to the rounding of its block partial sums: the pipeline up to the
matrix is shape-independent and bit-exact, and eigenvalues and gPb agree
within the tolerances below (the gPb bar is 1e-3).  Then the autotuner must
time every shape, record the fastest, and later eigensolves of that size
must use it; dmsc_autotune_clear forgets it.
"""
import ctypes

import numpy as np
import pytest

import damascene_amd as dm
from damascene_amd._lib import check, lib

pytestmark = pytest.mark.gpu

W, H, CELLS, SEED = 160, 120, 12, 20261202
TILES = [3208, 6404, 1616, 3216, 1608, 6408, 6402, 3204]
# (mixed step form, tile, block size)
MIXED_SHAPES = [(2, t, 0) for t in TILES] + [(1, 3208, 512), (1, 3208, 256), (0, 3208, 256), (0, 3208, 64)]
PLAIN_SHAPES = [(0, 3208, 256), (0, 3208, 64)]
# eigenvalue relative and gPb max-abs agreement with the default shape
TOL = {dm.PREC_MIXED: (1e-5, 1e-3), dm.PREC_FP32: (1e-3, 1e-3)}


@pytest.fixture(scope="module", autouse=True)
def lanczos_solver():
    """These are the Lanczos step kernels: the mixed path runs its Lanczos
    solver here (the block solver is its default)."""
    with dm.tuned(eigensolver=0):
        yield


def set_shape(fused, tile, k1):
    check(lib().dmsc_tune_set(b"lanczos_mixed_fused", fused))
    check(lib().dmsc_tune_set(b"lanczos_tile", tile))
    check(lib().dmsc_tune_set(b"lanczos_k1_threads", k1))


@pytest.fixture(scope="module")
def image(ctx):
    print(f"case: {W}x{H} cells={CELLS} seed={SEED}", flush=True)
    return ctx.upload(dm.generate_image(W, H, CELLS, SEED))


@pytest.fixture(scope="module")
def defaults(ctx, image):
    check(lib().dmsc_autotune_clear())
    set_shape(2, 3208, 0)
    return {p: ctx.gpb(image, precision=p) for p in (dm.PREC_FP32, dm.PREC_MIXED)}


def compare(out, ref, precision, label):
    np.testing.assert_array_equal(out["textons"].cpu().numpy(), ref["textons"].cpu().numpy())
    np.testing.assert_array_equal(out["nmax"].cpu().numpy().view(np.uint32), ref["nmax"].cpu().numpy().view(np.uint32))
    ev, rv = out["eigvals"][1:].astype(np.float64), ref["eigvals"][1:].astype(np.float64)
    rel = float(np.max(np.abs(ev - rv) / np.abs(rv)))
    gerr = float(np.abs(out["gpb"].cpu().numpy() - ref["gpb"].cpu().numpy()).max())
    print(f"{label}: iterations {out['lanczos_iterations']} (default {ref['lanczos_iterations']}), "
          f"eigval rel {rel:.2e}, gPb max-abs {gerr:.2e}", flush=True)
    etol, gtol = TOL[precision]
    assert rel < etol, f"{label}: eigenvalues differ by {rel:.2e}"
    assert gerr < gtol, f"{label}: gPb differs by {gerr:.2e}"


@pytest.mark.parametrize("precision,shape", [(dm.PREC_MIXED, s) for s in MIXED_SHAPES] +
                         [(dm.PREC_FP32, s) for s in PLAIN_SHAPES])
def test_step_shape_matches_default(ctx, image, defaults, precision, shape):
    print(f"case: precision {precision} shape {shape}", flush=True)
    set_shape(*shape)
    try:
        out = ctx.gpb(image, precision=precision)
    finally:
        set_shape(2, 3208, 0)
    compare(out, defaults[precision], precision, f"shape {shape}")


@pytest.mark.parametrize("precision", [dm.PREC_FP32, dm.PREC_MIXED])
def test_autotune_picks_and_applies(ctx, image, defaults, precision):
    print(f"case: autotune precision {precision}", flush=True)
    m = ctx.intervene(defaults[precision]["nmax"])
    best, trials = ctx.autotune_eigensolver(W, H, m, precision=precision, steps=100)
    print("trials", [(t["mixed_fused"], t["tile"], t["k1_threads"], round(t["ms_per_step"] * 1e3, 2))
                     for t in trials], "best", best, flush=True)
    assert len(trials) == (len(TILES) + 4 if precision == dm.PREC_MIXED else 2)
    assert all(t["ms_per_step"] > 0 for t in trials)
    assert best["ms_per_step"] == min(t["ms_per_step"] for t in trials)
    try:
        out = ctx.gpb(image, precision=precision)  # runs the tuned shape
        compare(out, defaults[precision], precision, "autotuned")
    finally:
        check(lib().dmsc_autotune_clear())
    again = ctx.gpb(image, precision=precision)  # default shape again: bit-identical to the first run
    assert np.array_equal(again["gpb"].cpu().numpy().view(np.uint32),
                          defaults[precision]["gpb"].cpu().numpy().view(np.uint32))


def test_autotune_spmv(ctx, image, defaults):
    print("case: SpMV launch shapes", flush=True)
    m = ctx.intervene(defaults[dm.PREC_FP32]["nmax"])
    ctx.convert_matrix(W, H, m)
    x = np.random.default_rng(3).standard_normal(W * H).astype(np.float32)
    import torch

    xd = torch.from_numpy(x).reshape(H, W).cuda()
    ref = ctx.spmv(W, H, m, xd).cpu().numpy()
    best, us = ctx.autotune_spmv(W, H, m, reps=50)
    print("SpMV us per launch", us, "best rows per lane", best, flush=True)
    assert best in (0, 1, 2, 4) and all(v > 0 for v in us.values())
    assert us[best] == min(us.values())
    v = ctypes.c_int()
    check(lib().dmsc_tune_get(b"spmv_rows_per_lane", ctypes.byref(v)))
    assert v.value == best
    # every shape forms each row's sum in the same diagonal order: same bits
    for shape in (0, 1, 2, 4):
        check(lib().dmsc_tune_set(b"spmv_rows_per_lane", shape))
        y = ctx.spmv(W, H, m, xd).cpu().numpy()
        assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), f"SpMV shape {shape}"
    check(lib().dmsc_tune_set(b"spmv_rows_per_lane", best))


def test_autotune_chebyshev_filter_tile(ctx, image):
    """dmsc_autotune_chebyshev times all 18 filter tiles of the block
    solver at one size and makes the fastest that size's tile; every tile
    gives bit-identical results, so the mixed pipeline's output after tuning
    equals the untuned one bit for bit, and dmsc_autotune_clear forgets it."""
    with dm.tuned(eigensolver=1):
        ref = ctx.gpb(image, precision=dm.PREC_MIXED)
        m = ctx.intervene(ref["nmax"])
        ctx.convert_matrix(W, H, m)
        best, us = ctx.autotune_chebyshev(W, H, m, products=20)
        print(f"filter tiles at {W}x{H}: best {best}; " + ", ".join(f"{k}: {v:.1f} us" for k, v in us.items()),
              flush=True)
        assert len(us) == 18 and best in us and us[best] == min(us.values())
        out = ctx.gpb(image, precision=dm.PREC_MIXED)
        check(lib().dmsc_autotune_clear())
    for f in ("eigvecs", "gpb"):
        assert np.array_equal(out[f].cpu().numpy().view(np.uint32), ref[f].cpu().numpy().view(np.uint32)), f
    assert out["lanczos_iterations"] == ref["lanczos_iterations"]
