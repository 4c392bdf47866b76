"""Eigensolver and end-to-end parity of the HIP path against the CPU oracle.

Exact: texton labels, mPb/combinedG and nonmax (the pipeline up to the
affinity matrix is bit-exact by contract).

Plain FP32 path: the oracle runs the same FP32 recurrence (FP32 SpMV in the
same diagonal order, FP32 vectors, FP64-accumulated inner products rounded
to FP32), so the GPU's Lanczos trajectory, Ritz values and gPb follow the
oracle's closely.

Mixed path: compared with an FP64 ground truth -- FP64 Lanczos with full
reorthogonalisation on the same L_sym, converged to residual 1e-10, then the
same finishing stages (D^-1/2 scaling, min-max, spectral Pb, gPb).
"""
import numpy as np
import pytest
import torch

import damascene_amd as dm
from damascene_amd._lib import check, lib
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

NVEC, TOL = 9, 1e-3


def host(t):
    return t.detach().cpu().numpy()


def subspace_angle(A, B):
    """Largest principal angle (radians) between the column spans of A and B."""
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B)
    s = np.linalg.svd(qa.T @ qb, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1, 1)))


@pytest.fixture(scope="module")
def case():
    W, H, CELLS, SEED = 160, 120, 12, 20261202
    print(f"case: {W}x{H} cells={CELLS} seed={SEED}", flush=True)
    rgb = dm.generate_image(W, H, CELLS, SEED)
    ref = O.pipeline(rgb, nvec=NVEC, tol=TOL)
    assert ref["rc"] == 0
    Wm = O.intervene(ref["nmax"])
    lsym, d, rs = O.convert_matrix(W, H, Wm)
    rc, v64, u64, _ = O.lanczos_fp64(W, H, lsym, nvec=NVEC, tol=1e-10, maxiter=2000)
    assert rc == 0, "FP64 ground truth did not converge"
    vecs = O.finish_eigvecs(W, H, rs, u64.astype(np.float32))
    spb = O.spectral_pb(v64.astype(np.float32), vecs, W, H)
    _, g64 = O.gpb(ref["combined_g"], spb, ref["nmax"])
    truth = dict(W=Wm, lsym=lsym, rs=rs, vals=v64, u=u64, gpb=g64)
    return W, H, rgb, ref, truth


def check_exact_stages(out, ref, W, H):
    n = W * H
    np.testing.assert_array_equal(host(out["textons"]), ref["textons"])
    np.testing.assert_array_equal(host(out["mpb"])[:, :n].reshape(8, H, W).view(np.uint32),
                                  ref["mpb"].view(np.uint32))
    np.testing.assert_array_equal(host(out["nmax"]).view(np.uint32), ref["nmax"].view(np.uint32))


def test_pipeline_plain(ctx, case):
    W, H, rgb, ref, truth = case
    out = ctx.gpb(ctx.upload(rgb), precision=dm.PREC_FP32, timing=True)
    print("stage ms", [round(v, 3) for v in out["stage_ms"]], "lanczos it", out["lanczos_iterations"], flush=True)
    check_exact_stages(out, ref, W, H)
    ev = out["eigvals"]
    rel = np.abs(ev[1:] - ref["eigvals"][1:]) / np.abs(ref["eigvals"][1:])
    gerr = np.abs(host(out["gpb"]) - ref["gpb"]).max()
    gtruth = np.abs(host(out["gpb"]) - truth["gpb"]).max()
    print(f"eigvals gpu {ev}\n oracle {ref['eigvals']}\n max rel {rel.max():.3e}; gPb max abs vs oracle "
          f"{gerr:.3e}, vs FP64 truth {gtruth:.3e}", flush=True)
    assert out["lanczos_iterations"] % 500 == 0
    assert rel.max() < 1e-4
    assert gerr <= 1e-3


def test_pipeline_mixed(ctx, case):
    W, H, rgb, ref, truth = case
    out = ctx.gpb(ctx.upload(rgb), precision=dm.PREC_MIXED, timing=True)
    print("stage ms", [round(v, 3) for v in out["stage_ms"]], "lanczos it", out["lanczos_iterations"], flush=True)
    check_exact_stages(out, ref, W, H)
    ev = out["eigvals"].astype(np.float64)
    rel = np.abs(ev[1:] - truth["vals"][1:]) / truth["vals"][1:]
    gtruth = np.abs(host(out["gpb"]) - truth["gpb"]).max()
    print(f"eigval max rel vs FP64 {rel.max():.3e}; gPb max abs vs FP64 truth {gtruth:.3e}", flush=True)
    assert rel.max() < 1e-4
    assert gtruth < 1e-3


@pytest.mark.parametrize("precision", [dm.PREC_FP32, dm.PREC_MIXED])
def test_eigensolve_subspace(ctx, case, precision):
    W, H, rgb, ref, truth = case
    n = W * H
    m = torch.from_numpy(np.pad(truth["W"], ((0, 0), (0, dm.pitch_in_floats(n) - n)))).cuda()
    vals, vecs, it = ctx.generalized_eigensolve(W, H, m, nvec=NVEC, tol=TOL, precision=precision)
    rel = np.abs(vals[1:] - truth["vals"][1:]) / truth["vals"][1:]
    # span{1, v_j} is invariant under the per-vector min-max map, so compare
    # the spans of u = D^1/2 v with the D^1/2-weighted constant projected out
    sq = 1.0 / truth["rs"].astype(np.float64)
    G = host(vecs).reshape(NVEC, n).astype(np.float64) * sq[None, :]
    R = truth["u"].astype(np.float64)
    u0 = R[0] / np.linalg.norm(R[0])
    G1 = G[1:] - np.outer(G[1:] @ u0, u0)
    R1 = R[1:] - np.outer(R[1:] @ u0, u0)
    ang = subspace_angle(G1.T, R1.T)
    print(f"precision {precision}: iterations {it}, eigval max rel {rel.max():.3e}, subspace angle {ang:.3e}",
          flush=True)
    if precision == dm.PREC_MIXED:
        assert rel.max() < 1e-4
        assert ang < 1e-4
    else:
        assert rel.max() < 1e-3
        assert ang < 0.2


def test_pipeline_texton64(ctx):
    """texton_choice 2 end to end (plain path) against the oracle's 64-texton
    pipeline: textons, mPb and nonmax exact; eigenvalues and gPb as for 32."""
    W, H, CELLS, SEED = 128, 96, 10, 20261203
    print(f"case: texton64 {W}x{H} cells={CELLS} seed={SEED}", flush=True)
    rgb = dm.generate_image(W, H, CELLS, SEED)
    ref = O.pipeline(rgb, nvec=NVEC, tol=TOL, ntex=64)
    assert ref["rc"] == 0
    out = ctx.gpb(ctx.upload(rgb), precision=dm.PREC_FP32, texton_choice=dm.TEXTON64)
    check_exact_stages(out, ref, W, H)
    assert host(out["textons"]).max() >= 32
    rel = np.abs(out["eigvals"][1:] - ref["eigvals"][1:]) / np.abs(ref["eigvals"][1:])
    gerr = np.abs(host(out["gpb"]) - ref["gpb"]).max()
    print(f"texton64: eigval max rel {rel.max():.3e}, gPb max abs vs oracle {gerr:.3e}", flush=True)
    assert rel.max() < 1e-4
    assert gerr <= 1e-3


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("precision,seed,fixture", [(dm.PREC_MIXED, 20261202, "truth_481x321.npz"),
                                                    (dm.PREC_MIXED, 20261204, "truth_481x321_s20261204.npz")])
def test_eigensolve_481x321_vs_truth(ctx, precision, seed, fixture, solver):
    """The full-size workload (481x321, 656 blocks per Lanczos step) on the
    oracle's own L_sym against committed FP64 truths (shift-invert eigsh,
    tests/golden/make_truth.py), with the mixed path's Lanczos solver
    (solver 0) and its Chebyshev block solver (1).  Seed 20261204 is the
    image on which a Lanczos stop bound of tol * 1e-3 left a subspace error
    of 1.7e-4."""
    import os
    W, H, CELLS, SEED = 481, 321, 40, seed
    print(f"case: eigensolve {W}x{H} precision {precision} seed {seed} solver {solver}", flush=True)
    truth = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", fixture))
    assert list(truth["meta"]) == [W, H, CELLS, SEED, NVEC]
    rgb = dm.generate_image(W, H, CELLS, SEED)
    tex, _ = O.find_textons(O.rgb_to_grey(rgb))
    L, a, b = O.rgb_to_lab(rgb)
    mpb, _ = O.combine(*O.local_cues(L, a, b, tex))
    Wm = O.intervene(O.nonmax(mpb))
    n = W * H
    m = torch.from_numpy(np.pad(Wm, ((0, 0), (0, dm.pitch_in_floats(n) - n)))).cuda()
    with dm.tuned(eigensolver=solver):
        vals, vecs, it = ctx.generalized_eigensolve(W, H, m, nvec=NVEC, tol=TOL, precision=precision)
    tv = truth["eigvals"][:NVEC]
    rel = np.abs(vals[1:] - tv[1:]) / tv[1:]
    sq = truth["sqrt_d"].astype(np.float64)
    u0 = sq / np.linalg.norm(sq)
    G = host(vecs).reshape(NVEC, n)[1:].astype(np.float64) * sq[None, :]
    G -= np.outer(G @ u0, u0)
    R = truth["u"].astype(np.float64)
    R -= np.outer(R @ u0, u0)
    ang = subspace_angle(G.T, R.T)
    print(f"481x321 precision {precision} solver {solver}: iterations {it}, eigval max rel {rel.max():.3e}, "
          f"subspace angle {ang:.3e}", flush=True)
    assert rel.max() < 1e-4
    assert ang < 1e-4


@pytest.mark.parametrize("precision,cap", [(dm.PREC_MIXED, 200), (dm.PREC_FP32, 250)])
def test_iteration_cap_returns_last_check(ctx, case, precision, cap):
    """lanczos.cu:519-560: at MAXITER the reference checks once more and uses
    that check's Ritz pairs whether or not they converged; a lowered cap
    ("lanczos_max_iter") takes the same path: status OK, iteration count =
    the cap, and the last check's values, which by Cauchy interlacing bound
    the converged eigenvalues from above (the j-th smallest screened Ritz
    value is at least the j-th smallest Ritz value, itself at least the j-th
    eigenvalue).  Both precisions on the Lanczos solver."""
    import ctypes

    from damascene_amd._lib import check, lib

    W, H, rgb, ref, truth = case
    print(f"case: iteration cap {cap}, precision {precision}", flush=True)
    with dm.tuned(eigensolver=0):
        full = ctx.gpb(ctx.upload(rgb), precision=precision)
        assert full["lanczos_iterations"] > cap
        with dm.tuned(lanczos_max_iter=cap):
            out = ctx.gpb(ctx.upload(rgb), precision=precision)
    v = ctypes.c_int()
    check(lib().dmsc_tune_get(b"lanczos_max_iter", ctypes.byref(v)))
    assert v.value == 6000
    assert out["lanczos_iterations"] == cap
    ev = out["eigvals"]
    assert np.all(np.isfinite(ev)) and np.all(np.isfinite(host(out["gpb"])))
    print(f"capped eigenvalues {ev} vs converged {full['eigvals']}", flush=True)
    assert np.all(ev[1:] >= full["eigvals"][1:] * (1 - 1e-4))
    assert np.all(np.diff(ev) >= 0)


@pytest.mark.parametrize("solver,bar", [(0, 5e-7), (1, 1e-6)])
def test_large_image_fp64_residuals(ctx, solver, bar):
    """1024x768 (786,432 pixels, 3,072 step tiles), mixed precision: the
    returned eigenpairs checked a posteriori in FP64 against the library's own
    L_sym (u = D^1/2 v with the D^1/2 direction projected out, Rayleigh
    quotient, residual |L u - lam u|): residuals at the solver's floor
    (Lanczos: its tol * 1e-4 bound plus output rounding; the block solver:
    the FP32 storage of its filtered block, about 6e-7) and eigenvalues
    within 1e-5 of their Rayleigh quotients; converged before the cap."""
    import bench as B

    W, H, CELLS, SEED = 1024, 768, 120, 20261202
    print(f"case: {W}x{H} cells={CELLS} seed={SEED} mixed solver {solver}", flush=True)
    with dm.tuned(eigensolver=solver):
        out = ctx.gpb(ctx.upload(dm.generate_image(W, H, CELLS, SEED)), precision=dm.PREC_MIXED)
    chk = B.fp64_residuals(ctx, out, W, H)
    print(f"iterations {out['lanczos_iterations']}, residual max {chk['residual_max']:.2e}, "
          f"eigval vs Rayleigh {chk['eigval_vs_rayleigh_max_rel']:.2e}", flush=True)
    assert out["lanczos_iterations"] < 6000
    assert chk["residual_max"] < bar
    assert chk["eigval_vs_rayleigh_max_rel"] < 1e-5


@pytest.mark.parametrize("w,h,seed", [(780, 1001, 20261202 + 142), (160, 120, 20261202), (61, 37, 7)])
def test_block_solver_large_image(ctx, w, h, seed):
    """The block eigensolver (tune key "eigensolver" = 1, the mixed default)
    against the Lanczos solver (both mixed precision): eigenvalues and gPb on a 780x1001
    image of the uneven batch, whose denser spectrum needs filter degrees of
    about 150, on the 160x120 test image, and on a 61x37 image narrower than
    one 64-column filter tile."""
    print(f"case: block solver {w}x{h} seed {seed}", flush=True)
    img = ctx.upload(dm.generate_image(w, h, max(4, w * h // 4000), seed))
    with dm.tuned(eigensolver=0):
        ref = ctx.gpb(img, precision=dm.PREC_MIXED)
    with dm.tuned(eigensolver=1):
        out = ctx.gpb(img, precision=dm.PREC_MIXED)
        rule, res = ctx.eigensolve_report()
    ev, rv = out["eigvals"][1:].astype(np.float64), ref["eigvals"][1:].astype(np.float64)
    rel = float(np.max(np.abs(ev - rv) / rv))
    g = float(np.abs(out["gpb"].cpu().numpy() - ref["gpb"].cpu().numpy()).max())
    import bench as B
    chk = B.fp64_residuals(ctx, out, w, h)
    print(f"{w}x{h}: Lanczos {ref['lanczos_iterations']} steps, block solver {out['lanczos_iterations']} block "
          f"products (stop rule {rule}, Gram residual {res:.2e}); eigval rel {rel:.2e}, gPb max-abs {g:.2e}; "
          f"FP64 residual max {chk['residual_max']:.2e}, eigval vs Rayleigh {chk['eigval_vs_rayleigh_max_rel']:.2e}",
          flush=True)
    assert rule in (dm.STOP_RESIDUAL, dm.STOP_STEADY, dm.STOP_FLOOR)
    assert rel < 1e-4
    assert g < 1e-3
    # the returned pairs themselves, in FP64 against the library's L_sym: at
    # the mixed bound (tol * 1e-4, or the FP32 floor when that stop fired)
    # plus what FP32 storage of the min-max mapped output vectors leaves
    # (rounding_floor: 61x37 has vectors whose varying part is small next to
    # their constant part, and both solvers' pairs sit at about 5e-6 there)
    bar = 2e-6 if rule == dm.STOP_FLOOR else 1e-6
    ref_chk = B.fp64_residuals(ctx, ref, w, h)
    print(f"FP64 residuals block {chk['residuals']} (output rounding floor {chk['rounding_floor']}); "
          f"Lanczos {ref_chk['residuals']}", flush=True)
    assert chk["residual_max"] < max(bar, 2 * ref_chk["residual_max"])
    assert chk["eigval_vs_rayleigh_max_rel"] < 1e-5


@pytest.mark.parametrize("w,h,seed", [(160, 120, 20261202), (61, 37, 7)])
@pytest.mark.parametrize("shape", [22, 28, 42, 324, 344, 342, 328, 348, 1324, 1328, 1336, 1337, 1344, 1348, 2324, 2328, 2337])
def test_block_solver_shapes(ctx, w, h, seed, shape):
    """Every filter-kernel tile of the block solver ("chebyshev_shape": rows
    per lane x 10 + float4 planes per lane) gives the default tile's (24)
    results bit for bit: each output sums its 81 products in the same order
    (column offset outer, row offset inner) whatever the tile."""
    print(f"case: block solver shape {shape} vs 24, {w}x{h}", flush=True)
    img = ctx.upload(dm.generate_image(w, h, max(4, w * h // 4000), seed))
    outs = {}
    for sh in (24, shape):
        with dm.tuned(eigensolver=1, chebyshev_shape=sh):
            outs[sh] = ctx.gpb(img, precision=dm.PREC_MIXED)
    a, b = outs[24], outs[shape]
    assert a["lanczos_iterations"] == b["lanczos_iterations"]
    assert np.array_equal(a["eigvals"].view(np.uint32), b["eigvals"].view(np.uint32))
    for f in ("eigvecs", "gpb"):
        assert np.array_equal(host(a[f]).view(np.uint32), host(b[f]).view(np.uint32)), f"{f} differs"


@pytest.mark.parametrize("idx", [40, 55])
def test_block_solver_finds_eigenvalue_uniform_start_lanczos_misses(ctx, idx):
    """Images 40 (408x318) and 55 (346x380) of the uneven 64-image batch: the
    Lanczos solver, started like the reference from the uniform vector
    1/sqrt(N) (lanczos.cu:64-72), returns nine eigenpairs that are each
    correct (FP64 residuals at the floor) but skips one of the nine smallest
    eigenvalues, whose eigenvector is nearly orthogonal to the start; the
    block solver (the mixed default) returns the true nine.
    Checked without a stored truth: both sets are eigenpairs, the block
    set is element-wise no larger, and it holds an eigenvalue below the
    Lanczos set's largest that the Lanczos set lacks."""
    import bench as B

    w, h, cells, seed = B.uneven_specs()[idx]
    print(f"case: uneven image {idx}: {w}x{h} cells {cells} seed {seed}", flush=True)
    img = ctx.upload(dm.generate_image(w, h, cells, seed))
    with dm.tuned(eigensolver=0):
        lz = ctx.gpb(img, precision=dm.PREC_MIXED)
    with dm.tuned(eigensolver=1):
        bk = ctx.gpb(img, precision=dm.PREC_MIXED)
    el, eb = lz["eigvals"][1:].astype(np.float64), bk["eigvals"][1:].astype(np.float64)
    rl, rb = B.fp64_residuals(ctx, lz, w, h), B.fp64_residuals(ctx, bk, w, h)
    missing = [float(v) for v in eb if np.min(np.abs(el - v) / v) > 1e-4]
    print(f"Lanczos {np.array2string(el, precision=4)} (FP64 residual {rl['residual_max']:.2e})\n"
          f"block   {np.array2string(eb, precision=4)} (FP64 residual {rb['residual_max']:.2e}); "
          f"missing from the Lanczos set: {missing}", flush=True)
    assert rl["residual_max"] < 1e-6 and rb["residual_max"] < 1e-6
    assert np.all(eb <= el * (1 + 1e-5))
    assert missing and min(missing) < el.max()


@pytest.mark.parametrize("start,first", [(0, 50), (1, 30)])
def test_block_solver_start_and_first_degree(ctx, start, first):
    """The block solver's other starting block (chebyshev_start 0: hashed
    uniform values) and another first filter degree reach the same
    eigenpairs as the defaults (cosine start, degree 50) on the 160x120
    image: eigenvalues to 1e-5 relative, FP64 residuals below 1e-6, gPb to
    1e-3."""
    import bench as B

    w, h = 160, 120
    print(f"case: block solver start {start} first degree {first}, {w}x{h}", flush=True)
    img = ctx.upload(dm.generate_image(w, h, 12, 20261202))
    with dm.tuned(eigensolver=1):
        ref = ctx.gpb(img, precision=dm.PREC_MIXED)
        with dm.tuned(chebyshev_start=start, chebyshev_first_degree=first):
            out = ctx.gpb(img, precision=dm.PREC_MIXED)
    ev, er = out["eigvals"][1:].astype(np.float64), ref["eigvals"][1:].astype(np.float64)
    assert np.max(np.abs(ev - er) / er) < 1e-5
    assert B.fp64_residuals(ctx, out, w, h)["residual_max"] < 1e-6
    assert np.abs(host(out["gpb"]) - host(ref["gpb"])).max() < 1e-3


def test_chebyshev_tile_choice(ctx):
    """dmsc_chebyshev_tile: with the tuning key at 0 the 64 x 21 read-ahead
    tile (2337) where its grid fits in one round of the device's block slots
    and the 64 x 16 one's does not (481x321 on a 256-CU MI355X: 512 against
    672 blocks), else 2328; a set key wins over the size rule."""
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    tile = ctx.chebyshev_tile(481, 321)
    assert tile == (2337 if cus == 256 else tile) and tile in (2337, 2328)
    assert ctx.chebyshev_tile(1024, 768) == 2328
    assert ctx.chebyshev_tile(321, 481) == 2328
    with dm.tuned(chebyshev_shape=1324):
        assert ctx.chebyshev_tile(481, 321) == 1324
