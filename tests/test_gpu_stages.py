"""Stage-by-stage parity of the HIP path (through the C ABI) against the CPU
oracle on the same seeded synthetic image.

Integer and byte work (textons, quantised labels, histograms via the
chi-squared planes, nonmax masks) and every FP32 stage whose operation order
is fixed by contract (colour, filter bank, local cues, combine, matrix
transform, SpMV, spectralPb, gPb) must match bit for bit.  The intervening
contour's affinity matrix is bit-exact too: kernels and oracle share one
project exp built from basic IEEE operations in place of the device expf.
"""
import ctypes

import numpy as np
import pytest
import torch

import damascene_amd as dm
from damascene_amd._lib import check, lib
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

W, H, CELLS, SEED = 160, 120, 12, 20261202


def host(t):
    return t.detach().cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(gpu, ref, what):
    g, r = bits(gpu), bits(ref)
    bad = np.count_nonzero(g != r)
    if bad:
        gf, rf = np.asarray(gpu, np.float32), np.asarray(ref, np.float32)
        idx = np.argwhere(g != r)[0]
        raise AssertionError(f"{what}: {bad} of {g.size} values differ; first at {tuple(idx)}: "
                             f"gpu {gf[tuple(idx)]!r} oracle {rf[tuple(idx)]!r}; "
                             f"max abs diff {np.nanmax(np.abs(gf - rf)):.3e}")


@pytest.fixture(scope="module")
def image():
    print(f"case: {W}x{H} cells={CELLS} seed={SEED}", flush=True)
    return dm.generate_image(W, H, CELLS, SEED)


@pytest.fixture(scope="module")
def colour(ctx, image):
    d_rgb = ctx.upload(image)
    grey = ctx.rgb_to_grey(d_rgb)
    L, a, b = ctx.rgb_to_lab(d_rgb)
    return d_rgb, grey, (L, a, b)


def test_colour(colour, image):
    _, grey, (L, a, b) = colour
    assert_bitexact(host(grey), O.rgb_to_grey(image), "grey")
    oL, oa, ob = O.rgb_to_lab(image)
    assert_bitexact(host(L), oL, "L")
    assert_bitexact(host(a), oa, "a")
    assert_bitexact(host(b), ob, "b")


@pytest.fixture(scope="module")
def responses(ctx, colour):
    return ctx.texton_responses(colour[1])


def test_texton_responses(responses, colour):
    ref = O.texton_responses(host(colour[1]))
    assert_bitexact(host(responses), ref, "texton responses")
    assert np.all(host(responses)[33] == 0)


def test_kmeans_labels(ctx, responses):
    labels, it = ctx.kmeans(responses)
    ref, rit = O.kmeans(host(responses))
    assert it == rit, f"iterations gpu {it} oracle {rit}"
    np.testing.assert_array_equal(host(labels), ref)


def test_kmeans_labels_mixed(ctx, responses):
    """BF16 MFMA distances with the certified-margin test and FP32 re-check
    of near ties: the same labels and iteration count as the exact path."""
    labels, it = ctx.kmeans(responses, precision=dm.PREC_MIXED)
    rech = ctypes.c_int()
    check(lib().dmsc_kmeans_rechecks(ctx.handle, ctypes.byref(rech)))
    n = labels.numel()
    print(f"mixed k-means: {it} iterations, {rech.value} exact re-checks of {it * n} assignments "
          f"({rech.value / (it * n):.2%})", flush=True)
    ref, rit = O.kmeans(host(responses))
    assert it == rit, f"iterations gpu {it} oracle {rit}"
    np.testing.assert_array_equal(host(labels), ref)
    assert 0 <= rech.value <= it * n


@pytest.fixture(scope="module")
def textons(ctx, colour):
    lab, _ = ctx.find_textons(colour[1])
    return lab


@pytest.fixture(scope="module")
def cues(ctx, colour, textons):
    L, a, b = colour[2]
    return ctx.local_cues(L, a, b, textons, raw=True)


def test_channel_gradients(cues, colour, textons):
    (_, raw) = cues
    L, a, b = [host(t) for t in colour[2]]
    raw = host(raw).reshape(4, 3, 8, H, W)
    q = [np.minimum((np.floor(c * np.float32(25))).astype(np.int32), 24) for c in (L, a, b)]
    specs = [(q[0], 25, 1, 2.0, (3, 5, 10)), (q[1], 25, 1, 1.0, (5, 10, 20)), (q[2], 25, 1, 1.0, (5, 10, 20)),
             (host(textons), 32, 0, 0.0, (5, 10, 20))]
    for ch, (lab, nb, blur, sig, radii) in enumerate(specs):
        ref = O.channel_gradients(lab, nb, blur, sig, radii)
        assert_bitexact(raw[ch], ref, f"raw chi-squared planes, cue {ch}")


def test_local_cues(cues, colour, textons):
    (got, _) = cues
    L, a, b = [host(t) for t in colour[2]]
    ref = O.local_cues(L, a, b, host(textons))
    for name, g, r in zip(("bg", "cga", "cgb", "tg"), got, ref):
        assert_bitexact(host(g).reshape(24, H, W), r, name)


@pytest.fixture(scope="module")
def combined(ctx, cues):
    (got, _) = cues
    full = [torch.nn.functional.pad(c, (0, dm.pitch_in_floats(W * H) - W * H)).contiguous() for c in got]
    return ctx.combine(W, H, *full), got


def test_combine_and_nonmax(ctx, combined):
    (mpb, cg), got = combined
    n = W * H
    ref_m, ref_c = O.combine(*[host(g).reshape(24, H, W) for g in got])
    assert_bitexact(host(mpb)[:, :n].reshape(8, H, W), ref_m, "mPb")
    assert_bitexact(host(cg)[:, :n].reshape(8, H, W), ref_c, "combinedG")
    nm = ctx.non_max_suppression(W, H, mpb)
    assert_bitexact(host(nm), O.nonmax(ref_m), "nonmax")


def test_intervene_convert_spmv(ctx, combined):
    (mpb, _), _ = combined
    nm = ctx.non_max_suppression(W, H, mpb)
    m = ctx.intervene(nm)
    n = W * H
    ref = O.intervene(host(nm))
    got = host(m)[:, :n]
    assert_bitexact(got, ref, "affinity matrix")
    # transform and SpMV from the GPU's own matrix: bit-exact
    src = got.copy()
    d, rs = ctx.convert_matrix(W, H, m)
    lsym, od, ors = O.convert_matrix(W, H, src)
    assert_bitexact(host(rs).ravel(), ors, "D^-1/2")
    assert_bitexact(host(m)[:, :n], lsym, "L_sym")
    x = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    y = ctx.spmv(W, H, m, torch.from_numpy(x).cuda())
    assert_bitexact(host(y).ravel(), O.spmv(W, H, lsym, x), "SpMV")


def test_spectral_pb_and_gpb(ctx, combined):
    (mpb, cg), _ = combined
    n = W * H
    rng = np.random.default_rng(7)
    vecs = rng.random((9, H, W), dtype=np.float32)
    vals = np.sort(rng.random(9).astype(np.float32)) * 1e-2 + 1e-4
    spb = ctx.spectral_pb(vals, torch.from_numpy(vecs).cuda())
    ref = O.spectral_pb(vals, vecs.reshape(9, n), W, H)
    assert_bitexact(host(spb)[:, :n].reshape(8, H, W), ref, "sPb")
    nm = ctx.non_max_suppression(W, H, mpb)
    all_, g = ctx.calc_gpb(W, H, cg, spb, nm)
    ra, rg = O.gpb(host(cg)[:, :n].reshape(8, H, W).copy(), ref, host(nm))
    assert_bitexact(host(all_)[:, :n].reshape(8, H, W), ra, "gPbAll")
    assert_bitexact(host(g), rg, "gPb")


def test_postprocess_and_normalize(ctx, combined):
    """skeleton.cu PostProcess and NormalizeGpbAll against the oracle: the
    thinned mask and iteration count exactly, values within 1e-6 (device
    expf in the sigmoid)."""
    (mpb, cg), _ = combined
    n = W * H
    nm = ctx.non_max_suppression(W, H, mpb)
    rng = np.random.default_rng(11)
    gpb = torch.from_numpy(rng.random((H, W), dtype=np.float32) * 1.2).cuda()
    g, thin, it = ctx.postprocess(gpb, nm)
    rg, rthin, rit = O.postprocess(host(gpb), host(nm))
    assert it == rit, f"iterations gpu {it} oracle {rit}"
    off = np.float32((1 / (1 + np.exp(np.float32(2.6433))) - np.float32(0.0667)) / np.float32(0.9333))
    np.testing.assert_array_equal(np.abs(host(thin) - off) > 1e-5, np.abs(rthin - off) > 1e-5)
    np.testing.assert_allclose(host(thin), rthin, atol=1e-6, rtol=0)
    np.testing.assert_allclose(host(g), rg, atol=1e-6, rtol=0)
    all_ = torch.from_numpy(np.pad(rng.random((8, n), dtype=np.float32) * 1.5 - 0.2,
                                   ((0, 0), (0, dm.pitch_in_floats(n) - n)))).cuda()
    got = ctx.normalize_gpb_all(all_, n)
    ref = O.normalize_gpb_all(host(all_)[:, :n].copy())
    np.testing.assert_allclose(host(got)[:, :n], ref, atol=1e-6, rtol=0)


# ---------------------------------------------------------------- 64 textons
# texton_choice 2 (texton.cu:65-67, kmeans.cu:37-55/106-153/278-286,
# rotate.cu:780-782, combine.cu:16,20): same bar as the 32-texton stages.

@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_kmeans64_labels(ctx, responses, precision):
    prec = dm.PREC_FP32 if precision == "fp32" else dm.PREC_MIXED
    print(f"case: kmeans k=64 {precision}", flush=True)
    labels, it = ctx.kmeans(responses, precision=prec, k=64)
    ref, rit = O.kmeans(host(responses), k=64)
    assert it == rit, f"iterations gpu {it} oracle {rit}"
    np.testing.assert_array_equal(host(labels), ref)
    assert host(labels).max() >= 32  # the upper half of the clusters is in use


@pytest.fixture(scope="module")
def textons64(ctx, colour):
    lab, _ = ctx.find_textons(colour[1], texton_choice=dm.TEXTON64)
    return lab


def test_find_textons64(textons64, colour):
    ref, _ = O.find_textons(host(colour[1]), ntex=64)
    np.testing.assert_array_equal(host(textons64), ref)


@pytest.fixture(scope="module")
def cues64(ctx, colour, textons64):
    L, a, b = colour[2]
    return ctx.local_cues(L, a, b, textons64, raw=True, texton_choice=dm.TEXTON64)


def test_texture_gradients64(cues64, textons64):
    (_, raw) = cues64
    raw = host(raw).reshape(4, 3, 8, H, W)
    ref = O.channel_gradients(host(textons64), 64, 0, 0.0, (5, 10, 20))
    assert_bitexact(raw[3], ref, "raw chi-squared planes, 64 textons")


def test_local_cues_and_combine64(ctx, cues64, colour, textons64):
    (got, _) = cues64
    L, a, b = [host(t) for t in colour[2]]
    ref = O.local_cues(L, a, b, host(textons64), ntex=64)
    for name, g, r in zip(("bg", "cga", "cgb", "tg"), got, ref):
        assert_bitexact(host(g).reshape(24, H, W), r, name + " (64 textons)")
    full = [torch.nn.functional.pad(c, (0, dm.pitch_in_floats(W * H) - W * H)).contiguous() for c in got]
    mpb, cg = ctx.combine(W, H, *full, texton_choice=dm.TEXTON64)
    n = W * H
    ref_m, ref_c = O.combine(*ref, ntex=64)
    assert_bitexact(host(mpb)[:, :n].reshape(8, H, W), ref_m, "mPb (64 textons)")
    assert_bitexact(host(cg)[:, :n].reshape(8, H, W), ref_c, "combinedG (64 textons)")
