"""The FP64 reference of the differentiable spectral Pb + combine
(oracle/grad_ref.py), pinned on the CPU before it checks the GPU backward.

* Forward: with the reference's weights (combine table, 0.0139,
  1/sqrt(lambda_i)) the FP64 forward reproduces the oracle pipeline's
  pre-mask gPb (max over orientations of its gPbAll) to FP32 rounding; the
  oracle pipeline itself is pinned against the reference's golden vectors.
* Gradient: automatic differentiation in FP64 against central finite
  differences with h = 1e-3 in FP64, for every parameter and for
  eigenvector entries in the interior and on the mirrored borders (where
  the adjoint of the 3-pixel reflection folds contributions back).  The
  cue weights and gamma get looser bounds: a step of 1e-3 on weights of
  about 0.01 moves some pixels' maximising orientation, a kink finite
  differences straddle; a dropped term or a wrong sign misses by far more.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import grad_ref as G  # noqa: E402
from oracle import pyoracle as O  # noqa: E402


def test_forward_matches_oracle_pipeline():
    import damascene_amd as dm

    w, h = 48, 40
    rgb = dm.generate_image(w, h, 6, 20261202)
    ref = O.pipeline(rgb, nvec=9, tol=1e-3)
    assert ref["rc"] == 0
    L, a, b = O.rgb_to_lab(rgb)
    tex, _ = O.find_textons(O.rgb_to_grey(rgb))
    cues = O.local_cues(L, a, b, tex)
    C = G.cue_planes(*[c.reshape(24, -1) for c in cues], w, h)
    # combinedG weights of the 32-texton table (combine.cu:21), as the oracle pipeline uses them
    wts = np.array([0, 0, .0028, .0041, .0042, .0047, .0033, .0033, .0035, .0035, .0035, .0188], np.float32)
    ev = ref["eigvals"]
    eta = np.zeros(9, np.float32)
    eta[1:] = np.float32(1) / np.sqrt(ev[1:]).astype(np.float32)
    prm = np.concatenate([wts, np.array([0.0139], np.float32), eta]).astype(np.float64)
    pre = G.forward64(torch.as_tensor(prm), torch.as_tensor(ref["eigvecs"].astype(np.float64)), C, w, h).numpy()
    want = ref["gpb_all"].max(axis=0)
    err = np.abs(pre - want).max()
    assert err < 1e-5 * max(1.0, np.abs(want).max()), f"FP64 forward vs oracle gPbAll max: {err:.2e}"


def test_gradient_against_finite_differences():
    w, h, nvec = 40, 32, 9
    rng = np.random.default_rng(5)
    C = torch.as_tensor(rng.random((12, 8, h, w)))
    v = rng.standard_normal((nvec, h, w)) * 0.05
    prm = np.concatenate([rng.random(12) * 0.02, [0.0139], [0.0], 30 + rng.random(nvec - 1) * 20])
    mask = rng.random((h, w)) < 0.2
    _, gp, gv = G.loss_and_grads(prm, v, C, mask, w, h)
    step = 1e-3

    def fd_param(j):
        a, b = prm.copy(), prm.copy()
        a[j] += step
        b[j] -= step
        return (G.loss_only(a, v, C, mask, w, h) - G.loss_only(b, v, C, mask, w, h)) / (2 * step)

    fd = np.array([fd_param(j) for j in range(len(prm))])
    np.testing.assert_allclose(gp[:12], fd[:12], rtol=1e-2)  # cue weights (argmax kinks)
    np.testing.assert_allclose(gp[12], fd[12], rtol=1e-3)  # gamma: a 7 % step, kinks too
    np.testing.assert_allclose(gp[13:], fd[13:], rtol=1e-6, atol=1e-12)  # eta
    assert gp[13] == 0.0
    for (i, y, x) in [(1, 3, 4), (2, 0, 0), (8, h - 1, w - 1), (4, 10, 20), (5, 2, w - 2), (3, h - 3, 1)]:
        a, b = v.copy(), v.copy()
        a[i, y, x] += step
        b[i, y, x] -= step
        f = (G.loss_only(prm, a, C, mask, w, h) - G.loss_only(prm, b, C, mask, w, h)) / (2 * step)
        assert abs(gv[i, y, x] - f) <= 1e-5 * abs(f) + 1e-12, f"eigvec {(i, y, x)}: {gv[i, y, x]} vs {f}"
    assert np.all(gv[0] == 0)
