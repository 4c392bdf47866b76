"""Independent pins of oracle functions whose operation order the shared
fixtures leave open (CPU only).

Each check restates what the reference computes from its source, by a
different route than the oracle (numpy in FP64, closed forms, planted
inputs), chosen so that a plausible mistake fails it:

* texton filter bank (texton.cu:39-57): FP64 direct correlation with
  clamp-to-edge borders and the [x][y] coefficient order, all 34 filters;
  a transposed filter, a shifted centre or a mirrored border fails.
* parabola (parabola.cu:75-124, parabola_kernel.cu:1-30): FP64 correlation
  of the symmetrically mirrored gradient planes (gradient.cu:240-241 mirrors
  them with mirrorImage) with the [dy][dx] filters of the same orientation.
* intervening contour (intervening.cu:44-283, 298-475, 477-511): images
  holding straight lines of mPb, where the max-accumulated boundary value
  along any monotone pixel path is known in closed form from the edges
  the reference reads (a horizontal step between columns c and c+1 reads
  column c; a vertical step between rows r and r+1 reads row r; a diagonal
  step reads both), including the corner entries the reference forces to
  the zero-boundary affinity (intervening.cu:459-471) and the 0.5 (a + b)
  symmetrisation.
* rotated chi-squared orientations (rotate.cu:805-870, gradient.cu:210-281):
  step edges at each of the 8 orientations k pi / 8 put the largest
  chi-squared value in the slot the rotation geometry predicts.
* Cullum-Willoughby screening (lanczos.cu:279-384): tridiagonals with
  planted copies, near-copies and spurious values.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O


# ---------------------------------------------------------------- texton bank
def test_texton_convolve_against_fp64_correlation():
    rng = np.random.default_rng(5)
    h, w = 37, 45  # smaller than 2 r + 1 = 19 in neither direction; clamped borders on all four sides
    grey = rng.random((h, w), dtype=np.float32)
    got = O.texton_responses(grey)
    coeffs, radii, nbound = O.texton_filters()
    assert nbound == 9010 - 19 * 19  # the last filter's coefficients are never bound (texton.cu:104-109)
    g = grey.astype(np.float64)
    ci = 0
    for f in range(34):
        r = int(radii[f])
        L = 2 * r + 1
        # coefficient index runs over compX outer, compY inner (texton.cu:47-52)
        K = coeffs[ci:ci + L * L].astype(np.float64).reshape(L, L)  # K[i][j]: x offset i - r, y offset j - r
        if f == 33:
            K = np.zeros_like(K)
        ci += L * L
        pad = np.pad(g, r, mode="edge")  # tex2D clamp-to-edge addressing
        ref = np.zeros((h, w))
        for i in range(L):
            for j in range(L):
                ref += K[i, j] * pad[j:j + h, i:i + w]
        scale = max(1e-3, float(np.abs(ref).max()))
        err = float(np.abs(got[f].astype(np.float64) - ref).max()) / scale
        assert err <= 1e-5, f"filter {f} (radius {r}): max error {err:.2e} of its largest response"
    assert np.all(got[33] == 0)


# ---------------------------------------------------------------- parabola
@pytest.mark.parametrize("radius", [3, 5, 10, 20])
def test_parabola_against_fp64_correlation(radius):
    rng = np.random.default_rng(radius)
    h, w, border = 29, 34, 30
    g8 = rng.random((8, h, w), dtype=np.float32)
    filt = O.parabola_filters(radius)
    out = np.zeros((8, h, w), np.float32)
    O.lib().orc_parabola(w, h, border, radius, O.P(g8), O.P(filt), O.P(out))
    L = 2 * radius + 1
    for o in range(8):
        m = np.pad(g8[o].astype(np.float64), border, mode="symmetric")  # mirrorImage
        F = filt[o].astype(np.float64)  # [v + r][u + r]: row offset v, column offset u
        ref = np.zeros((h, w))
        for v in range(L):
            for u in range(L):
                ref += F[v, u] * m[border - radius + v:border - radius + v + h,
                                   border - radius + u:border - radius + u + w]
        err = float(np.abs(out[o] - ref).max()) / float(np.abs(ref).max())
        assert err <= 1e-5, f"orientation {o}: max error {err:.2e}"


# ---------------------------------------------------------------- intervening contour
def affinity(ic):
    """intervening.cu:28-42 in FP64: exp(-(1 - C(ic)) / sigma), sigma 0.1."""
    val = 1.8682 + (ic / 0.3130) * -1.3113
    c = 1.0 / (1.0 + math.exp(-val))
    return math.exp(-(1.0 - c) * 10.0)


def disc():
    """The radius-5 stencil in scan order: diagonal d <-> offset (dy, dx)."""
    return [(dy, dx) for dy in range(-5, 6) for dx in range(-5, 6) if dy * dy + dx * dx <= 25]


def corner_fixes(w, h):
    """intervening.cu:459-471: (pixel, neighbour) entries set to the affinity
    of a zero boundary whatever was walked (the reference's fixEntry)."""
    fixes = set()
    for x0 in range(w):
        for y0 in range(h):
            rxa, rya, rxb, ryb = max(0, x0 - 5), max(0, y0 - 5), min(x0 + 5, w - 1), min(y0 + 5, h - 1)
            if y0 == rya + 1 and x0 in (rxa, rxb):
                fixes.add(((x0, y0), (x0, y0 - 1)))
            if x0 == rxa + 1 and y0 in (rya, ryb):
                fixes.add(((x0, y0), (x0 - 1, y0)))
            if x0 == rxb - 1 and y0 in (rya, ryb):
                fixes.add(((x0, y0), (x0 + 1, y0)))
            if y0 == ryb - 1 and x0 in (rxa, rxb):
                fixes.add(((x0, y0), (x0, y0 + 1)))
    return fixes


def line_value(lines, a, b):
    """Max boundary value on a monotone pixel path whose coordinate (column
    for vertical lines, row for horizontal ones) runs from a to b: a line at
    position c with value v is read by every path that steps from c to c + 1
    (or back), and by none whose range stays on one side of c; a path that
    ends on c without crossing it may or may not read it (None)."""
    lo, hi = min(a, b), max(a, b)
    best = 0.0
    for c, v in lines:
        if lo <= c < hi:
            best = max(best, v)
        elif hi == c:
            return None
    return best


@pytest.mark.parametrize("kind,lines,n_fix", [
    ("col", [(11, 0.6)], 0),             # one vertical line inside the image
    ("row", [(9, 0.45)], 0),             # one horizontal line
    ("col", [(8, 0.3), (11, 0.7)], 0),   # two lines: the walk keeps the maximum
    ("col", [(0, 0.8)], 4),              # left column: corner entries of pixels (1, 0), (1, h-1)
    ("col", [(22, 0.8)], 4),             # column w - 2: corner entries of (w-2, 0), (w-2, h-1)
    ("row", [(0, 0.8)], 4),              # top row: corner entries of (0, 1), (w-1, 1)
    ("row", [(18, 0.8)], 4),             # row h - 2: corner entries of (0, h-2), (w-1, h-2)
])
def test_intervene_straight_lines(kind, lines, n_fix):
    w, h = 24, 20
    mpb = np.zeros((h, w), np.float32)
    for c, v in lines:
        if kind == "col":
            mpb[:, c] = v
        else:
            mpb[c, :] = v
    M = O.intervene(mpb).reshape(81, h, w).astype(np.float64)
    flat = O.intervene(np.zeros((h, w), np.float32)).reshape(81, h, w)
    fixes = corner_fixes(w, h)
    f0 = affinity(0.0)
    checked = fixed_checked = 0
    for d, (dy, dx) in enumerate(disc()):
        for y0 in range(h):
            for x0 in range(w):
                x1, y1 = x0 + dx, y0 + dy
                if not (0 <= x1 < w and 0 <= y1 < h):
                    assert M[d, y0, x0] == 0.0  # neighbours outside the image are never written
                    continue
                if (dy, dx) == (0, 0):
                    assert M[d, y0, x0] == 1.0
                    continue
                if abs(flat[d, y0, x0] - f0) > 2e-6 * f0:
                    continue  # an entry no walk reaches in either direction keeps less than f(0)
                ic = line_value(lines, x0, x1) if kind == "col" else line_value(lines, y0, y1)
                if ic is None:
                    continue
                fwd = f0 if ((x0, y0), (x1, y1)) in fixes else affinity(ic)
                bwd = f0 if ((x1, y1), (x0, y0)) in fixes else affinity(ic)
                want = 0.5 * (fwd + bwd)  # symmetrizeMatrix
                assert abs(M[d, y0, x0] - want) <= 2e-6 * want, \
                    f"pixel ({x0},{y0}) neighbour ({x1},{y1}): {M[d, y0, x0]:.7f} vs {want:.7f}"
                checked += 1
                if fwd != bwd:
                    fixed_checked += 1
    assert checked > 20000
    # each forced entry is compared twice (both symmetric slots), and only
    # where the line makes it differ from the walked value
    assert fixed_checked == n_fix


# ---------------------------------------------------------------- rotated chi-squared
@pytest.mark.parametrize("k", range(8))
def test_rotated_gradients_peak_in_predicted_slot(k):
    """A step edge at angle k pi / 8 (counter-clockwise from the x axis, y
    up) through the image centre.  The rotation tables map image direction
    (dx, dy) (dy down) to (dx - t dy, t dx + dy) with t = tan(o pi / 8)
    (rotate.cu:815-826; the -3 pi / 8 table turns by pi / 2 first and gives
    the same rotation by 3 pi / 8), so the rotated frame's rows run along
    image angle o pi / 8: slot o + 4 (the top/bottom split, gradient.cu:240)
    responds to edges at o pi / 8 and slot o (the left/right split) to edges
    at pi / 2 + o pi / 8.  Edge angle k pi / 8 therefore peaks in slot
    (k + 4) mod 8, at every pixel along the edge and every scale."""
    h = w = 97
    a = k * math.pi / 8
    nx, ny = math.sin(a), math.cos(a)  # normal of edge direction (cos a, -sin a) in (x, row) coordinates
    ys, xs = np.mgrid[0:h, 0:w]
    side = nx * (xs - 48) + ny * (ys - 48)
    labels = (side > 0.25).astype(np.int32)
    g = O.channel_gradients(labels, 32, 0, 0.0, (5, 10, 20))  # [3 scales][8 slots][h][w]
    want = (k + 4) % 8
    near = (np.abs(side - 0.25) < 0.75) & (np.hypot(xs - 48, ys - 48) < 12)
    assert near.sum() >= 10
    for s in range(3):
        slots = g[s][:, near]  # [8][pixels]
        peaks = np.bincount(slots.argmax(axis=0), minlength=8)
        # the mean over the edge pixels peaks in the predicted slot, and so do
        # nearly all pixels singly (a pixel whose half-discs split the edge's
        # staircase unevenly may tip to a neighbouring slot at radius 5)
        assert int(slots.mean(axis=1).argmax()) == want, f"scale {s}: mean peaks {slots.mean(axis=1)}, want {want}"
        assert peaks[want] >= 0.9 * near.sum(), f"scale {s}: peaks in slots {peaks}, want {want}"
        assert float(np.median(slots[want])) > 4.0  # a clean edge: chi-squared near its maximum 8


# ---------------------------------------------------------------- Cullum screening
def screen(alpha, beta, eig_check=30):
    alpha = np.asarray(alpha, np.float32)
    beta = np.asarray(beta, np.float32)
    vals = np.zeros(eig_check, np.float64)
    n = O.lib().orc_cullum_screen(len(alpha) - 1, O.P(alpha), O.P(beta), eig_check, O.P(vals))
    assert n >= 0
    return vals[:n]


def test_cullum_screening_planted():
    # T = [a0] (+) T-hat when beta_0 = 0: every eigenvalue of T-hat is also one of T
    # and spurious by the T-hat test, so only a0 survives
    that = [0.1, 0.3, 0.5, 0.7]
    np.testing.assert_allclose(screen([0.2] + that, [0.0] * 5), [np.float32(0.2)], rtol=1e-12)
    # a copy closer than LUMPTOL = 1e-5 is lumped into the first value, which is then a
    # duplicate and exempt from the spurious test; a near-copy 2e-5 away is neither
    that = [0.1, np.float32(0.1) + np.float32(5e-6), 0.3, 0.3 + 2e-5, 0.5]
    got = screen([0.2] + that, [0.0] * 6)
    np.testing.assert_allclose(got, [np.float32(0.1), np.float32(0.2)], rtol=1e-12)
    # coupled first row: values of T away from T-hat's survive
    rng = np.random.default_rng(3)
    a = rng.random(40).astype(np.float32)
    b = (0.3 * rng.random(40)).astype(np.float32)
    T = np.diag(a.astype(np.float64)) + np.diag(b[:39].astype(np.float64), 1) + np.diag(b[:39].astype(np.float64), -1)
    ev = np.linalg.eigvalsh(T)
    evh = np.linalg.eigvalsh(T[1:, 1:])
    got = screen(a, b, eig_check=40)
    acc, dup = [ev[0]], [False]  # lump first (marking the kept value of a lump) ...
    for p, v in zip(ev, ev[1:]):
        if v - p < 1e-5:
            dup[-1] = True
        else:
            acc.append(v)
            dup.append(False)
    # ... then drop non-duplicates within 1e-10 (relative) of an eigenvalue of T-hat
    want = [v for v, dp in zip(acc, dup) if dp or np.min(np.abs(v - evh) / abs(v)) > 1e-10]
    np.testing.assert_allclose(got, want, rtol=1e-9)
    # only the first eig_check values of T take part
    got5 = screen(a, b, eig_check=5)
    assert len(got5) <= 5 and np.all(got5 <= ev[4] * (1 + 1e-12))


# ---------------------------------------------------------------- timing mode
def test_simd_timing_mode_is_bit_identical():
    """The oracle's SIMD timing mode (register-tiled correlations and SpMV,
    used by the benchmark's CPU baseline) runs each output's operations in
    the scalar loops' order: texton responses, parabola, spectral Pb and the
    SpMV equal the scalar results bit for bit, including the tile edges
    (sizes not multiples of the tile)."""
    rng = np.random.default_rng(11)
    h, w = 45, 71
    grey = rng.random((h, w), dtype=np.float32)
    g8 = rng.random((8, h, w), dtype=np.float32)
    vecs = rng.random((9, h, w), dtype=np.float32)
    vals = np.sort(rng.random(9)).astype(np.float32) + np.float32(1e-3)
    A = rng.standard_normal((81, h * w)).astype(np.float32)
    x = rng.standard_normal(h * w).astype(np.float32)
    outs = []
    for simd in (0, 1):
        O.lib().orc_set_simd(simd)
        try:
            par = np.zeros((8, h, w), np.float32)
            O.lib().orc_parabola(w, h, 30, 10, O.P(g8), O.P(O.parabola_filters(10)), O.P(par))
            outs.append([O.texton_responses(grey), par, O.spectral_pb(vals, vecs, w, h), O.spmv(w, h, A, x)])
        finally:
            O.lib().orc_set_simd(0)
    for name, a, b in zip(("texton", "parabola", "spectral Pb", "SpMV"), *outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{name} differs in the SIMD mode"
