"""Python binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Never used by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_host.so")

_o = None


def build(ref=False):
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            build()
        _o = ctypes.CDLL(LIB)
    return _o


def ref_lib():
    """The reference's own host code compiled into oracle/_ref (None if absent)."""
    if not os.path.exists(REF_LIB):
        return None
    return ctypes.CDLL(REF_LIB)


def P(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f32(*shape):
    return np.zeros(shape, np.float32)


# ------------------------------------------------------------------ tables
def texton_filters():
    c, r = f32(9010), np.zeros(34, np.int32)
    nb = lib().orc_texton_filters(P(c), P(r))
    return c, r, nb


def parabola_filters(radius):
    L = 2 * radius + 1
    out = f32(8, L, L)
    lib().orc_parabola_filters(radius, P(out))
    return out


def spb_filters():
    out = f32(8, 49)
    lib().orc_spb_filters(P(out))
    return out


def rotation_tables(width, height, theta_pi):
    n = max(width, height) * 4 + 16
    bx, by, gx, gy = [np.zeros(n, np.int32) for _ in range(4)]
    nw, nh = ctypes.c_int(), ctypes.c_int()
    lib().orc_rotation_tables(width, height, ctypes.c_float(theta_pi), ctypes.byref(nw), ctypes.byref(nh),
                              P(bx), P(by), P(gx), P(gy))
    return nw.value, nh.value, bx[:width], by[:width], gx[:height], gy[:height]


def blur_kernel(sigma):
    k = f32(64)
    r = ctypes.c_int()
    n = lib().orc_blur_kernel(ctypes.c_float(sigma), P(k), ctypes.byref(r))
    return k[:n], r.value


def stencil(radius, width):
    offs, diag = np.zeros(4 * radius * radius + 16, np.int32), np.zeros((2 * radius + 1) ** 2, np.int32)
    n = lib().orc_stencil(radius, width, P(offs), P(diag))
    return offs[:n], diag


def gamma_lut(gamma=2.5):
    out = f32(256)
    lib().orc_gamma_lut(ctypes.c_float(gamma), P(out))
    return out


def cbrtf(x):
    o = lib()
    o.orc_cbrtf.restype = ctypes.c_float
    o.orc_cbrtf.argtypes = [ctypes.c_float]
    return np.array([o.orc_cbrtf(float(v)) for v in np.asarray(x, np.float32).ravel()], np.float32)


def expf(x):
    o = lib()
    o.orc_expf.restype = ctypes.c_float
    o.orc_expf.argtypes = [ctypes.c_float]
    return np.array([o.orc_expf(float(v)) for v in np.asarray(x, np.float32).ravel()], np.float32)


# ------------------------------------------------------------------ stages
def rgb_to_grey(rgb):
    h, w = rgb.shape
    out = f32(h, w)
    lib().orc_rgb_to_grey(w, h, P(np.ascontiguousarray(rgb, np.uint32)), P(out))
    return out


def rgb_to_lab(rgb, gamma=2.5, normalize=True):
    h, w = rgb.shape
    L, a, b = f32(h, w), f32(h, w), f32(h, w)
    lib().orc_rgb_to_lab(w, h, ctypes.c_float(gamma), P(np.ascontiguousarray(rgb, np.uint32)), P(L), P(a), P(b))
    if normalize:
        lib().orc_normalize_lab(w, h, P(L), P(a), P(b))
    return L, a, b


def texton_responses(grey):
    h, w = grey.shape
    out = f32(34, h, w)
    lib().orc_texton_convolve(w, h, P(np.ascontiguousarray(grey, np.float32)), P(out))
    return out


def kmeans(resp, max_iter=15, k=32):
    _, h, w = resp.shape
    labels = np.zeros((h, w), np.int32)
    it = ctypes.c_int()
    lib().orc_kmeans(w, h, k, 34, P(np.ascontiguousarray(resp, np.float32)), P(labels), max_iter, ctypes.byref(it))
    return labels, it.value


def find_textons(grey, ntex=32):
    h, w = grey.shape
    labels = np.zeros((h, w), np.int32)
    it = ctypes.c_int()
    lib().orc_find_textons(w, h, P(np.ascontiguousarray(grey, np.float32)), ntex, P(labels), ctypes.byref(it))
    return labels, it.value


def integral_images(labels, nbins):
    H, W = labels.shape
    out = np.zeros((H, W, nbins), np.int32)
    lib().orc_integral_images(W, H, nbins, P(np.ascontiguousarray(labels, np.int32)), nbins, P(out))
    return out


def channel_gradients(labels, nbins, blur, sigma, radii3):
    h, w = labels.shape
    out = f32(3, 8, h, w)
    r = np.array(radii3, np.int32)
    lib().orc_channel_gradients(w, h, nbins, int(blur), ctypes.c_float(sigma), P(r),
                                P(np.ascontiguousarray(labels, np.int32)), P(out))
    return out


def local_cues(L, a, b, textons, ntex=32):
    h, w = L.shape
    outs = [f32(24, h, w) for _ in range(4)]
    lib().orc_local_cues(w, h, P(L), P(a), P(b), P(np.ascontiguousarray(textons, np.int32)), ntex,
                         *[P(o) for o in outs])
    return outs


def combine(bg, cga, cgb, tg, ntex=32):
    _, h, w = bg.shape
    mpb, cg = f32(8, h, w), f32(8, h, w)
    lib().orc_combine(w, h, P(bg), P(cga), P(cgb), P(tg), ntex, P(mpb), P(cg))
    return mpb, cg


def nonmax(pb8):
    _, h, w = pb8.shape
    out = f32(h, w)
    lib().orc_nonmax(w, h, P(np.ascontiguousarray(pb8, np.float32)), P(out))
    return out


def intervene(mpb, sigma=0.1):
    h, w = mpb.shape
    m = f32(81, h * w)
    lib().orc_intervene(w, h, P(np.ascontiguousarray(mpb, np.float32)), P(m), ctypes.c_float(sigma))
    return m


def convert_matrix(w, h, m):
    m = np.array(m, np.float32, copy=True)
    d, rs = f32(h * w), f32(h * w)
    lib().orc_convert_matrix(w, h, P(m), P(d), P(rs))
    return m, d, rs


def spmv(w, h, m, x):
    y = f32(h * w)
    lib().orc_spmv(w, h, P(np.ascontiguousarray(m, np.float32)), P(np.ascontiguousarray(x, np.float32)), P(y))
    return y


def tridiag_eigvals(alpha, beta, il, iu):
    a = np.ascontiguousarray(alpha, np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    w = np.zeros(iu - il + 1)
    lib().orc_tridiag_eigvals(len(a), P(a), P(b), il, iu, P(w))
    return w


def tridiag_eigvec(alpha, beta, lam):
    a = np.ascontiguousarray(alpha, np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    v = np.zeros(len(a))
    lib().orc_tridiag_eigvec.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib().orc_tridiag_eigvec(len(a), P(a), P(b), lam, P(v), None, 0)
    return v


def lanczos_fp32(w, h, lsym, nvec=9, tol=1e-3):
    vals, vecs = f32(nvec), f32(nvec, h * w)
    it = ctypes.c_int()
    rc = lib().orc_lanczos_fp32(w, h, P(np.ascontiguousarray(lsym, np.float32)), nvec, ctypes.c_float(tol), P(vals),
                                P(vecs), ctypes.byref(it))
    return rc, vals, vecs, it.value


def lanczos_fp64(w, h, lsym, nvec=9, tol=1e-10, maxiter=3000):
    vals, vecs = np.zeros(nvec), np.zeros((nvec, h * w))
    it = ctypes.c_int()
    lib().orc_lanczos_fp64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rc = lib().orc_lanczos_fp64(w, h, P(np.ascontiguousarray(lsym, np.float32)), nvec, tol, maxiter, P(vals), P(vecs),
                                ctypes.byref(it))
    return rc, vals, vecs, it.value


def finish_eigvecs(w, h, rs, vecs):
    v = np.array(vecs, np.float32, copy=True)
    lib().orc_finish_eigvecs(w, h, v.shape[0], P(np.ascontiguousarray(rs, np.float32)), P(v))
    return v


def generalized_eigensolve(w, h, W, nvec=9, tol=1e-3):
    vals, vecs = f32(nvec), f32(nvec, h * w)
    it = ctypes.c_int()
    rc = lib().orc_generalized_eigensolve(w, h, P(np.ascontiguousarray(W, np.float32)), nvec, ctypes.c_float(tol),
                                          P(vals), P(vecs), ctypes.byref(it))
    return rc, vals, vecs, it.value


def spectral_pb(eigvals, eigvecs, w, h):
    nvec = eigvecs.shape[0]
    out = f32(8, h, w)
    lib().orc_spectral_pb(w, h, nvec, P(np.ascontiguousarray(eigvals, np.float32)),
                          P(np.ascontiguousarray(eigvecs, np.float32)), P(out))
    return out


def gpb(cg, spb, nmax):
    _, h, w = cg.shape
    all_, g = f32(8, h, w), f32(h, w)
    lib().orc_gpb(w, h, P(cg), P(spb), P(np.ascontiguousarray(nmax, np.float32)), P(all_), P(g))
    return all_, g


def pipeline(rgb, nvec=9, tol=1e-3, ntex=32):
    """Whole plain-FP32 pipeline; returns a dict of host arrays and stage ms."""
    h, w = rgb.shape
    r = dict(textons=np.zeros((h, w), np.int32), mpb=f32(8, h, w), combined_g=f32(8, h, w), nmax=f32(h, w),
             eigvals=f32(nvec), eigvecs=f32(nvec, h, w), spb=f32(8, h, w), gpb_all=f32(8, h, w), gpb=f32(h, w))
    ms = np.zeros(6)
    rc = lib().orc_gpb_pipeline(w, h, P(np.ascontiguousarray(rgb, np.uint32)), nvec, ctypes.c_float(tol),
                                P(r["textons"]), P(r["mpb"]), P(r["combined_g"]), P(r["nmax"]), P(r["eigvals"]),
                                P(r["eigvecs"]), P(r["spb"]), P(r["gpb_all"]), P(r["gpb"]), P(ms), ntex)
    r["rc"] = rc
    r["stage_ms"] = ms
    return r


def skeleton_pass(img, lut):
    h, w = img.shape
    out = f32(h, w)
    removed = lib().orc_skeleton_pass(w, h, lut, P(np.ascontiguousarray(img, np.float32)), P(out))
    return out, removed


def postprocess(gpb, nmax):
    """Returns (sigmoid(gPb), thinned sigmoid(gPb), iterations)."""
    h, w = nmax.shape
    g = np.array(gpb, np.float32, copy=True).reshape(h, w)
    thin = f32(h, w)
    it = lib().orc_postprocess(w, h, P(g), P(np.ascontiguousarray(nmax, np.float32)), P(thin))
    return g, thin, it


def normalize_gpb_all(all_):
    a = np.array(all_, np.float32, copy=True)
    lib().orc_normalize_gpb_all(a.shape[-1] if a.ndim == 2 else a[0].size, a.shape[0], a[0].size, P(a))
    return a
