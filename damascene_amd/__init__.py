"""damascene_amd -- MI355X-native gPb (global probability of boundary).

Python host side over the C ABI of libdmsc.so (include/dmsc.h).  Device
memory is held in torch tensors (plumbing only); every stage runs in the
library's HIP kernels.  Stage functions mirror the reference's stage API
(bryancatanzaro/damascene include/*.h): find_textons, kmeans, local_cues,
combine, non_max_suppression, intervene, generalized_eigensolve,
spectral_pb, calc_gpb, plus gpb() for the whole pipeline.
"""
import ctypes
import os

import numpy as np

from ._lib import BatchImage, DmscError, GpbOutputs, GpbParams, StepShape, check, lib

__all__ = [
    "DmscError", "Context", "generate_image", "boundary_mask", "load_ppm", "pitch_in_floats", "gpb",
    "PREC_FP32", "PREC_MIXED",
]

PREC_FP32 = 0
PREC_MIXED = 1
STOP_RESIDUAL, STOP_STEADY, STOP_FLOOR, STOP_CAP = 1, 2, 3, 4
TEXTON32 = 1
TEXTON64 = 2


def _p(t):
    """Raw pointer of a tensor / ndarray (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(t.data_ptr())


def tune_get(key):
    """Current value of a process-wide tuning key (dmsc_tune_get)."""
    v = ctypes.c_int()
    check(lib().dmsc_tune_get(key.encode(), ctypes.byref(v)))
    return v.value


class tuned:
    """Context manager: set tuning keys (dmsc_tune_set) for a block of code
    and restore their previous values afterwards, e.g.
    ``with dm.tuned(eigensolver=0): ctx.gpb(...)``."""

    def __init__(self, **keys):
        self.keys = keys
        self.old = {}

    def __enter__(self):
        for k, v in self.keys.items():
            self.old[k] = tune_get(k)
            check(lib().dmsc_tune_set(k.encode(), int(v)))
        return self

    def __exit__(self, *exc):
        for k, v in self.old.items():
            check(lib().dmsc_tune_set(k.encode(), v))
        return False


def pitch_in_floats(n):
    return lib().dmsc_pitch_in_floats(int(n))


def generate_image(w, h, ncells, seed, with_cells=False):
    """Synthetic RGB8 test image packed one uint32 per pixel (R in bits 0-7)."""
    rgb = np.zeros((h, w), np.uint32)
    cells = np.zeros((h, w), np.int32) if with_cells else None
    check(lib().dmsc_generate_image(w, h, ncells, ctypes.c_uint64(seed), _p(rgb), _p(cells)))
    return (rgb, cells) if with_cells else rgb


def load_ppm(path):
    """convert.cu:88-95 loadPPM_rgbU: a binary PPM (P6, maxval 255) as a
    [h, w] uint32 array, R in bits 0-7, G 8-15, B 16-23."""
    w, h = ctypes.c_int(), ctypes.c_int()
    check(lib().dmsc_load_ppm(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), None))
    rgb = np.zeros((h.value, w.value), np.uint32)
    check(lib().dmsc_load_ppm(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), _p(rgb)))
    return rgb


def boundary_mask(cells):
    h, w = cells.shape
    m = np.zeros((h, w), np.uint8)
    check(lib().dmsc_boundary_mask(w, h, _p(np.ascontiguousarray(cells, np.int32)), _p(m)))
    return m


class Context:
    """A device + stream + scratch arenas (dmsc_ctx)."""

    def __init__(self, device=0, stream=None):
        import torch

        self._torch = torch
        self.device = torch.device("cuda", device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        h = ctypes.c_void_p()
        check(lib().dmsc_ctx_create(ctypes.byref(h), device, ctypes.c_void_p(stream.cuda_stream)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().dmsc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().dmsc_ctx_synchronize(self.handle))

    # -------------------------------------------------------------- buffers
    def empty(self, *shape, dtype=None):
        return self._torch.empty(*shape, dtype=dtype or self._torch.float32, device=self.device)

    def zeros(self, *shape, dtype=None):
        return self._torch.zeros(*shape, dtype=dtype or self._torch.float32, device=self.device)

    def upload(self, arr):
        t = self._torch.from_numpy(np.ascontiguousarray(arr))
        if t.dtype == self._torch.uint32:
            t = t.view(self._torch.int32)
        return t.to(self.device)

    # -------------------------------------------------------------- stages
    def rgb_to_grey(self, rgb):
        h, w = rgb.shape
        out = self.empty(h, w)
        check(lib().dmsc_rgb_to_grey(self.handle, w, h, _p(rgb), _p(out)))
        return out

    def rgb_to_lab(self, rgb, gamma=2.5, normalize=True):
        h, w = rgb.shape
        L, a, b = self.empty(h, w), self.empty(h, w), self.empty(h, w)
        check(lib().dmsc_rgb_to_lab(self.handle, w, h, gamma, _p(rgb), _p(L), _p(a), _p(b)))
        if normalize:
            check(lib().dmsc_normalize_lab(self.handle, w, h, _p(L), _p(a), _p(b)))
        return L, a, b

    def mirror_image(self, img, border):
        """convert.h:75/97 mirrorImage: symmetric-reflect padding of a float32
        or int32 [h, w] image by `border` pixels on every side."""
        h, w = img.shape
        out = self.empty(h + 2 * border, w + 2 * border, dtype=img.dtype)
        fn = lib().dmsc_mirror_image_i if img.dtype == self._torch.int32 else lib().dmsc_mirror_image_f
        check(fn(self.handle, w, h, border, _p(img), _p(out)))
        return out

    def unmirror_image(self, img, border):
        """convert.h:120 unMirrorImage: the interior of a mirrored float32
        image."""
        bh, bw = img.shape
        out = self.empty(bh - 2 * border, bw - 2 * border)
        check(lib().dmsc_unmirror_image_f(self.handle, bw, bh, border, _p(img), _p(out)))
        return out

    def quantize_image(self, img, nbins):
        """convert.h:140 quantizeImage: floor(v * nbins), nbins -> nbins - 1."""
        h, w = img.shape
        out = self.empty(h, w, dtype=self._torch.int32)
        check(lib().dmsc_quantize_image(self.handle, w, h, nbins, _p(img), _p(out)))
        return out

    def texton_responses(self, grey):
        h, w = grey.shape
        out = self.empty(34, h, w)
        check(lib().dmsc_texton_responses(self.handle, w, h, _p(grey), _p(out)))
        return out

    def kmeans(self, resp, precision=PREC_FP32, max_iter=15, k=32):
        _, h, w = resp.shape
        labels = self.empty(h, w, dtype=self._torch.int32)
        ch, it = ctypes.c_int(), ctypes.c_int()
        check(lib().dmsc_kmeans(self.handle, w, h, k, 34, _p(resp), _p(labels), max_iter, 0, precision,
                                ctypes.byref(ch), ctypes.byref(it)))
        return labels, it.value

    def find_textons(self, grey, precision=PREC_FP32, texton_choice=TEXTON32):
        h, w = grey.shape
        labels = self.empty(h, w, dtype=self._torch.int32)
        it = ctypes.c_int()
        check(lib().dmsc_find_textons(self.handle, w, h, _p(grey), _p(labels), texton_choice, precision,
                                      ctypes.byref(it)))
        return labels, it.value

    def local_cues(self, L, a, b, textons, raw=False, texton_choice=TEXTON32):
        """Returns (bg, cga, cgb, tg) as [24, N] views of pitched stacks (and raw [4,3,8,N])."""
        h, w = L.shape
        n = w * h
        pitch = pitch_in_floats(n)
        cues = [self.zeros(24, pitch) for _ in range(4)]
        if raw:
            r = self.empty(4, 3, 8, n)
            check(lib().dmsc_local_cues_raw(self.handle, w, h, _p(L), _p(a), _p(b), _p(textons),
                                            *[_p(c) for c in cues], pitch, _p(r), texton_choice))
            return [c[:, :n] for c in cues], r
        check(lib().dmsc_local_cues(self.handle, w, h, _p(L), _p(a), _p(b), _p(textons),
                                    *[_p(c) for c in cues], pitch, texton_choice))
        return [c[:, :n] for c in cues]

    def combine(self, w, h, bg, cga, cgb, tg, texton_choice=TEXTON32):
        """Cue stacks given as pitched [24, pitch] tensors; returns pitched (mpb, combinedG)."""
        pitch = bg.shape[1]
        mpb, cg = self.zeros(8, pitch), self.zeros(8, pitch)
        check(lib().dmsc_combine(self.handle, w, h, pitch, _p(bg), _p(cga), _p(cgb), _p(tg), _p(mpb), _p(cg),
                                 texton_choice))
        return mpb, cg

    def non_max_suppression(self, w, h, pb):
        out = self.empty(h, w)
        check(lib().dmsc_nonmax(self.handle, w, h, _p(pb), pb.shape[1], _p(out)))
        return out

    def intervene(self, mpb, sigma=0.1):
        h, w = mpb.shape
        pitch = pitch_in_floats(w * h)
        m = self.empty(81, pitch)
        check(lib().dmsc_intervene(self.handle, w, h, _p(mpb), _p(m), pitch, sigma))
        return m

    def convert_matrix(self, w, h, m):
        d, rs = self.empty(h, w), self.empty(h, w)
        check(lib().dmsc_convert_matrix(self.handle, w, h, _p(m), m.shape[1], _p(d), _p(rs)))
        return d, rs

    def autotune_eigensolver(self, w, h, m, precision=PREC_FP32, steps=200):
        """dmsc_autotune_eigensolver on the DIA matrix m (converted in place):
        times every step-kernel shape and keeps the fastest for this size and
        precision.  Returns (best, trials) as dicts."""
        best = StepShape()
        tried = (StepShape * 32)()
        nt = ctypes.c_int()
        check(lib().dmsc_autotune_eigensolver(self.handle, w, h, _p(m), m.shape[1], precision, steps,
                                              ctypes.byref(best), tried, 32, ctypes.byref(nt)))

        def d(x):
            return {"mixed_fused": x.mixed_fused, "tile": x.tile, "k1_threads": x.k1_threads,
                    "ms_per_step": x.ms_per_step}
        return d(best), [d(tried[i]) for i in range(min(nt.value, 32))]

    def autotune_spmv(self, w, h, m, reps=100):
        """dmsc_autotune_spmv: returns (rows per lane selected, {shape: us})."""
        x = self._torch.rand(h, w, device=self.device)
        y = self.empty(h, w)
        best = ctypes.c_int()
        us = (ctypes.c_float * 4)()
        check(lib().dmsc_autotune_spmv(self.handle, w, h, _p(m), m.shape[1], _p(x), _p(y), reps,
                                       ctypes.byref(best), us))
        return best.value, {s: us[i] for i, s in enumerate((0, 1, 2, 4))}

    def chebyshev_tile(self, w, h):
        """dmsc_chebyshev_tile: the block solver's filter tile at w x h."""
        v = ctypes.c_int()
        check(lib().dmsc_chebyshev_tile(self.handle, w, h, ctypes.byref(v)))
        return v.value

    def autotune_chebyshev(self, w, h, m, products=100):
        """dmsc_autotune_chebyshev on a converted matrix m: returns (filter tile
        selected for w x h, {tile: us per step})."""
        best, nt = ctypes.c_int(), ctypes.c_int()
        shapes = (ctypes.c_int * 32)()
        us = (ctypes.c_float * 32)()
        check(lib().dmsc_autotune_chebyshev(self.handle, w, h, _p(m), m.shape[1], products, ctypes.byref(best),
                                            shapes, us, 32, ctypes.byref(nt)))
        return best.value, {shapes[i]: us[i] for i in range(min(nt.value, 32))}

    def spmv(self, w, h, m, x):
        y = self.empty(h, w)
        check(lib().dmsc_stencil_spmv(self.handle, w, h, _p(m), m.shape[1], _p(x), _p(y)))
        return y

    def generalized_eigensolve(self, w, h, m, nvec=9, tol=1e-3, precision=PREC_FP32):
        vals = np.zeros(nvec, np.float32)
        vecs = self.empty(nvec, h, w)
        it = ctypes.c_int()
        check(lib().dmsc_generalized_eigensolve(self.handle, w, h, _p(m), m.shape[1], nvec, tol, precision,
                                                _p(vals), _p(vecs), ctypes.byref(it)))
        return vals, vecs, it.value

    def eigensolve_report(self):
        """How the last eigensolve on this context stopped: (stop rule, the
        largest residual of its wanted pairs); rules as DMSC_STOP_* in
        dmsc.h (1 residual bound met, 2 vectors steady, 3 FP32 floor, 4
        iteration cap)."""
        rule, res = ctypes.c_int(), ctypes.c_double()
        check(lib().dmsc_eigensolve_report(self.handle, ctypes.byref(rule), ctypes.byref(res)))
        return rule.value, res.value

    def spectral_pb(self, eigvals, eigvecs):
        nvec, h, w = eigvecs.shape
        pitch = pitch_in_floats(w * h)
        out = self.zeros(8, pitch)
        ev = np.ascontiguousarray(eigvals, np.float32)
        check(lib().dmsc_spectral_pb(self.handle, _p(ev), _p(eigvecs), w, h, nvec, _p(out), pitch))
        return out

    def calc_gpb(self, w, h, cg, spb, nmax):
        pitch = cg.shape[1]
        all_ = self.zeros(8, pitch)
        g = self.empty(h, w)
        check(lib().dmsc_calc_gpb(self.handle, w * h, pitch, 8, _p(cg), _p(spb), _p(nmax), _p(all_), _p(g)))
        return all_, g

    def postprocess(self, gpb, nmax):
        """skeleton.cu PostProcess: returns (sigmoid(gPb), thinned sigmoid(gPb),
        thinning iterations); the input gPb is left unchanged."""
        h, w = nmax.shape
        g = gpb.reshape(h, w).clone()
        thin = self.empty(h, w)
        it = ctypes.c_int()
        check(lib().dmsc_postprocess(self.handle, w, h, _p(g), _p(nmax), _p(thin), ctypes.byref(it)))
        return g, thin, it.value

    def normalize_gpb_all(self, gpb_all, npix):
        """skeleton.cu NormalizeGpbAll on a copy of the [8][pitch] planes."""
        out = gpb_all.clone()
        check(lib().dmsc_normalize_gpb_all(self.handle, npix, out.shape[0], out.shape[1], _p(out)))
        return out

    def _gpb_buffers(self, w, h, n_eig):
        n = w * h
        pitch = pitch_in_floats(n)
        torch = self._torch
        out = {
            "textons": self.empty(h, w, dtype=torch.int32),
            "mpb": self.zeros(8, pitch), "combined_g": self.zeros(8, pitch),
            "nmax": self.empty(h, w), "eigvecs": self.empty(n_eig, h, w),
            "spb": self.zeros(8, pitch), "gpb_all": self.zeros(8, pitch), "gpb": self.empty(h, w),
            "pitch": pitch,
        }
        o = GpbOutputs()
        o.d_textons, o.d_mpb, o.d_combined_g = out["textons"].data_ptr(), out["mpb"].data_ptr(), out["combined_g"].data_ptr()
        o.d_nmax, o.d_eigvecs = out["nmax"].data_ptr(), out["eigvecs"].data_ptr()
        o.d_spb, o.d_gpb_all, o.d_gpb = out["spb"].data_ptr(), out["gpb_all"].data_ptr(), out["gpb"].data_ptr()
        o.pitch = pitch
        return out, o

    @staticmethod
    def _gpb_results(out, o, n_eig):
        out["eigvals"] = np.array(o.h_eigvals[:n_eig], np.float32)
        out["kmeans_iterations"] = o.kmeans_iterations
        out["lanczos_iterations"] = o.lanczos_iterations
        out["stage_ms"] = list(o.stage_ms[:8])
        return out

    def gpb(self, rgb, n_eig=9, tol=1e-3, precision=PREC_FP32, timing=False, texton_choice=TEXTON32):
        """Whole pipeline on a device RGB8 image (int32/uint32 [h, w] tensor)."""
        h, w = rgb.shape
        out, o = self._gpb_buffers(w, h, n_eig)
        p = GpbParams(n_eig, tol, texton_choice, precision)
        check(lib().dmsc_gpb(self.handle, w, h, _p(rgb), ctypes.byref(p), ctypes.byref(o), 1 if timing else 0))
        return self._gpb_results(out, o, n_eig)

    def gpb_batch(self, rgbs, n_eig=9, tol=1e-3, precision=PREC_FP32, texton_choice=TEXTON32, lanes=0,
                  buffers=None):
        """Whole pipeline on a list of device RGB8 images of any sizes
        (dmsc_gpb_batch: concurrent lanes, largest images first).  Returns one
        output dict per image, as gpb() does.  `buffers` (what
        gpb_batch_buffers returned for images of the same sizes) reuses
        output tensors across calls."""
        if buffers is None:
            buffers = self.gpb_batch_buffers(rgbs, n_eig)
        outs, objs = buffers
        items = (BatchImage * len(rgbs))()
        for i, rgb in enumerate(rgbs):
            h, w = rgb.shape
            items[i].w, items[i].h, items[i].d_rgb = w, h, rgb.data_ptr()
            items[i].out = ctypes.pointer(objs[i])
        p = GpbParams(n_eig, tol, texton_choice, precision)
        check(lib().dmsc_gpb_batch(self.handle, len(rgbs), items, ctypes.byref(p), lanes))
        return [self._gpb_results(outs[i], objs[i], n_eig) for i in range(len(rgbs))]

    def gpb_batch_buffers(self, rgbs, n_eig=9):
        pairs = [self._gpb_buffers(r.shape[1], r.shape[0], n_eig) for r in rgbs]
        return [q[0] for q in pairs], [q[1] for q in pairs]
