"""Differentiable spectral Pb + combine behind torch.autograd.

The forward and backward are the library's HIP kernels
(dmsc_spb_combine_forward / dmsc_spb_combine_backward).  The parameters
are one float32 device vector

    params = [w_0 .. w_11, gamma, eta_0 .. eta_{nvec-1}]

w_k the combinedG weight of cue x scale plane k (combine.cu:21 order: Bg
scales 0-2, Cga, Cgb, Tg), gamma the sPb weight (0.0139, globalPb.cu:13) and
eta_i the weight of eigenvector i (1/sqrt(lambda_i), spectralPb.cu:759;
eta_0 is unused).  The output is gPbPre = max over orientations of
gPbAll_o = sum_k w_k C_{k,o} + gamma sum_i eta_i |K_o * v_i| -- the gPb before
the non-max mask of globalPb.cu:43-47.

The loss of the boundary fit is a class-balanced binary cross-entropy on the
logit a0 + a1 gPbPre, with the reference's PostProcess sigmoid constants
(skeleton.cu:79-95: a0 = -2.6433, a1 = 10.7998): positives weighted by the
fraction of negatives and negatives by the fraction of positives.
"""
import ctypes

import numpy as np

from ._lib import check, lib

A0, A1 = -2.6433, 10.7998


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def reference_params(eigvals, texton_choice=1, device=None):
    """The reference's weights: combine table, 0.0139, 1/sqrt(lambda_i)
    (computed in float32 as the pipeline does)."""
    import torch

    coef = (ctypes.c_float * 12)()
    wts = (ctypes.c_float * 12)()
    check(lib().dmsc_combine_weights(texton_choice, coef, wts))
    ev = np.asarray(eigvals, np.float32)
    eta = np.zeros(len(ev), np.float32)
    eta[1:] = np.float32(1.0) / np.sqrt(ev[1:]).astype(np.float32)
    v = np.concatenate([np.array(wts[:], np.float32), np.array([0.0139], np.float32), eta])
    return torch.from_numpy(v).to(device) if device is not None else torch.from_numpy(v)


class SpbCombine:
    """Cue stacks (Bg, Cga, Cgb, Tg as [24, cue_pitch] device tensors) of one
    image, fixed across calls; __call__(params, eigvecs) -> gPbPre [h, w]."""

    def __init__(self, ctx, w, h, cues):
        self.ctx, self.w, self.h = ctx, w, h
        self.cues = [c.contiguous() for c in cues]
        self.cue_pitch = self.cues[0].shape[1]

    def forward(self, params, eigvecs, gpb_all=None):
        import torch

        nvec = eigvecs.shape[0]
        assert params.dtype == torch.float32 and params.numel() == 13 + nvec and params.is_contiguous()
        assert eigvecs.dtype == torch.float32 and eigvecs.is_contiguous()
        pre = torch.empty(self.h, self.w, device=params.device)
        amax = torch.empty(self.h, self.w, dtype=torch.int32, device=params.device)
        bg, cga, cgb, tg = self.cues
        pitch = gpb_all.shape[1] if gpb_all is not None else 0
        check(lib().dmsc_spb_combine_forward(self.ctx.handle, self.w, self.h, self.cue_pitch, _p(bg), _p(cga), _p(cgb),
                                             _p(tg), nvec, _p(eigvecs), _p(params),
                                             _p(gpb_all) if gpb_all is not None else None, pitch, _p(pre), _p(amax)))
        return pre, amax

    def backward(self, params, eigvecs, amax, grad):
        import torch

        nvec = eigvecs.shape[0]
        gp = torch.empty(13 + nvec, device=params.device)
        gv = torch.empty_like(eigvecs)
        bg, cga, cgb, tg = self.cues
        check(lib().dmsc_spb_combine_backward(self.ctx.handle, self.w, self.h, self.cue_pitch, _p(bg), _p(cga),
                                              _p(cgb), _p(tg), nvec, _p(eigvecs), _p(params), _p(amax),
                                              _p(grad.contiguous()), _p(gp), _p(gv)))
        return gp, gv

    def __call__(self, params, eigvecs):
        return _SpbCombineFn.apply(params, eigvecs, self)


def _make_fn():
    import torch

    class Fn(torch.autograd.Function):
        @staticmethod
        def forward(fctx, params, eigvecs, op):
            pre, amax = op.forward(params.contiguous(), eigvecs.contiguous())
            fctx.op = op
            fctx.save_for_backward(params, eigvecs, amax)
            return pre

        @staticmethod
        def backward(fctx, grad):
            params, eigvecs, amax = fctx.saved_tensors
            gp, gv = fctx.op.backward(params.contiguous(), eigvecs.contiguous(), amax, grad)
            return gp, gv, None

    return Fn


class _Lazy:
    fn = None

    def apply(self, *a):
        if _Lazy.fn is None:
            _Lazy.fn = _make_fn()
        return _Lazy.fn.apply(*a)


_SpbCombineFn = _Lazy()


def balanced_bce(gpb_pre, mask):
    """Class-balanced BCE of sigmoid(A0 + A1 gPbPre) against a boolean
    boundary mask (mean over pixels)."""
    import torch

    y = mask.to(gpb_pre.dtype)
    pos = y.mean()
    wgt = (1 - pos) * y + pos * (1 - y)
    z = A0 + A1 * gpb_pre
    return torch.nn.functional.binary_cross_entropy_with_logits(z, y, weight=wgt, reduction="mean")
