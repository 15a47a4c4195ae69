"""Differentiable FFT for PyTorch-ROCm, with a hand-written adjoint.

FFTFunction.apply(x, weight=None, sign=-1) returns y = diag(weight) F x along the last
dimension (F the unnormalised DFT of direction `sign`, weight a real float64 vector of
length n fused into the kernel's store). Its backward is the adjoint
F^H diag(weight) g: the opposite-direction Stockham schedule with the weight fused into
the load of the first pass -- the conjugate-twiddle transform, since the backward tables
are the forward ones without the negated imaginary part (LeanFFTW/Core/FastPlan.lean:27-37).
For real weights and PyTorch's conjugate-Wirtinger convention, the gradient of a real loss
L(y) with respect to x is exactly that: grad_x = F^H (weight * grad_y).

The weight is treated as a constant (no gradient flows to it).
"""

from __future__ import annotations

import torch


def _lf():
    import sys

    return sys.modules[__name__.rsplit(".", 1)[0]]


_plans: dict = {}


def _plan(n: int, sign: int, howmany: int):
    key = (n, sign, howmany, torch.cuda.current_device())
    p = _plans.get(key)
    if p is None:
        p = _lf().Plan(n, sign, howmany)
        _plans[key] = p
    return p


class FFTFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, weight=None, sign=-1):
        if x.dtype != torch.complex128 or not x.is_cuda:
            raise TypeError("FFTFunction takes complex128 CUDA tensors")
        x = x.contiguous()
        n = x.shape[-1]
        howmany = x.numel() // n
        out = torch.empty_like(x)
        plan = _plan(n, sign, howmany)
        if weight is None:
            plan.call(x, out)
        else:
            weight = weight.to(device=x.device, dtype=torch.float64).contiguous()
            plan.call_weighted(x, out, post=weight)
        ctx.save_for_backward(weight if weight is not None else torch.empty(0, device=x.device))
        ctx.meta = (n, sign, howmany, weight is not None)
        return out

    @staticmethod
    def backward(ctx, g):
        (w,) = ctx.saved_tensors
        n, sign, howmany, weighted = ctx.meta
        g = g.contiguous()
        out = torch.empty_like(g)
        plan = _plan(n, -sign, howmany)
        if weighted:
            plan.call_weighted(g, out, pre=w)
        else:
            plan.call(g, out)
        return out, None, None


def fft_autograd(x, weight=None, sign: int = -1):
    return FFTFunction.apply(x, weight, sign)


def weighted_power_loss(x, w):
    """loss = sum(|w * FFT(x)|^2) along the last dimension, differentiable in x."""
    y = FFTFunction.apply(x, w, -1)
    return (y.real * y.real + y.imag * y.imag).sum()
