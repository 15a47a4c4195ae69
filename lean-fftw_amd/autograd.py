"""Differentiable FFT for PyTorch-ROCm, with a hand-written adjoint.

FFTFunction.apply(x, weight=None, sign=-1) returns y = diag(weight) F x along the last
dimension (F the unnormalised DFT of direction `sign`, weight a real float64 vector of
length n fused into the kernel's store). Its backward is the adjoint
F^H diag(weight) g: the opposite-direction Stockham schedule with the weight fused into
the load of the first pass -- the conjugate-twiddle transform, since the backward tables
are the forward ones without the negated imaginary part (LeanFFTW/Core/FastPlan.lean:27-37).
For real weights and PyTorch's conjugate-Wirtinger convention, the gradient of a real loss
L(y) with respect to x is exactly that: grad_x = F^H (weight * grad_y).

WeightedPowerLoss fuses the whole loss sum |w * F x|^2 (config of the differentiable
benchmark) into one forward and one adjoint kernel. The weight is treated as a constant
(no gradient flows to it).
"""

from __future__ import annotations

import torch


def _lf():
    import sys

    return sys.modules[__name__.rsplit(".", 1)[0]]


_plans: dict = {}


def _plan(n: int, sign: int, howmany: int):
    # one plan per stream (a plan's intermediate must not be shared by executions that
    # overlap on two streams); under a graph capture, the plan made before it
    return _lf()._cached_plan(n, sign, howmany, _plans)


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


class WeightedPowerLoss(torch.autograd.Function):
    """loss = sum |w * F x|^2 over a batch of rows, fused.

    Forward: one kernel writes y = w * F x and, in the same stores, per-workgroup partial
    sums of |y|^2 (a fixed-order reduction, so the loss is deterministic); the loss is the
    sum of those partials. Backward: grad_x = F^H (2 g w * y), g = d(out)/d(loss): one
    adjoint kernel reading the saved y with the real weight (2 g) w formed in its load from
    w and the device scalar g (no elementwise launch for the weight).
    """

    @staticmethod
    def forward(ctx, x, w):
        if x.dtype != torch.complex128 or not x.is_cuda:
            raise TypeError("WeightedPowerLoss takes complex128 CUDA tensors")
        x = x.contiguous()
        n = x.shape[-1]
        howmany = x.numel() // n
        w = w.to(device=x.device, dtype=torch.float64).contiguous()
        plan = _plan(n, -1, howmany)
        y = torch.empty_like(x)
        parts = torch.empty(plan.partials, dtype=torch.float64, device=x.device)
        plan.call_weighted_sumsq(x, y, parts, post=w)
        ctx.save_for_backward(y, w)
        ctx.meta = (n, howmany)
        return parts.sum()

    @staticmethod
    def backward(ctx, g):
        y, w = ctx.saved_tensors
        n, howmany = ctx.meta
        g = g.to(device=y.device, dtype=torch.float64).reshape(1).contiguous()
        out = torch.empty_like(y)
        _plan(n, 1, howmany).call_weighted_scaled(y, out, pre=w, scale=g, factor=2.0)
        return out, None


def weighted_power_loss(x, w):
    """loss = sum(|w * FFT(x)|^2) along the last dimension, differentiable in x."""
    return WeightedPowerLoss.apply(x, w)
