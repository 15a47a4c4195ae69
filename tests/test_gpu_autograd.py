"""The differentiable FFT: forward and adjoint kernels against the oracle, and the gradient
of loss = sum(|w * FFT(x)|^2) against central differences (h = 1e-6, 64 random
coordinates) and against the oracle's backward refTransform."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

DEV = "cuda:0"


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def test_forward_and_adjoint_bit_exact(lf, oracle):
    N, hm = 4096, 8
    x = oracle.lcg_input(N, seed=31, howmany=hm)
    w = np.random.default_rng(7).uniform(0, 1, N)
    xd = torch.from_numpy(x).to(DEV).requires_grad_(True)
    wd = torch.from_numpy(w).to(DEV)
    y = lf.FFTFunction.apply(xd, wd, -1)
    rs = lf.Plan(N, -1, hm).radices
    assert bits_equal(y.detach().cpu().numpy(), oracle.ref_transform(True, rs, x) * w)
    g = oracle.lcg_input(N, seed=32, howmany=hm)
    (gx,) = torch.autograd.grad(y, xd, torch.from_numpy(g).to(DEV))
    rsb = lf.Plan(N, 1, hm).radices
    ref = oracle.ref_transform(False, rsb, g * w)  # F^H (w g): backward transform, weight on load
    assert bits_equal(gx.cpu().numpy(), ref)


def test_unweighted_adjoint_is_backward_transform(lf, oracle):
    N, hm = 1000, 4
    x = oracle.lcg_input(N, seed=41, howmany=hm)
    xd = torch.from_numpy(x).to(DEV).requires_grad_(True)
    y = lf.fft_autograd(xd)
    g = oracle.lcg_input(N, seed=42, howmany=hm)
    (gx,) = torch.autograd.grad(y, xd, torch.from_numpy(g).to(DEV))
    assert bits_equal(gx.cpu().numpy(), oracle.ref_transform(False, lf.Plan(N, 1, hm).radices, g))


def test_gradient_central_differences():
    """Config of the differentiable benchmark: x (512, 4096) complex128, w ~ U[0,1)."""
    from conftest import load_product

    lf = load_product()
    g = torch.Generator(device=DEV).manual_seed(2026)
    B, N = 512, 4096
    x = torch.complex(torch.rand(B, N, dtype=torch.float64, device=DEV, generator=g) - 0.5,
                      torch.rand(B, N, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    w = torch.rand(N, dtype=torch.float64, device=DEV, generator=g)
    xr = x.clone().requires_grad_(True)
    loss = lf.weighted_power_loss(xr, w)
    loss.backward()
    grad = xr.grad
    rng = np.random.default_rng(11)
    h = 1e-6
    scale = grad.abs().max().item()
    for _ in range(64):
        b, n = int(rng.integers(B)), int(rng.integers(N))
        part = int(rng.integers(2))
        d = torch.zeros_like(x)
        d[b, n] = h if part == 0 else 1j * h
        with torch.no_grad():
            lp = lf.weighted_power_loss(x + d, w).item()
            lm = lf.weighted_power_loss(x - d, w).item()
        fd = (lp - lm) / (2 * h)
        an = grad[b, n].real.item() if part == 0 else grad[b, n].imag.item()
        assert abs(fd - an) <= 2e-5 * scale + 1e-6 * abs(an), (b, n, part, fd, an)


def test_gradient_matches_closed_form(lf, oracle):
    """grad_x = F^H (2 w^2 F x): the backward transform of the weighted forward output."""
    B, N = 4, 4096
    x = oracle.lcg_input(N, seed=51, howmany=B)
    w = np.random.default_rng(9).uniform(0, 1, N)
    xd = torch.from_numpy(x).to(DEV).requires_grad_(True)
    loss = lf.weighted_power_loss(xd, torch.from_numpy(w).to(DEV))
    loss.backward()
    rs = lf.Plan(N, -1, B).radices
    rsb = lf.Plan(N, 1, B).radices
    y = oracle.ref_transform(True, rs, x) * w
    ref = oracle.ref_transform(False, rsb, (2.0 * y) * w)
    got = xd.grad.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
