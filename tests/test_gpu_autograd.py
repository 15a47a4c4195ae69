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
    """grad_x = F^H (2 w^2 F x): the backward transform of the weighted forward output, bit
    for bit (the adjoint's weight (2 g) w with g = 1 is 2 w, and y (2 w) = (2 y) w exactly)."""
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
    assert bits_equal(got, ref)


@pytest.mark.parametrize("N,B", [(16, 300), (1000, 3), (4096, 8), (16384, 2), (65536, 2)])
def test_scaled_input_weights_bit_exact(lf, oracle, N, B):
    """Input weights times a device scalar (mfft_execute_weighted_scaled): equal, bit for bit,
    to the unscaled entry point with the weight vector (2 s) w formed on the device, and to
    the oracle (the generic stage kernel with staged copies and without, the fixed register
    kernel, the split kernel, a two-stage plan)."""
    x = oracle.lcg_input(N, seed=70 + N, howmany=B)
    w = np.random.default_rng(N + 1).uniform(0, 1, N)
    s = 0.37
    p = lf.Plan(N, 1, B)
    xd = torch.from_numpy(x).to(DEV)
    wd = torch.from_numpy(w).to(DEV)
    sd = torch.tensor([s], dtype=torch.float64, device=DEV)
    y1 = torch.empty_like(xd)
    p.call_weighted_scaled(xd, y1, pre=wd, scale=sd, factor=2.0)
    y2 = torch.empty_like(xd)
    p.call_weighted(xd, y2, pre=(2.0 * sd) * wd)
    torch.cuda.synchronize()
    assert bits_equal(y1.cpu().numpy(), y2.cpu().numpy())
    ref = oracle.ref_transform(False, p.radices, x * ((2.0 * s) * w))
    assert bits_equal(y1.cpu().numpy(), ref)


def test_loss_gradient_with_upstream_scale(lf, oracle):
    """d(c loss)/dx for an upstream gradient c: the adjoint's weight (2 c) w is formed in the
    kernel from the device scalar c, bit-exact against the oracle."""
    B, N = 4, 4096
    x = oracle.lcg_input(N, seed=53, howmany=B)
    w = np.random.default_rng(10).uniform(0, 1, N)
    xd = torch.from_numpy(x).to(DEV).requires_grad_(True)
    loss = lf.weighted_power_loss(xd, torch.from_numpy(w).to(DEV))
    c = 0.731
    (gx,) = torch.autograd.grad(loss, xd, torch.tensor(c, dtype=torch.float64, device=DEV))
    y = oracle.ref_transform(True, lf.Plan(N, -1, B).radices, x) * w
    ref = oracle.ref_transform(False, lf.Plan(N, 1, B).radices, y * ((2.0 * c) * w))
    assert bits_equal(gx.cpu().numpy(), ref)


def test_fused_loss_matches_unfused(lf, oracle):
    """The fused loss equals sum |y|^2 of the weighted forward transform, and its gradient
    the adjoint of the unfused composition, to rounding."""
    B, N = 16, 4096
    x = oracle.lcg_input(N, seed=61, howmany=B)
    w = np.random.default_rng(12).uniform(0, 1, N)
    xd = torch.from_numpy(x).to(DEV).requires_grad_(True)
    wd = torch.from_numpy(w).to(DEV)
    loss = lf.weighted_power_loss(xd, wd)
    loss.backward()
    y = oracle.ref_transform(True, lf.Plan(N, -1, B).radices, x) * w
    ref_loss = float(np.sum(y.real ** 2 + y.imag ** 2))
    assert abs(loss.item() - ref_loss) <= 1e-12 * ref_loss
    x2 = torch.from_numpy(x).to(DEV).requires_grad_(True)
    yy = lf.FFTFunction.apply(x2, wd, -1)
    (yy.real * yy.real + yy.imag * yy.imag).sum().backward()
    assert torch.allclose(xd.grad, x2.grad, rtol=0, atol=1e-12 * x2.grad.abs().max().item())


def test_partials_cover_two_stage_plans(lf, oracle):
    """Sum-of-squares partials of a chunked two-stage plan (N = 65536)."""
    B, N = 6, 65536
    x = oracle.lcg_input(N, seed=62, howmany=B)
    p = lf.Plan(N, -1, B)
    xd = torch.from_numpy(x).to(DEV)
    y = torch.empty_like(xd)
    parts = torch.full((p.partials,), float("nan"), dtype=torch.float64, device=DEV)
    p.call_weighted_sumsq(xd, y, parts)
    torch.cuda.synchronize()
    assert not torch.isnan(parts).any()
    ref = (y.abs() ** 2).sum().item()
    assert abs(parts.sum().item() - ref) <= 1e-12 * ref


@pytest.mark.parametrize("N,B", [(32, 70), (512, 7), (1024, 3), (8192, 2), (16384, 3), (2100, 5)])
def test_partials_every_kernel_shape(lf, oracle, N, B):
    """Weighted transform with sum-of-squares partials on the register-resident kernel (several
    transforms per workgroup with a partial last workgroup, one per workgroup, 16 elements per
    thread) and on a large LDS tile: outputs bit-exact, partials sum to sum |out|^2."""
    x = oracle.lcg_input(N, seed=63 + N, howmany=B)
    w = np.random.default_rng(N).uniform(0, 1, N)
    p = lf.Plan(N, -1, B)
    xd = torch.from_numpy(x).to(DEV)
    wd = torch.from_numpy(w).to(DEV)
    y = torch.empty_like(xd)
    parts = torch.full((p.partials,), float("nan"), dtype=torch.float64, device=DEV)
    p.call_weighted_sumsq(xd, y, parts, pre=wd, post=wd)
    torch.cuda.synchronize()
    ref = oracle.ref_transform(True, p.radices, x * w) * w
    got = y.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert not torch.isnan(parts).any()
    tot = float(np.sum(ref.real ** 2 + ref.imag ** 2))
    assert abs(parts.sum().item() - tot) <= 1e-12 * tot


def test_graph_captured_fwd_bwd_replays(lf):
    """The fused loss and its backward captured in a HIP graph: every replay, with eager work
    in between, matches the eager result bit for bit.

    Each step keeps only the detached loss. Keeping the loss with its graph would keep x's
    gradient accumulator alive from the first eager step, and with it that step's stream (the
    legacy default stream): the captured backward would then synchronise the capture stream
    with the legacy stream (the engine's producer/consumer and leaf-stream events), which a
    global-mode capture forbids."""
    B, N = 64, 4096
    g = torch.Generator(device=DEV).manual_seed(77)
    x = torch.complex(torch.rand(B, N, dtype=torch.float64, device=DEV, generator=g) - 0.5,
                      torch.rand(B, N, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    x.requires_grad_(True)
    w = torch.rand(N, dtype=torch.float64, device=DEV, generator=g)
    out = {}

    def step():
        x.grad = None
        loss = lf.weighted_power_loss(x, w)
        loss.backward()
        out["loss"] = loss.detach()

    step()
    ref_loss = out["loss"].detach().clone()
    ref_grad = x.grad.detach().clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(3):
            step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for i in range(4):
        junk = torch.randn(1 << 20, device=DEV) * 2.0  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out["loss"], ref_loss), i
        assert torch.equal(x.grad, ref_grad), i
        del junk


def test_graph_captured_forward_and_adjoint(lf):
    """The fused forward loss and the explicit adjoint launch (what the Function's backward
    runs) captured in HIP graphs: every replay, with eager work in between, matches the
    eager result bit for bit."""
    B, N = 64, 4096
    g = torch.Generator(device=DEV).manual_seed(78)
    x = torch.complex(torch.rand(B, N, dtype=torch.float64, device=DEV, generator=g) - 0.5,
                      torch.rand(B, N, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    w = torch.rand(N, dtype=torch.float64, device=DEV, generator=g)
    fwd = lf.Plan(N, -1, B)
    bwd = lf.Plan(N, 1, B)
    y = torch.empty_like(x)
    gx = torch.empty_like(x)
    parts = torch.empty(fwd.partials, dtype=torch.float64, device=DEV)
    scale = torch.empty_like(w)
    two = torch.full((), 2.0, dtype=torch.float64, device=DEV)

    def step():
        fwd.call_weighted_sumsq(x, y, parts, post=w)
        torch.mul(w, two, out=scale)
        bwd.call_weighted(y, gx, pre=scale)

    step()
    torch.cuda.synchronize()
    ref_parts, ref_gx = parts.clone(), gx.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for i in range(4):
        parts.zero_()
        gx.zero_()
        junk = torch.randn(1 << 20, device=DEV) * 2.0  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(parts, ref_parts), i
        assert torch.equal(gx, ref_gx), i
        del junk
