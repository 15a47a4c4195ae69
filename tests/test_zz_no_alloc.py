"""mfft_execute_static (the allocation-free entry point) against mfft_execute, through the C
ABI: bit-identical outputs on consecutive calls (the second smaller than the first, so the
shared static scratch holds the previous call's contents), on batches larger than the
scratch (chunked), on column and 2-D plans, on fused real-data plans, and replayed from a
HIP graph."""

import ctypes

import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

DEV = "cuda:0"
SCRATCH = 64 << 20  # MFFT_STATIC_SCRATCH_BYTES


def rand(shape, seed, dtype=torch.complex128):
    g = torch.Generator(device=DEV).manual_seed(seed)
    if dtype == torch.float64:
        return torch.rand(shape, dtype=torch.float64, device=DEV, generator=g) - 0.5
    return torch.complex(torch.rand(shape, dtype=torch.float64, device=DEV, generator=g) - 0.5,
                         torch.rand(shape, dtype=torch.float64, device=DEV, generator=g) - 0.5)


def abi_call(lf, name, plan, x, out):
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(lf.lib(), name)(plan._h, ctypes.c_void_p(x.data_ptr()),
                                 ctypes.c_void_p(out.data_ptr()), stream)
    torch.cuda.synchronize()
    return rc


def same(a, b):
    return torch.equal(a.view(torch.float64), b.view(torch.float64))


def check(lf, plan, x, out_like):
    ref = torch.full_like(out_like, 7.0)
    got = torch.full_like(out_like, -7.0)
    assert abi_call(lf, "mfft_execute", plan, x, ref) == 0
    assert abi_call(lf, "mfft_execute_static", plan, x, got) == 0
    assert same(got, ref), (plan.radices, plan.stages if hasattr(plan, "stages") else None)


def test_static_matches_execute_two_calls(lf):
    """A batch whose intermediate is 4.7x the scratch (runs in 5 chunks), then a smaller one
    of another size, then an in-place call."""
    for n, hm, seed in ((32768, 600, 1), (262144, 9, 2), (65536, 3, 3)):
        assert hm * n * 16 > SCRATCH or (n, hm) != (32768, 600)
        x = rand((hm, n), seed)
        plan = lf.Plan(n, lf.FORWARD, hm)
        assert len(plan.stages) == 2
        check(lf, plan, x, x)
    n, hm = 131072, 20
    x = rand((hm, n), 4)
    plan = lf.Plan(n, lf.BACKWARD, hm)
    ref = torch.empty_like(x)
    assert abi_call(lf, "mfft_execute", plan, x, ref) == 0
    y = x.clone()
    assert abi_call(lf, "mfft_execute_static", plan, y, y) == 0
    assert same(y, ref)


def test_static_single_stage_and_composites(lf):
    """Single-stage plans (no intermediate), a two-stage column layout, a 2-D plan whose
    columns run two stages, and real-data plans whose maps fuse (into register kernels, into
    a staged tile); an unfused real-data plan is refused with MFFT_ERR_UNSUPPORTED."""
    x = rand((1000, 1024), 5)
    check(lf, lf.Plan(1024, lf.FORWARD, 1000), x, x)
    # columns of a 8192 x 64 row-major array: 64 transforms of 8192 points, stride 64
    a = rand((8192, 64), 6)
    cols = lf.Plan(8192, lf.FORWARD, 64, istride=64, idist=1, ostride=64, odist=1)
    check(lf, cols, a, a)
    b = rand((3, 4096, 512), 7)
    p2 = lf.Plan2D(4096, 512, lf.FORWARD, 3)
    check(lf, p2, b, b)
    r = rand((40, 65536), 8, torch.float64)
    pr = lf.RealPlan(65536, lf.FORWARD, 40)
    out = torch.empty((40, 65536 // 2 + 1), dtype=torch.complex128, device=DEV)
    check(lf, pr, r, out)
    pt = lf.RealPlan(360, lf.FORWARD, 8)  # the maps in a staged tile's copies
    rt = rand((8, 360), 9, torch.float64)
    check(lf, pt, rt, torch.empty((8, 181), dtype=torch.complex128, device=DEV))
    p2 = lf.RealPlan(10368, lf.FORWARD, 3)  # the maps in two tile stages (an intermediate)
    r2 = rand((3, 10368), 12, torch.float64)
    check(lf, p2, r2, torch.empty((3, 5185), dtype=torch.complex128, device=DEV))
    pu = lf.RealPlan(5000, lf.FORWARD, 2)  # two stages without map variants: no fused maps
    ru = rand((2, 5000), 11, torch.float64)
    outu = torch.empty((2, 2501), dtype=torch.complex128, device=DEV)
    assert abi_call(lf, "mfft_execute_static", pu, ru, outu) == 2  # MFFT_ERR_UNSUPPORTED


def test_static_replayed_from_graph(lf):
    """Captured in a HIP graph and replayed with eager work in between: every replay equals
    the eager mfft_execute result."""
    n, hm = 65536, 150
    x = rand((hm, n), 10)
    plan = lf.Plan(n, lf.FORWARD, hm)
    ref = torch.empty_like(x)
    assert abi_call(lf, "mfft_execute", plan, x, ref) == 0
    out = torch.empty_like(x)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.call_static(x, out)  # first call on the device: the symbol lookup
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.call_static(x, out)
    for i in range(3):
        out.zero_()
        junk = torch.randn(1 << 20, device=DEV) * 2.0
        graph.replay()
        torch.cuda.synchronize()
        assert same(out, ref), i
        del junk
