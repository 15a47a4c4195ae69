"""GPU parity of the composite transforms: real-data (r2c / c2r) and 2-D complex.

The reference's generic engine computes both through its complex line transform
(LeanFFTW/Core/Problem.lean:142-149 for LineOp.r2c / c2r, :165-182 alongAxis, :222-238
planProblemCore). Bar: bit-exact equality with the same composition evaluated with the
oracle's refTransform for the plans' radix sequences (embedding, first n/2 + 1 outputs,
Hermitian extension, real part, axis 0 then axis 1), and agreement with numpy's
independent FFTs (pocketfft) within a relative L2 of 1e-14, which pins the index maps
themselves (a wrong half length, mirror index or axis order fails it).
"""

import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():  # pragma: no cover - CPU-only containers
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

DEV = "cuda:0"
TOL = 1e-14


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rel_l2(got, ref):
    return float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))


def real_input(n, hm, seed):
    return np.random.default_rng(seed).standard_normal((hm, n))


def half_input(n, hm, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((hm, n // 2 + 1)) + 1j * rng.standard_normal((hm, n // 2 + 1))


def herm_ext(h, n):
    """hermExt (Problem.lean:138-140): h_k for k <= n/2, conj(h_{n-k}) above."""
    k = np.arange(n)
    return np.where(k <= n // 2, h[:, np.minimum(k, n // 2)],
                    np.conj(h[:, np.minimum((n - k) % n, n // 2)]))


# powers of two from 128 on run with the maps fused into the register kernels (16384 through
# the embedded batch on the input side), batches that fill a staged tile (powers of two up to
# 64 points, every other single-stage size up to 4096) with the maps in the tile copies
# (60, 105, 1000, 2100 here; every generated shape in tests/test_gpu_tile.py), the two-stage
# sizes 10368, 75600 and 165375 with the maps in their tile stages' first loads and last
# stores; the others (smaller batches, other two-stage sizes that are not powers of two)
# through the element-map kernels
REAL_SIZES = [2, 3, 4, 5, 8, 15, 16, 32, 60, 64, 105, 128, 256, 512, 1000, 1024, 2048, 2100,
              4096, 8192, 10368, 16384, 32768, 65536, 75600, 131072, 165375, 262144]


@pytest.mark.parametrize("n", REAL_SIZES)
def test_r2c_bit_exact(lf, oracle, n):
    hm = max(1, min(6, (1 << 17) // n))
    x = real_input(n, hm, seed=n)
    plan = lf.RealPlan(n, lf.FORWARD, hm)
    got = plan.exec(torch.from_numpy(x).to(DEV))
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    # the plan's radices: a real-data plan keeps the register-kernel schedule of its complex
    # part (at 1024 and 8192 points not the plain complex default)
    full = oracle.ref_transform(True, plan.radices, x.astype(np.complex128))
    assert bits_equal(got, full[:, : n // 2 + 1])
    assert rel_l2(got, np.fft.rfft(x, axis=1)) < TOL


@pytest.mark.parametrize("n", REAL_SIZES)
def test_c2r_bit_exact(lf, oracle, n):
    hm = max(1, min(6, (1 << 17) // n))
    h = half_input(n, hm, seed=n + 1)
    plan = lf.RealPlan(n, lf.BACKWARD, hm)
    got = plan.exec(torch.from_numpy(h).to(DEV))
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    full = oracle.ref_transform(False, plan.radices, herm_ext(h, n))
    assert bits_equal(got, np.ascontiguousarray(full.real))
    # numpy's irfft drops the imaginary parts of bins 0 and n/2, as the real part does exactly
    assert rel_l2(got, np.fft.irfft(h, n, axis=1) * n) < TOL


def test_real_round_trip_full_size(lf):
    """c2r(r2c(x)) = n x at 2^18 points, 16 transforms (a size-independent property)."""
    n, hm = 1 << 18, 16
    x = torch.randn(hm, n, dtype=torch.float64, device=DEV, generator=torch.Generator(DEV).manual_seed(3))
    fwd, bwd = lf.RealPlan(n, lf.FORWARD, hm), lf.RealPlan(n, lf.BACKWARD, hm)
    y = bwd.exec(fwd.exec(x))
    torch.cuda.synchronize()
    err = float(torch.linalg.vector_norm(y / n - x) / torch.linalg.vector_norm(x))
    assert err < 1e-15 * 18


@pytest.mark.parametrize("n", [2100, 4095, 10368, 75600, 165375])
def test_real_full_batch_fused_maps(lf, oracle, n):
    """The fused real-data paths of sizes that are not powers of two (staged tiles with the maps
    in their copies: 2100, 4095; tile stages with the maps in their first / last passes: 10368,
    75600, 165375) at the full batch of 2^24 / n transforms: c2r(r2c(x)) = n x, and the first,
    middle and last transforms of each direction bit for bit against the oracle."""
    hm = (1 << 24) // n
    g = torch.Generator(DEV).manual_seed(n)
    x = torch.rand(hm, n, dtype=torch.float64, device=DEV, generator=g) - 0.5
    fwd, bwd = lf.RealPlan(n, lf.FORWARD, hm), lf.RealPlan(n, lf.BACKWARD, hm)
    h = fwd.exec(x)
    y = bwd.exec(h)
    torch.cuda.synchronize()
    err = float(torch.linalg.vector_norm(y / n - x) / torch.linalg.vector_norm(x))
    assert err < 1e-15 * 20, err
    rows = [0, hm // 2, hm - 1]
    xs = x[rows].cpu().numpy()
    full = oracle.ref_transform(True, fwd.radices, xs.astype(np.complex128))
    assert bits_equal(h[rows].cpu().numpy(), full[:, : n // 2 + 1]), (n, "r2c")
    hs = h[rows].cpu().numpy()
    back = oracle.ref_transform(False, bwd.radices, herm_ext(hs, n))
    assert bits_equal(y[rows].cpu().numpy(), np.ascontiguousarray(back.real)), (n, "c2r")


def test_real_plan_rejects_in_place_and_composite_only_entry_points(lf):
    plan = lf.RealPlan(64, lf.FORWARD, 2)
    assert lf.lib().mfft_plan_kind(plan._h) == 1
    assert lf.lib().mfft_plan_kind(lf.RealPlan(64, lf.BACKWARD, 2)._h) == 2
    assert lf.lib().mfft_plan_kind(lf.Plan2D(4, 8)._h) == 3
    assert lf.lib().mfft_plan_kind(lf.Plan(64)._h) == 0
    buf = torch.zeros(4 * 64, dtype=torch.float64, device=DEV)
    p = ctypes.c_void_p(buf.data_ptr())
    assert lf.lib().mfft_execute(plan._h, p, p, None) == 2  # MFFT_ERR_UNSUPPORTED
    assert lf.lib().mfft_execute_weighted(plan._h, p, p, None, None, None) == 2
    assert lf.lib().mfft_plan_partials(plan._h) == -1


def test_real_empty_batch_and_host_execute(lf, oracle):
    n = 1000
    empty = lf.RealPlan(n, lf.FORWARD, 0)
    z = torch.zeros(1, device=DEV, dtype=torch.complex128)
    assert lf.lib().mfft_execute(empty._h, ctypes.c_void_p(z.data_ptr()),
                                 ctypes.c_void_p(z.data_ptr() + 16), None) == 0
    x = real_input(n, 3, seed=5)
    out = np.zeros((3, n // 2 + 1), dtype=np.complex128)
    dp = ctypes.POINTER(ctypes.c_double)
    plan = lf.RealPlan(n, lf.FORWARD, 3)
    assert lf.lib().mfft_execute_host(plan._h, x.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    full = oracle.ref_transform(True, plan.radices, x.astype(np.complex128))
    assert bits_equal(out, full[:, : n // 2 + 1])


def oracle_2d(lf, oracle, fwd, x):
    """Axis 0 then axis 1 with the oracle, each with the radices of the plan's part."""
    hm, n0, n1 = x.shape
    sign = lf.FORWARD if fwd else lf.BACKWARD
    rs0 = lf.Plan(n0, sign, n1, istride=n1, idist=1, ostride=n1, odist=1).radices
    rs1 = lf.Plan(n1, sign, n0 * hm, istride=1, idist=n1, ostride=1, odist=n1).radices
    cols = np.ascontiguousarray(x.transpose(0, 2, 1)).reshape(hm * n1, n0)
    a = oracle.ref_transform(fwd, rs0, cols).reshape(hm, n1, n0).transpose(0, 2, 1)
    rows = np.ascontiguousarray(a).reshape(hm * n0, n1)
    return oracle.ref_transform(fwd, rs1, rows).reshape(hm, n0, n1)


# (128, 128), (64, 256), (256, 64), (64, 128), (128, 64), (32, 256), (256, 32): the
# single-pass register kernel (both axes in one workgroup)
SHAPES_2D = [(2, 2, 1), (3, 5, 2), (8, 8, 3), (16, 1000, 1), (64, 64, 4), (12, 4096, 1),
             (128, 128, 5), (64, 256, 3), (256, 64, 3), (64, 128, 4), (128, 64, 4), (32, 256, 3),
             (256, 32, 3), (256, 256, 2), (1000, 12, 1), (1024, 1024, 1)]


@pytest.mark.parametrize("n0,n1,hm", SHAPES_2D)
@pytest.mark.parametrize("fwd", [True, False])
def test_2d_bit_exact(lf, oracle, n0, n1, hm, fwd):
    rng = np.random.default_rng(n0 * 7919 + n1 + fwd)
    x = rng.standard_normal((hm, n0, n1)) + 1j * rng.standard_normal((hm, n0, n1))
    plan = lf.Plan2D(n0, n1, lf.FORWARD if fwd else lf.BACKWARD, hm)
    got = plan.exec(torch.from_numpy(x).to(DEV))
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    assert bits_equal(got, oracle_2d(lf, oracle, fwd, x))
    ref = np.fft.fft2(x) if fwd else np.fft.ifft2(x) * (n0 * n1)
    assert rel_l2(got, ref) < TOL


@pytest.mark.parametrize("n0,n1,hm", [(128, 96, 2), (128, 128, 3), (256, 32, 2)])
def test_2d_in_place(lf, oracle, n0, n1, hm):
    rng = np.random.default_rng(11 + n1)
    x = rng.standard_normal((hm, n0, n1)) + 1j * rng.standard_normal((hm, n0, n1))
    plan = lf.Plan2D(n0, n1, lf.FORWARD, hm)
    xd = torch.from_numpy(x).to(DEV)
    plan.call(xd, xd)
    torch.cuda.synchronize()
    assert bits_equal(xd.cpu().numpy(), oracle_2d(lf, oracle, True, x))


def test_2d_round_trip_full_size(lf):
    """Backward(forward(x)) = n0 n1 x for one 2048 x 2048 array."""
    n0 = n1 = 2048
    g = torch.Generator(DEV).manual_seed(7)
    x = torch.randn(n0, n1, dtype=torch.complex128, device=DEV, generator=g)
    y = lf.Plan2D(n0, n1, lf.BACKWARD).exec(lf.Plan2D(n0, n1, lf.FORWARD).exec(x))
    torch.cuda.synchronize()
    err = float(torch.linalg.vector_norm(y / (n0 * n1) - x) / torch.linalg.vector_norm(x))
    assert err < 1e-15 * 22


def test_spec_rdft2_and_2d(lf, oracle):
    """The reference's generic spec (FFI.lean:13-21) for rdft2 rank 1 and dft rank 2."""
    n, hm, nh = 1000, 3, 501
    spec = [1, -1, 1, 1, 0, 0, 2, 3, n, 1, 2, hm, n, 2 * nh]  # r2c, output offset 2 doubles
    plan = lf.SpecPlan(spec)
    x = real_input(n, hm, seed=9)
    out = torch.zeros(2 * nh * hm + 2, dtype=torch.float64, device=DEV)
    plan.call(torch.from_numpy(x.ravel().copy()).to(DEV), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()[2:].view(np.complex128).reshape(hm, nh)
    full = oracle.ref_transform(True, lf.Plan(n, lf.FORWARD, hm).radices, x.astype(np.complex128))
    assert bits_equal(got, full[:, :nh])

    spec = [1, 1, 1, 1, 0, 1, 0, 0, n, 2, 1, hm, 2 * nh, n]  # c2r
    plan = lf.SpecPlan(spec)
    h = half_input(n, hm, seed=10)
    out = torch.zeros(n * hm, dtype=torch.float64, device=DEV)
    plan.call(torch.from_numpy(h.view(np.float64).ravel().copy()).to(DEV), out)
    torch.cuda.synchronize()
    full = oracle.ref_transform(False, lf.Plan(n, lf.BACKWARD, hm).radices, herm_ext(h, n))
    assert bits_equal(out.cpu().numpy().reshape(hm, n), np.ascontiguousarray(full.real))

    n0, n1 = 24, 40
    spec = [0, -1, 2, 1, 0, 1, 0, 1, n0, 2 * n1, 2 * n1, n1, 2, 2, 2, 2 * n0 * n1, 2 * n0 * n1]
    plan = lf.SpecPlan(spec)
    rng = np.random.default_rng(12)
    x = rng.standard_normal((2, n0, n1)) + 1j * rng.standard_normal((2, n0, n1))
    out = torch.zeros(2 * n0 * n1 * 2, dtype=torch.float64, device=DEV)
    plan.call(torch.from_numpy(x.view(np.float64).ravel().copy()).to(DEV), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.complex128).reshape(2, n0, n1)
    assert bits_equal(got, oracle_2d(lf, oracle, True, x))

    # layouts outside the contiguous forms stay with the caller's own path
    assert lf.plan_problem([1, -1, 1, 0, 0, 1, 0, 1, 8, 1, 1]) is None
    assert lf.plan_problem([0, -1, 2, 0, 0, 1, 0, 1, 4, 2, 2, 4, 8, 8]) is None
