"""GPU parity: the HIP path against the oracle on identical seeded inputs.

Bar: bit-exact equality with the oracle's refTransform for the same radix sequence
(every butterfly, twiddle table and pass is the reference's). Accuracy: relative L2
error against the binary128 DFT within 2x of the reference's published error.
Large configurations: size-independent properties (inverse round trip, linearity,
Parseval) plus bit-exact spot checks of whole transforms.
"""

import math

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():  # pragma: no cover - CPU-only containers
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

DEV = "cuda:0"


def to_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


def run(lf, plan, x, inplace=False):
    xd = to_dev(x)
    out = xd if inplace else torch.empty_like(xd)
    plan.call(xd, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def first_diff(a, b):
    a = np.ascontiguousarray(a).view(np.uint64).ravel()
    b = np.ascontiguousarray(b).view(np.uint64).ravel()
    idx = np.nonzero(a != b)[0]
    return None if idx.size == 0 else int(idx[0])


# ------------------------------------------------------------------ bit-exact vs oracle

POW2 = [2 ** k for k in range(1, 19)]
SMOOTH = [3, 5, 6, 7, 9, 10, 11, 12, 13, 14, 15, 60, 360, 1000, 2100, 6, 143, 1960, 4725,
          10368, 27000, 75600, 165375]


@pytest.mark.parametrize("N", POW2 + SMOOTH)
@pytest.mark.parametrize("fwd", [True, False])
def test_default_plan_bit_exact(lf, oracle, N, fwd):
    howmany = max(1, min(8, (1 << 16) // N))
    x = oracle.lcg_input(N, seed=oracle.LCG_SEED ^ N ^ (fwd << 40), howmany=howmany)
    plan = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, howmany)
    got = run(lf, plan, x)
    ref = oracle.ref_transform(fwd, plan.radices, x)
    assert bits_equal(got, ref), (plan.radices, plan.stages, first_diff(got, ref))


# plans the reference chose (results/repro-2026-11-04/xcheck-patient.jsonl, bounds.jsonl)
REF_PLANS = [
    (1024, [8, 4, 4, 8], 0), (1024, [8, 8, 8, 2], 0), (4096, [8, 16, 8, 4], 2),
    (512, [16, 2, 4, 4], 0), (8192, [8, 16, 8, 8], 2), (262144, [8, 8, 8, 8, 8, 8], 3),
    (80, [8, 10], 0), (108, [6, 6, 3], 0), (210, [15, 7, 2], 0), (504, [8, 7, 3, 3], 0),
    (1000, [8, 5, 5, 5], 0), (1960, [8, 7, 7, 5], 0), (4725, [7, 5, 9, 15], 2),
    (10368, [16, 6, 9, 12], 2), (27000, [8, 15, 9, 5, 5], 2), (75600, [8, 7, 9, 10, 15], 2),
    (165375, [15, 15, 15, 7, 7], 2), (6, [6], 0), (12, [12], 0), (15, [15], 0), (16, [16], 0),
]


@pytest.mark.parametrize("N,rs,split", REF_PLANS)
def test_reference_plans_bit_exact(lf, oracle, N, rs, split):
    howmany = max(1, min(4, (1 << 16) // N))
    x = oracle.lcg_input(N, seed=12345 + N, howmany=howmany)
    for fwd in (True, False):
        plan = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, howmany, radices=rs, split=split)
        got = run(lf, plan, x)
        ref = oracle.ref_transform(fwd, rs, x)
        assert bits_equal(got, ref), (rs, split, fwd, first_diff(got, ref))


@pytest.mark.parametrize("N,rs,split", [(4096, [16, 16, 16], 1), (4096, [16, 16, 16], 2),
                                        (2048, [2, 4, 8, 16, 2], 3), (30030, [13, 11, 7, 6, 5], 2),
                                        (65536, [16, 16, 16, 16], 2), (96, [3, 2, 16], 1),
                                        (5120, [8, 8, 2, 8, 5], 3), (10240, [5, 8, 8, 8, 4], 2),
                                        (6144, [6, 2, 8, 8, 8], 2),
                                        (32768, [8, 8, 2, 8, 8, 4], 3),
                                        (131072, [8, 8, 4, 8, 8, 8], 3)])
def test_stage_splits_bit_exact(lf, oracle, N, rs, split):
    """The same schedule cut into kernel stages at different points gives the same bits,
    including register-resident stages with a partial last tile (5120: 40 columns of 128 in
    tiles of 16; 10240: 40 residue classes of 256 in tiles of 16; 6144: 12 residue classes of
    512 in tiles of 8), both directions."""
    x = oracle.lcg_input(N, seed=77 + N, howmany=3)
    for fwd in (True, False):
        plan = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, 3, radices=rs, split=split)
        got = run(lf, plan, x)
        assert bits_equal(got, oracle.ref_transform(fwd, rs, x)), (N, rs, split, fwd)


@pytest.mark.parametrize("N,rs", [(2048, [8, 8, 8, 4]), (4096, [8, 8, 8, 8]), (8192, [8, 8, 8, 8, 2]),
                                  (8192, [8, 4, 4, 4, 4, 4]), (16384, [16, 16, 16, 4]),
                                  (16384, [8, 8, 8, 8, 4]), (16384, [4, 4, 4, 4, 4, 4, 4]),
                                  (4096, [4, 8, 2, 8, 8]), (1024, [2, 4, 8, 4, 4]),
                                  (8192, [16, 16, 8, 4]), (8192, [16, 4, 16, 8]),
                                  (16384, [16, 8, 16, 8])])
def test_register_schedules_bit_exact(lf, oracle, N, rs):
    """The register-resident kernel's fixed (compile-time) schedules and its run-time
    dispatched ones, both directions, in and out of place."""
    x = oracle.lcg_input(N, seed=31 + N + len(rs), howmany=3)
    for fwd in (True, False):
        plan = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, 3, radices=rs)
        ref = oracle.ref_transform(fwd, rs, x)
        assert bits_equal(run(lf, plan, x), ref), (N, rs, fwd)
        assert bits_equal(run(lf, plan, x, inplace=True), ref), (N, rs, fwd, "in place")


def test_inplace_and_batch_tail(lf, oracle):
    for N, howmany in ((1024, 7), (60, 1001), (8192, 5), (2, 3333)):
        x = oracle.lcg_input(N, seed=5 + N, howmany=howmany)
        plan = lf.Plan(N, lf.FORWARD, howmany)
        got = run(lf, plan, x, inplace=True)
        assert bits_equal(got, oracle.ref_transform(True, plan.radices, x)), N


def test_strided_plan_many(lf, oracle):
    """Transposed batch layout (istride = howmany, idist = 1) and a padded output."""
    N, hm = 360, 24
    x = oracle.lcg_input(N, seed=99, howmany=hm)  # (hm, N)
    xt = np.ascontiguousarray(x.T)  # element j of transform h at j*hm + h
    plan = lf.Plan(N, lf.FORWARD, hm, istride=hm, idist=1, ostride=1, odist=N + 8)
    xd = to_dev(xt)
    out = torch.zeros(hm * (N + 8), dtype=torch.complex128, device=DEV)
    plan.call(xd, out)
    got = out.cpu().numpy().reshape(hm, N + 8)
    ref = oracle.ref_transform(True, plan.radices, x)
    assert bits_equal(got[:, :N], ref)
    assert np.all(got[:, N:] == 0)  # nothing written outside the layout


COLUMN_CASES = [(64, 40), (128, 19), (256, 33), (512, 9), (1024, 17), (2048, 5), (360, 12),
                (4096, 24), (8192, 9), (10368, 7), (65536, 3), (1 << 18, 2)]


@pytest.mark.parametrize("N,hm", COLUMN_CASES)
@pytest.mark.parametrize("fwd", [True, False])
def test_column_layout_bit_exact(lf, oracle, N, hm, fwd):
    """Columns of a row-major (N, hm) array (istride = ostride = hm, idist = odist = 1):
    one-stage tiles of adjacent transforms, register stage kernels, and two-stage column
    plans (N > 2048) with partial tiles of transforms; out of place and in place."""
    x = oracle.lcg_input(N, seed=N * 3 + hm, howmany=hm)
    xt = np.ascontiguousarray(x.T)
    plan = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, hm, istride=hm, idist=1, ostride=hm,
                   odist=1)
    ref = np.ascontiguousarray(oracle.ref_transform(fwd, plan.radices, x).T)
    assert bits_equal(run(lf, plan, xt), ref), (plan.radices, plan.stages)
    assert bits_equal(run(lf, plan, xt, inplace=True), ref), (plan.radices, plan.stages)


def test_guard_words_untouched(lf, oracle):
    """Outputs land exactly on [off, off + howmany*N): guard words around them survive."""
    N, hm = 4096, 3
    x = oracle.lcg_input(N, seed=3, howmany=hm)
    plan = lf.Plan(N, lf.FORWARD, hm)
    buf = torch.full((hm * N + 64,), complex(7.0, -7.0), dtype=torch.complex128, device=DEV)
    out = buf[32:32 + hm * N]
    plan.call(to_dev(x).reshape(-1), out)
    got = buf.cpu().numpy()
    assert np.all(got[:32] == complex(7, -7)) and np.all(got[32 + hm * N:] == complex(7, -7))
    assert bits_equal(got[32:32 + hm * N].reshape(hm, N), oracle.ref_transform(True, plan.radices, x))


def test_weighted_transform_bit_exact(lf, oracle):
    N, hm = 4096, 4
    x = oracle.lcg_input(N, seed=11, howmany=hm)
    w = np.random.default_rng(2).uniform(0, 1, N)
    plan = lf.Plan(N, lf.FORWARD, hm)
    xd = to_dev(x)
    wd = torch.from_numpy(w).to(DEV)
    out = torch.empty_like(xd)
    plan.call_weighted(xd, out, pre=wd, post=wd)
    torch.cuda.synchronize()
    xw = x * w  # the same binary64 product the kernel forms on load
    ref = oracle.ref_transform(True, plan.radices, xw) * w
    assert bits_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("N", [8192, 16384])
def test_split_kernel_weights_and_partials(lf, oracle, N):
    """The split register kernel (the default schedules of 8192 and 16384 points) with input
    and output weights, each alone and together, and with per-workgroup sums of |out|^2."""
    hm = 3
    x = oracle.lcg_input(N, seed=19 + N, howmany=hm)
    w = np.random.default_rng(N).uniform(0, 1, N)
    plan = lf.Plan(N, lf.FORWARD, hm)
    assert plan.radices == ([16, 4, 16, 8] if N == 8192 else [16, 8, 16, 8])
    xd = to_dev(x)
    wd = torch.from_numpy(w).to(DEV)
    for pre, post in ((wd, None), (None, wd), (wd, wd)):
        out = torch.empty_like(xd)
        plan.call_weighted(xd, out, pre=pre, post=post)
        torch.cuda.synchronize()
        ref = oracle.ref_transform(True, plan.radices, x * w if pre is not None else x)
        if post is not None:
            ref = ref * w
        assert bits_equal(out.cpu().numpy(), ref), (pre is not None, post is not None)
    parts = torch.full((plan.partials,), float("nan"), dtype=torch.float64, device=DEV)
    out = torch.empty_like(xd)
    plan.call_weighted_sumsq(xd, out, parts, post=wd)
    torch.cuda.synchronize()
    ref = oracle.ref_transform(True, plan.radices, x) * w
    assert bits_equal(out.cpu().numpy(), ref)
    assert plan.partials == hm and not torch.isnan(parts).any()
    rows = (ref.real ** 2 + ref.imag ** 2).sum(axis=1)
    assert np.allclose(parts.cpu().numpy(), rows, rtol=1e-13, atol=0)


def test_spec_plan(lf, oracle):
    """The reference's generic spec: complex, rank 1, vrank 1, interleaved, offsets."""
    N, hm = 1000, 5
    # strides in doubles: element stride 2, batch distance 2N; outputs offset by 4 doubles
    spec = [0, -1, 1, 1, 0, 1, 4, 5, N, 2, 2, hm, 2 * N, 2 * N]
    plan = lf.SpecPlan(spec)
    x = oracle.lcg_input(N, seed=8, howmany=hm)
    xin = torch.from_numpy(x.view(np.float64).reshape(-1).copy()).to(DEV)
    out = torch.zeros(2 * N * hm + 4, dtype=torch.float64, device=DEV)
    plan.call(xin, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()[4:].view(np.complex128).reshape(hm, N)
    ref_plan = lf.Plan(N, lf.FORWARD, hm)
    assert bits_equal(got, oracle.ref_transform(True, ref_plan.radices, x))
    assert lf.plan_problem([1, -1, 1, 0, 0, 1, 0, 1, 8, 1, 1]) is None  # rdft2: unsupported


def test_unsupported_sizes_return_none(lf):
    for n in (0, 1, 17, 34, 19 * 2, (1 << 24) * 2):
        with pytest.raises(lf.MfftError):
            lf.Plan(n, lf.FORWARD, 1)


def test_host_pointer_execute(lf, oracle):
    N, hm = 2100, 3
    x = oracle.lcg_input(N, seed=21, howmany=hm)
    plan = lf.Plan(N, lf.BACKWARD, hm)
    out = np.zeros_like(x)
    plan.execute_host(np.ascontiguousarray(x), out)
    assert bits_equal(out, oracle.ref_transform(False, plan.radices, x))


def test_deterministic_rerun(lf, oracle):
    N, hm = 65536, 16
    x = oracle.lcg_input(N, seed=1, howmany=hm)
    plan = lf.Plan(N, lf.FORWARD, hm)
    a = run(lf, plan, x)
    b = run(lf, plan, x)
    assert bits_equal(a, b)


# ------------------------------------------------------------------ accuracy vs reference


# every size with a published reference L2: 2^1..2^18 and the 18 cross-checked 13-smooth
# sizes (results/repro-2026-11-04/xcheck-patient.jsonl), config 3's published ones included;
# and config 3's unpublished sizes 60, 360 and 2100, whose bar is 2x the nearest published
# size's (log distance: 64, 504 and 2048)
ACCURACY_SIZES = [2 ** k for k in range(1, 19)] + [6, 9, 12, 15, 18, 24, 36, 80, 108, 210, 504,
                                                   1000, 1960, 4725, 10368, 27000, 75600, 165375,
                                                   60, 360, 2100]


def reference_l2(N):
    """(reference L2, the size it was published for). Published sizes: the larger of
    libbench2's --accuracy figure (emr-i12) and the cross-check against long-double FFTW
    (xcheck); unpublished ones: the nearest published size in log distance, over those two
    sets and the non-power-of-two --accuracy run (spr-nonpow2-i9)."""
    import json
    import os

    from conftest import GOLDEN

    acc = json.load(open(os.path.join(GOLDEN, "accuracy_ref.json")))
    pub = {}
    for k, f in (("pow2", "fwd_l2"), ("xcheck", "lean_l2")):
        for n, v in acc[k].items():
            pub.setdefault(int(n), []).append(v[f])
    if N in pub:
        return max(pub[N]), N
    for n, v in acc["nonpow2"]["lean-i8"].items():
        pub.setdefault(int(n), []).append(v["fwd_l2"])
    near = min(pub, key=lambda m: (abs(math.log(m / N)), m))
    return max(pub[near]), near


@pytest.mark.parametrize("N", ACCURACY_SIZES)
def test_accuracy_within_2x_reference(lf, oracle, N):
    import json
    import os

    # at N = 4 libbench2's inputs make the transform exact (3e-44), the cross-check reads
    # 5.6e-17: the bar takes the larger published figure
    ref, ref_n = reference_l2(N)
    rounds = 4 if N >= 65536 else 10
    x = oracle.lcg_input(N, seed=4242 + N, howmany=rounds)
    plan = lf.Plan(N, lf.FORWARD, rounds)
    got = run(lf, plan, x)
    err = oracle.rel_l2(True, x, got)
    log = os.environ.get("MFFT_ACCURACY_LOG")
    if log:  # one JSON line per size, for the per-size table of the design notes
        with open(log, "a") as f:
            f.write(json.dumps({"N": N, "radices": plan.radices, "l2_mean": float(err.mean()),
                                "l2_max": float(err.max()), "reference_l2": ref,
                                "reference_n": ref_n}) + "\n")
    # libbench2's --accuracy averages the relative L1/L2 errors over its rounds (and takes
    # the max of L-inf); the bar is 2x the reference's published figure for the mean, with
    # every single round within 3x.
    # (N = 2 and 4 are exact in binary64: the reference reports 0 and 3e-44, and what is left
    # here is the binary128 comparison's own rounding, so the bar has a floor of 1e-32)
    assert err.mean() <= max(2 * ref, 1e-32), (N, err.mean(), ref, ref_n)
    assert err.max() <= max(3 * ref, 1e-32), (N, err.max(), ref, ref_n)
    # and a floor: a binary64 transform of these inputs has an error of the order of the
    # reference's (0.74-1.3x of it at every size measured), so a metric reading far below it
    # is itself broken
    if N >= 8:
        assert err.mean() >= 0.2 * ref, (N, err.mean(), ref, ref_n)


# ------------------------------------------------------------------ full-size properties


@pytest.mark.parametrize("k", range(1, 19))
def test_full_size_pow2_properties(lf, oracle, k):
    """Config of the batched sweep: N = 2^k, batch 2^24 / N (256 MB). Round trip
    backward(forward(x)) = N x, linearity, Parseval, and a bit-exact check of three whole
    transforms (first, middle, last) against the oracle."""
    N = 1 << k
    hm = (1 << 24) // N
    g = torch.Generator(device=DEV).manual_seed(k)
    x = (torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g) - 0.5) + 1j * (
        torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    fwd = lf.Plan(N, lf.FORWARD, hm)
    bwd = lf.Plan(N, lf.BACKWARD, hm)
    y = torch.empty_like(x)
    fwd.call(x, y)
    z = torch.empty_like(x)
    bwd.call(y, z)
    torch.cuda.synchronize()
    err = ((z / N - x).abs().max() / x.abs().max()).item()
    assert err < 1e-14 * math.log2(N) + 1e-15
    # Parseval: sum |y|^2 = N sum |x|^2
    ex = (x.abs() ** 2).sum().item()
    ey = (y.abs() ** 2).sum().item()
    assert abs(ey - N * ex) <= 1e-12 * N * ex
    # linearity on two rows
    a, b = x[:1], x[-1:]
    s = (a + 2.0 * b).contiguous()
    p1 = lf.Plan(N, lf.FORWARD, 1)
    ya, yb, ys = (torch.empty_like(a) for _ in range(3))
    p1.call(a.contiguous(), ya)
    p1.call(b.contiguous(), yb)
    p1.call(s, ys)
    torch.cuda.synchronize()
    assert ((ys - (ya + 2.0 * yb)).abs().max() / ys.abs().max()).item() < 1e-13
    rows = [0, hm // 2, hm - 1]
    xs = x[rows].cpu().numpy()
    ref = oracle.ref_transform(True, fwd.radices, xs)
    assert bits_equal(y[rows].cpu().numpy(), ref), (N, fwd.radices, fwd.stages)


CONFIG3 = [6, 60, 360, 1000, 2100, 10368, 75600, 165375]


@pytest.mark.parametrize("N", CONFIG3)
def test_full_size_mixed_properties(lf, oracle, N):
    """Config of the mixed-radix batch: N 13-smooth, batch ceil(2^24 / N) (at least 256 MB,
    the XCD-remapped and padded grids of the two-stage sizes at their real extent). Round
    trip backward(forward(x)) = N x, Parseval, and the first, middle and last transforms bit
    for bit against the oracle."""
    hm = -(-(1 << 24) // N)
    g = torch.Generator(device=DEV).manual_seed(N)
    x = (torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g) - 0.5) + 1j * (
        torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    fwd = lf.Plan(N, lf.FORWARD, hm)
    bwd = lf.Plan(N, lf.BACKWARD, hm)
    y = torch.empty_like(x)
    fwd.call(x, y)
    z = torch.empty_like(x)
    bwd.call(y, z)
    torch.cuda.synchronize()
    err = ((z / N - x).abs().max() / x.abs().max()).item()
    assert err < 1e-14 * math.log2(N) + 1e-15
    ex = (x.abs() ** 2).sum().item()
    ey = (y.abs() ** 2).sum().item()
    assert abs(ey - N * ex) <= 1e-12 * N * ex
    rows = [0, hm // 2, hm - 1]
    ref = oracle.ref_transform(True, fwd.radices, x[rows].cpu().numpy())
    assert bits_equal(y[rows].cpu().numpy(), ref), (N, fwd.radices, fwd.stages)
    refb = oracle.ref_transform(False, bwd.radices, y[rows].cpu().numpy())
    assert bits_equal(z[rows].cpu().numpy(), refb), (N, bwd.radices, bwd.stages)


def test_fft_on_two_streams_same_shape(lf):
    """lf.fft keeps a plan per stream, so one shape running on two streams at once (two-stage
    sizes own an intermediate) gives each stream its own plan and the eager results."""
    N, hm = 65536, 64
    g = torch.Generator(device=DEV).manual_seed(9)
    xs = [torch.complex(torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g),
                        torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g))
          for _ in range(2)]
    refs = [lf.fft(x) for x in xs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = []
    for _ in range(3):
        outs = []
        for s, x in zip(streams, xs):
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                outs.append(lf.fft(x))
        for s in streams:
            torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            assert torch.equal(o, r)


@pytest.mark.parametrize("N,hm", [(4096, 32), (65536, 16)])
def test_fft_captured_on_torch_capture_stream(lf, N, hm):
    """lf.fft inside torch.cuda.graph, which captures on a stream of its own: the capture takes
    the plan made before it (creating one would need synchronous copies) and every replay, with
    eager work in between, matches the eager result bit for bit."""
    g = torch.Generator(device=DEV).manual_seed(N + 3)
    x = torch.complex(torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g) - 0.5,
                      torch.rand(hm, N, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    ref = lf.fft(x)
    torch.cuda.synchronize()
    before = len(lf._plan_cache)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = lf.fft(x)
    assert len(lf._plan_cache) == before
    for i in range(3):
        out.zero_()
        junk = torch.randn(1 << 20, device=DEV) * 2.0  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), i
        del junk


# ------------------------------------------------------------------ three kernel stages


@pytest.mark.parametrize("N,rs,passes", [(4096, [16, 16, 16], [1, 1, 1]), (360, [3, 4, 2, 3, 5], [2, 1, 2]),
                                         (210, [5, 7, 2, 3], [1, 1, 2]), (65536, [8, 8, 4, 8, 8, 4], [2, 2, 2]),
                                         (10368, [12, 8, 12, 9], [1, 2, 1])])
@pytest.mark.parametrize("fwd", [True, False])
def test_three_stage_forced_bit_exact(lf, oracle, N, rs, passes, fwd):
    """Small sizes forced into three kernel stages (x -> T -> U -> y; the middle stage with
    sub-problems along grid.y and transforms along grid.z): bit-exact against refTransform
    with the same radices, out of place and in place, a batch of several transforms."""
    hm = 5
    x = oracle.lcg_input(N, seed=N ^ 77 ^ (fwd << 3), howmany=hm)
    p = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, hm, radices=rs, passes=passes)
    assert p.stages == passes
    ref = oracle.ref_transform(fwd, rs, x)
    assert bits_equal(run(lf, p, x), ref)
    assert bits_equal(run(lf, p, x, inplace=True), ref)


@pytest.mark.parametrize("N", [371293, 1 << 23])
def test_three_stage_sizes_bit_exact(lf, oracle, N):
    """Sizes no two-way split fits (13^5, 2^23) run as three stages by default, bit-exact
    against the oracle (one forward transform each; its exact-integer tables take seconds)."""
    p = lf.Plan(N, lf.FORWARD, 2)
    assert len(p.stages) == 3, p.stages
    x = oracle.lcg_input(N, seed=N % 1009, howmany=2)
    assert bits_equal(run(lf, p, x)[1], oracle.ref_transform(True, p.radices, x[1]))


@pytest.mark.parametrize("N", [1 << 24, 3 ** 14, 13 ** 6])
def test_three_stage_largest_properties(lf, N):
    """The largest sizes of the reference's fast domain (2^24 = 256 MiB per transform, 3^14,
    13^6): round trip backward(forward(x)) = N x and Parseval."""
    p = lf.Plan(N, lf.FORWARD, 1)
    q = lf.Plan(N, lf.BACKWARD, 1)
    assert len(p.stages) == 3
    g = torch.Generator(device=DEV).manual_seed(N % 997)
    x = (torch.rand(N, dtype=torch.float64, device=DEV, generator=g) - 0.5) + 1j * (
        torch.rand(N, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    p.call(x, y)
    q.call(y, z)
    torch.cuda.synchronize()
    assert ((z / N - x).abs().max() / x.abs().max()).item() < 1e-14 * math.log2(N)
    ex, ey = (x.abs() ** 2).sum().item(), (y.abs() ** 2).sum().item()
    assert abs(ey - N * ex) <= 1e-11 * N * ex
