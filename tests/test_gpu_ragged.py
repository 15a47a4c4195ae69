"""Ragged batches (two persistent launches from size-sorted queues): every transform
bit-identical to the oracle's refTransform with the radix sequence the uniform path uses for
that size."""

import numpy as np
import pytest

from conftest import gpu_available, load_product

pytestmark = pytest.mark.gpu

if not gpu_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

DEV = "cuda:0"


def check_ragged(lf, oracle, sizes, sign, seed, reps=1):
    sizes = [int(s) for s in sizes]
    total = sum(sizes)
    rp = lf.RaggedPlan(sizes, sign)
    x = oracle.lcg_input(total, seed=seed)
    xd = torch.from_numpy(x).to(DEV)
    out = torch.full_like(xd, complex(3.0, 3.0))
    for _ in range(reps):
        rp.call(xd, out)
    rp.status()
    got = out.cpu().numpy()
    off = 0
    radices = {}
    for n in sorted(set(sizes)):
        radices[n] = rp.class_plan(n)[0]
    bad = []
    for i, n in enumerate(sizes):
        ref = oracle.ref_transform(sign < 0, radices[n], x[off:off + n])
        if not np.array_equal(got[off:off + n].view(np.uint64), ref.view(np.uint64)):
            bad.append((i, n))
        off += n
    assert not bad, bad[:10]


def test_ragged_small_mix(lf, oracle):
    sizes = [8, 16, 6, 1000, 8, 2048, 12, 360, 13, 15, 2, 3, 7, 1024, 2100, 60, 8, 8, 9]
    check_ragged(lf, oracle, sizes, -1, 1)


def test_ragged_two_stage_sizes(lf, oracle):
    sizes = [4096, 65536, 10368, 3, 65520, 2187, 4096, 30030, 12, 2401, 16384, 6]
    check_ragged(lf, oracle, sizes, -1, 2)
    check_ragged(lf, oracle, sizes, 1, 3)


def test_ragged_benchmark_draw_prefix(lf, oracle):
    """The first 600 sizes of the benchmark's log-uniform draw, run twice (the queue heads
    are reset by every call)."""
    w = load_product().workloads
    sizes = w.ragged_sizes()[:600]
    check_ragged(lf, oracle, sizes, -1, 4, reps=2)


def test_ragged_full_draw_against_uniform(lf):
    """The whole 10,000-transform draw: the persistent launches equal per-transform plain
    launches of the uniform path with the same schedules, bit for bit."""
    w = load_product().workloads
    sizes = w.ragged_sizes()
    total = sum(sizes)
    g = torch.Generator(device=DEV).manual_seed(5)
    x = torch.complex(torch.rand(total, dtype=torch.float64, device=DEV, generator=g) - 0.5,
                      torch.rand(total, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    out = torch.empty_like(x)
    rp = lf.RaggedPlan(sizes, -1)
    rp.call(x, out)
    rp.status()
    ref = torch.empty_like(x)
    plans = {}
    off = 0
    for n in sizes:
        p = plans.get(n)
        if p is None:
            rs, split = rp.class_plan(n)
            p = plans[n] = lf.Plan(n, -1, 1, radices=rs, split=split)
        p.call(x[off:off + n], ref[off:off + n])
        off += n
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.float64), ref.view(torch.float64))


def test_ragged_full_draw_against_oracle(lf, oracle):
    """The whole 10,000-transform benchmark draw against the oracle: the first transform of
    every one of its 1,424 sizes and every 50th transform, bit for bit (the oracle runs in 8
    worker processes: its exact-integer twiddle tables for 1,424 sizes take minutes on one
    core)."""
    w = load_product().workloads
    sizes = [int(n) for n in w.ragged_sizes()]
    total = sum(sizes)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    g = torch.Generator(device=DEV).manual_seed(6)
    x = torch.complex(torch.rand(total, dtype=torch.float64, device=DEV, generator=g) - 0.5,
                      torch.rand(total, dtype=torch.float64, device=DEV, generator=g) - 0.5)
    out = torch.empty_like(x)
    rp = lf.RaggedPlan(sizes, -1)
    rp.call(x, out)
    rp.status()
    first = {}
    for i, n in enumerate(sizes):
        first.setdefault(n, i)
    pick = sorted(set(first.values()) | set(range(0, len(sizes), 50)))
    xs = x.cpu().numpy()
    got = out.cpu().numpy()
    items = [(rp.class_plan(sizes[i])[0], xs[starts[i]:starts[i] + sizes[i]]) for i in pick]
    refs = oracle.ref_transform_parallel(True, items, workers=8)
    bad = [(i, sizes[i]) for i, r in zip(pick, refs)
           if not np.array_equal(got[starts[i]:starts[i] + sizes[i]].view(np.uint64),
                                 r.view(np.uint64))]
    assert len(pick) >= 1424 + 150 and not bad, bad[:10]


def test_ragged_rejects_inplace_and_bad_sizes(lf):
    with pytest.raises(lf.MfftError):
        lf.RaggedPlan([8, 17, 16], -1)
    rp = lf.RaggedPlan([8, 4096], -1)
    x = torch.zeros(4104, dtype=torch.complex128, device=DEV)
    with pytest.raises(lf.MfftError):
        rp.call(x, x)
