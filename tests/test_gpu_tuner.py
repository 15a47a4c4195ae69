"""The measuring planner (MFFT_MEASURE): whatever schedule it times fastest, the result is
bit-identical to the oracle's refTransform with that radix sequence; the choice is cached in
the JSON-lines wisdom file and read back; lines that do not describe a valid schedule for
the size are ignored."""

import json

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


def run_plan(p, x):
    xd = torch.from_numpy(x).to(DEV)
    y = torch.empty_like(xd)
    p.call(xd, y)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("n,sign", [(1024, -1), (360, 1), (8192, -1), (65536, -1), (30030, 1)])
def test_measured_plan_bit_exact(lf, oracle, tmp_path, n, sign):
    lf.set_wisdom_path(str(tmp_path / "wisdom.jsonl"))
    try:
        hm = 3
        p = lf.Plan(n, sign, hm, flags=lf.MEASURE)
        rs = p.radices
        assert int(np.prod(rs)) == n
        x = oracle.lcg_input(n, seed=n % 97, howmany=hm)
        assert bits_equal(run_plan(p, x), oracle.ref_transform(sign < 0, rs, x))
    finally:
        lf.set_wisdom_path(None)


def test_wisdom_round_trip_and_invalid_lines(lf, oracle, tmp_path):
    path = tmp_path / "wisdom.jsonl"
    lf.set_wisdom_path(str(path))
    try:
        n, hm = 2048, 5
        p1 = lf.Plan(n, -1, hm, flags=lf.MEASURE)
        lines = [json.loads(s) for s in path.read_text().splitlines()]
        assert len(lines) == 1 and lines[0]["rs"] == p1.radices
        key = lines[0]["key"]
        # a second plan of the same class reads the cache instead of measuring again
        p2 = lf.Plan(n, -1, hm, flags=lf.MEASURE)
        assert p2.radices == p1.radices
        assert len(path.read_text().splitlines()) == 1
        # a hand-written line with another valid schedule is obeyed (the last valid line wins)
        forced = [2] * 11
        with open(path, "a") as f:
            f.write(json.dumps({"key": key, "rs": forced, "split": 0, "us": 1.0}) + "\n")
            # invalid: wrong product, radix 17, split past the end -- all ignored
            f.write(json.dumps({"key": key, "rs": [4, 4], "split": 0, "us": 0.5}) + "\n")
            f.write(json.dumps({"key": key, "rs": [17, 2], "split": 0, "us": 0.5}) + "\n")
            f.write(json.dumps({"key": key, "rs": forced, "split": 11, "us": 0.5}) + "\n")
        p3 = lf.Plan(n, -1, hm, flags=lf.MEASURE)
        assert p3.radices == forced
        x = oracle.lcg_input(n, seed=3, howmany=hm)
        assert bits_equal(run_plan(p3, x), oracle.ref_transform(True, forced, x))
    finally:
        lf.set_wisdom_path(None)


def test_measured_strided_plan(lf, oracle, tmp_path):
    """Transposed layout (istride = howmany, idist = 1) measured on scratch buffers of its own
    span; the user's arrays are never touched during planning."""
    lf.set_wisdom_path(str(tmp_path / "wisdom.jsonl"))
    try:
        n, hm = 512, 6
        x = oracle.lcg_input(n, seed=9, howmany=hm)
        xt = np.ascontiguousarray(x.reshape(hm, n).T).reshape(-1)
        p = lf.Plan(n, -1, hm, istride=hm, idist=1, ostride=hm, odist=1, flags=lf.MEASURE)
        xd = torch.from_numpy(xt).to(DEV)
        y = torch.empty_like(xd)
        p.call(xd, y)
        torch.cuda.synchronize()
        got = y.cpu().numpy().reshape(n, hm).T.reshape(-1)
        assert bits_equal(got, oracle.ref_transform(True, p.radices, x).reshape(-1))
    finally:
        lf.set_wisdom_path(None)


TUNE_SIZES = [2 ** k for k in range(1, 19)] + [6, 60, 360, 1000, 2100, 10368, 75600, 165375]


def _time(plan, x, y, reps=3, k=10):
    """us per call: the fastest of `reps` timings of `k` back-to-back calls."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        for _ in range(k):
            plan.call(x, y)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e3 / k)
    return best


@pytest.mark.parametrize("n", TUNE_SIZES)
def test_measure_not_slower_than_estimate(lf, tmp_path, n):
    """MEASURE races the planner's own schedule (always a candidate) against the cost-model
    schedules and every tile kernel of their stages; on a 64 MB batch the plan it keeps is no
    slower than ESTIMATE's (both timed interleaved, min of 5 rounds of 3 timings of 10
    back-to-back calls; 5% + 1 us of noise allowed)."""
    lf.set_wisdom_path(str(tmp_path / "wisdom.jsonl"))
    try:
        hm = max(1, (1 << 22) // n)
        pe = lf.Plan(n, -1, hm)
        pm = lf.Plan(n, -1, hm, flags=lf.MEASURE)
        g = torch.Generator(device=DEV).manual_seed(n)
        x = torch.complex(torch.rand(hm, n, dtype=torch.float64, device=DEV, generator=g),
                          torch.rand(hm, n, dtype=torch.float64, device=DEV, generator=g))
        y = torch.empty_like(x)
        for p in (pe, pm):
            p.call(x, y)
        te = tm = float("inf")
        for _ in range(5):
            te = min(te, _time(pe, x, y))
            tm = min(tm, _time(pm, x, y))
        assert tm <= 1.05 * te + 1.0, (n, pe.radices, pe.stages, pm.radices, pm.stages, te, tm)
        line = json.loads((tmp_path / "wisdom.jsonl").read_text().splitlines()[-1])
        assert line["rs"] == pm.radices and len(line["tg"]) == 2 and len(line["tn"]) == 2
    finally:
        lf.set_wisdom_path(None)


def test_wisdom_tile_choice_is_obeyed(lf, oracle, tmp_path):
    """A wisdom line naming the kernel of each stage: 'no tile' runs the register kernel, a
    tile that does not fit falls back to the planner's own kernel; either way the result is
    the oracle's for that radix sequence."""
    path = tmp_path / "wisdom.jsonl"
    lf.set_wisdom_path(str(path))
    try:
        n, hm = 1024, 64
        lf.Plan(n, -1, hm, flags=lf.MEASURE)
        key = json.loads(path.read_text().splitlines()[-1])["key"]
        x = oracle.lcg_input(n, seed=12, howmany=hm)
        for tg, tn in (([0, -1], [0, -1]), ([5, -1], [999, -1]), ([2, -1], [256, -1])):
            with open(path, "a") as f:
                f.write(json.dumps({"key": key, "rs": [8, 8, 8, 2], "split": 0, "tg": tg,
                                    "tn": tn, "us": 1.0}) + "\n")
            p = lf.Plan(n, -1, hm, flags=lf.MEASURE)
            assert p.radices == [8, 8, 8, 2]
            assert bits_equal(run_plan(p, x), oracle.ref_transform(True, [8, 8, 8, 2], x))
    finally:
        lf.set_wisdom_path(None)


@pytest.mark.parametrize("n", [10368, 75600, 165375, 32768, 262144])
def test_measure_not_slower_at_full_batch(lf, tmp_path, n):
    """At the benchmark's own batch (ceil(2^24 / n) transforms, 256 MB), where tile widths and
    the XCD-aware order depend on the launch's sub-problem count, MEASURE (timed at that batch)
    is no slower than ESTIMATE."""
    lf.set_wisdom_path(str(tmp_path / "wisdom.jsonl"))
    try:
        hm = -(-(1 << 24) // n)
        pe = lf.Plan(n, -1, hm)
        pm = lf.Plan(n, -1, hm, flags=lf.MEASURE)
        g = torch.Generator(device=DEV).manual_seed(n)
        x = torch.complex(torch.rand(hm, n, dtype=torch.float64, device=DEV, generator=g),
                          torch.rand(hm, n, dtype=torch.float64, device=DEV, generator=g))
        y = torch.empty_like(x)
        te = tm = float("inf")
        for _ in range(4):
            te = min(te, _time(pe, x, y, reps=2, k=4))
            tm = min(tm, _time(pm, x, y, reps=2, k=4))
        assert tm <= 1.03 * te + 1.0, (n, pe.radices, pe.stages, pm.radices, pm.stages, te, tm)
    finally:
        lf.set_wisdom_path(None)
