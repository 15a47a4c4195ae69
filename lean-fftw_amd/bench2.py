#!/usr/bin/env python3
"""libbench2-shaped command line for the GPU library.

The reference benchmarks every implementation as a libbench2 backend
(bench/backends/bench-leanfftw.c:57-306) driven by its protocol runner, which runs
``CMD -s <problem>`` and ``CMD -v1 --verify <problem>`` and parses their output
(tools/bench.py:45-81). This program accepts the same arguments and prints the same
lines, so the runner can take the GPU path as one more implementation::

    python3 tools/bench.py run --impl gpu="python3 /path/to/lean-fftw_amd/bench2.py" ...

Problems (libbench2's problem strings, the subset the protocol uses): ``[i|o][c|r][f|b]N``
and ``[i|o]c[f|b]NxM`` -- in place or out of place, complex or real (r2c forward, c2r
backward), forward or backward, 1-D or 2-D. Data lives in device memory for the whole run,
as the adapter keeps it in cache (``doit`` loops the transform ``iter`` times).

Speed (``-s``): planning is timed once (``setup``); then, as libbench2 does
(docs/BENCHMARK_PROTOCOL.md:15-19), the iteration count doubles until one measurement takes
at least ``--time-min`` seconds (10 ms), eight measurements are taken and the minimum per
transform is reported, timed with HIP events on the stream the transforms run on. The
``mflops`` figure is libbench2's convention: 5 N log2 N / t for complex transforms,
2.5 N log2 N / t for real ones.

Verification (``--verify``): the three properties libbench2's verifier tests -- linearity
F(a x + b y) = a F(x) + b F(y), the impulse response (a unit impulse transforms to all
ones; for c2r, all-ones half spectra transform to N times the impulse), and the time-shift
theorem F(x shifted by one) = F(x) w^k along each axis -- each reported as the largest
relative deviation max|lhs - rhs| / max|rhs| over random inputs. They are this program's
own statements of those tests, not libbench2's code, so the values are comparable in
magnitude, not digit for digit.
"""
import argparse
import importlib.util
import math
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))

_PROBLEM = re.compile(r"^([io])([cr])([fb])(\d+)(?:x(\d+))?$")


def parse_problem(s: str) -> dict:
    """Decode a libbench2 problem string of the supported subset (raises ValueError)."""
    m = _PROBLEM.match(s)
    if not m:
        raise ValueError(f"unsupported problem string: {s!r}")
    inplace, kind, direction, n0, n1 = m.groups()
    dims = [int(n0)] + ([int(n1)] if n1 else [])
    if kind == "r" and len(dims) != 1:
        raise ValueError("real problems are one-dimensional here")
    if kind == "r" and inplace == "i":
        raise ValueError("real transforms run out of place")
    if any(d < 1 for d in dims):
        raise ValueError("sizes must be positive")
    return {"name": s, "inplace": inplace == "i", "real": kind == "r",
            "forward": direction == "f", "dims": dims}


def mflops(prob: dict, seconds: float) -> float:
    n = math.prod(prob["dims"])
    flops = (2.5 if prob["real"] else 5.0) * n * math.log2(n) if n > 1 else 0.0
    return flops / seconds / 1e6


def fmt_time(seconds: float) -> str:
    for unit, scale in (("ns", 1e-9), ("us", 1e-6), ("ms", 1e-3)):
        if seconds < 1000 * scale:
            return f"{seconds / scale:.2f} {unit}"
    return f"{seconds:.2f} s"


def load_library():
    spec = importlib.util.spec_from_file_location(
        "lean_fftw_amd", os.path.join(HERE, "__init__.py"), submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules.setdefault("lean_fftw_amd", mod)
    spec.loader.exec_module(mod)
    return mod


class Problem:
    """A planned problem with its device buffers (the adapter's setup / doit / done)."""

    def __init__(self, lf, torch, prob: dict):
        self.lf, self.torch, self.prob = lf, torch, prob
        dev = "cuda:0"
        sign = lf.FORWARD if prob["forward"] else lf.BACKWARD
        dims = prob["dims"]
        n = dims[0]
        if prob["real"]:
            self.plan = lf.RealPlan(n, sign, 1)
            nh = n // 2 + 1
            if prob["forward"]:
                self.x = torch.zeros(1, n, dtype=torch.float64, device=dev)
                self.y = torch.zeros(1, nh, dtype=torch.complex128, device=dev)
            else:
                self.x = torch.zeros(1, nh, dtype=torch.complex128, device=dev)
                self.y = torch.zeros(1, n, dtype=torch.float64, device=dev)
        else:
            if len(dims) == 2:
                self.plan = lf.Plan2D(dims[0], dims[1], sign, 1)
            else:
                self.plan = lf.Plan(n, sign, 1)
            self.x = torch.zeros(*dims, dtype=torch.complex128, device=dev)
            self.y = self.x if prob["inplace"] else torch.zeros_like(self.x)

    def doit(self, iters: int):
        for _ in range(iters):
            self.plan.call(self.x, self.y)

    def apply(self, host):
        """Transform one host array (numpy) and return the result on the host."""
        torch = self.torch
        self.x.copy_(torch.from_numpy(host).reshape(self.x.shape))
        self.plan.call(self.x, self.y)
        torch.cuda.synchronize()
        return self.y.cpu().numpy().reshape(self.y.shape).copy()


def speed(lf, torch, prob: dict, time_min: float, reps: int = 8) -> str:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p = Problem(lf, torch, prob)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    p.x.zero_()  # libbench2 zeroes the input after setup
    p.doit(1)
    torch.cuda.synchronize()

    def measure(iters):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        p.doit(iters)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / 1e3

    iters = 1
    while measure(iters) < time_min and iters < (1 << 30):
        iters *= 2
    t = min(measure(iters) for _ in range(reps)) / iters
    return (f"Problem: {prob['name']}, setup: {fmt_time(setup)}, time: {fmt_time(t)}, "
            f"``mflops'': {mflops(prob, t):.3f}")


def _rel(lhs, rhs):
    import numpy as np

    scale = float(np.max(np.abs(rhs)))
    return float(np.max(np.abs(lhs - rhs))) / (scale if scale > 0 else 1.0)


def verify(lf, torch, prob: dict, rounds: int = 3, seed: int = 1) -> tuple:
    """Linearity, impulse and time-shift errors (each the max over `rounds` draws)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    p = Problem(lf, torch, prob)
    dims = prob["dims"]
    s = -1.0 if prob["forward"] else 1.0
    real_in = prob["real"] and prob["forward"]
    half_in = prob["real"] and not prob["forward"]
    n = dims[0]
    nh = n // 2 + 1

    def rand_input():
        if real_in:
            return rng.random(n) - 0.5
        if half_in:
            h = (rng.random(nh) - 0.5) + 1j * (rng.random(nh) - 0.5)
            h[0] = h[0].real  # a Hermitian half spectrum: real DC (and Nyquist) bins
            if n % 2 == 0:
                h[-1] = h[-1].real
            return h
        shape = tuple(dims)
        return (rng.random(shape) - 0.5) + 1j * (rng.random(shape) - 0.5)

    e_lin = e_imp = e_shift = 0.0
    for _ in range(rounds):
        x, y = rand_input(), rand_input()
        if prob["real"]:
            a, b = rng.random() - 0.5, rng.random() - 0.5
        else:
            a = complex(rng.random() - 0.5, rng.random() - 0.5)
            b = complex(rng.random() - 0.5, rng.random() - 0.5)
        fx, fy = p.apply(x), p.apply(y)
        e_lin = max(e_lin, _rel(p.apply(a * x + b * y), a * fx + b * fy))

        if half_in:
            # shift on the output side: c2r(h_k w^{-k}) is the real signal rolled by one
            w = np.exp(-2j * np.pi * np.arange(nh) / n)
            e_shift = max(e_shift, _rel(p.apply(x * w), np.roll(fx, 1, axis=-1)))
        else:
            for ax in range(len(dims)):
                m = dims[ax]
                k = np.arange(fx.shape[-len(dims) + ax] if not prob["real"] else nh)
                shape = [1] * fx.ndim
                shape[fx.ndim - len(dims) + ax] = k.size
                w = np.exp(s * 2j * np.pi * k / m).reshape(shape)
                rolled = np.roll(x, 1, axis=ax - len(dims))
                e_shift = max(e_shift, _rel(p.apply(rolled), fx * w))

    if half_in:
        ones = np.ones(nh, dtype=np.complex128)
        imp = np.zeros(n)
        imp[0] = n
        e_imp = _rel(p.apply(ones), imp[None, :])
    else:
        d = np.zeros(n if real_in else tuple(dims), dtype=np.float64 if real_in else np.complex128)
        d.flat[0] = 1.0
        out = p.apply(d)
        e_imp = _rel(out, np.ones_like(out))
    return e_lin, e_imp, e_shift


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("-s", "--speed", metavar="PROBLEM")
    ap.add_argument("-y", "--verify", metavar="PROBLEM")
    ap.add_argument("-v", "--verbose", nargs="?", const="1", default="0",
                    help="libbench2's verbosity flag (-v1); accepted for compatibility")
    ap.add_argument("--time-min", type=float, default=0.010,
                    help="minimum seconds per timed measurement (libbench2's time_min)")
    a = ap.parse_args(argv)
    if not a.speed and not a.verify:
        ap.error("give -s PROBLEM or --verify PROBLEM")
    try:
        prob = parse_problem(a.speed or a.verify)
    except ValueError as e:
        print(f"bench2: {e}", file=sys.stderr)
        return 2
    import torch

    lf = load_library()
    try:
        if a.speed:
            print(speed(lf, torch, prob, a.time_min), flush=True)
        else:
            e = verify(lf, torch, prob)
            print(f"{prob['name']} {e[0]:g} {e[1]:g} {e[2]:g}", flush=True)
    except lf.MfftError as e:
        print(f"bench2: {prob['name']}: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
