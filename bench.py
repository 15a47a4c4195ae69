#!/usr/bin/env python3
"""Benchmark of the FP64 complex FFT hot path on one MI355X.

Default workload (the metric's configuration): the batched power-of-two sweep -- complex
FP64 forward, out of place, N = 2^k for k = 1..18, batch 2^24/N per size (256 MB in and
256 MB out per size), uniform[-0.5, 0.5) re/im, inputs resident in HBM. One step is one
pass over all 18 sizes (18 batched transforms). value = algorithmic bytes (32 N per
transform: 16N read + 16N written) of the whole step divided by the step time, GB/s.

    python bench.py [--steps K] [--warmup W] [--config sweep|single|mixed|ragged|autograd]

Prints one JSON line. `roofline` is measured live with HIP events on the stream the
kernels run on; `cpu_baseline` times the oracle (the C restatement of the reference's
refTransform, one pinned core) with libbench2's timing rule, beside the reference's published
single-core times. The default line also carries the metric's other two parts: the ragged
10,000-transform batch (`ragged`) and the differentiable transform's forward + backward over
forward (`autograd`).
"""

from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "FP64 c2c FFT GB/s & % HBM roofline; ragged-batch & fwd+bwd vs plain GPU"
HBM_PEAK_GBS = 8000.0


def load_package():
    if "lean_fftw_amd" in sys.modules:
        return sys.modules["lean_fftw_amd"]
    pkg = os.path.join(ROOT, "lean-fftw_amd")
    spec = importlib.util.spec_from_file_location(
        "lean_fftw_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg]
    )
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lean_fftw_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def uniform_complex(torch, shape, seed):
    g = torch.Generator(device="cuda:0").manual_seed(seed)
    re = torch.rand(shape, dtype=torch.float64, device="cuda:0", generator=g) - 0.5
    im = torch.rand(shape, dtype=torch.float64, device="cuda:0", generator=g) - 0.5
    return torch.complex(re, im)


def time_steps(torch, step, steps, warmup):
    """W untimed steps, then K timed steps bracketed by synchronisation; HIP events on the
    current stream. Returns seconds per step."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / steps


# ------------------------------------------------------------------------------ CPU baseline

# The reference's published single-core times (ns per transform, one transform, out of place,
# forward): its assembly engine i12 on Emerald Rapids at 2.1 GHz, one core pinned, median of
# 10 process runs of libbench2's min-of-8 (the reference's docs/RESULTS.md:70-89,
# results/emr-i12/summary.txt).
REF_PUBLISHED_NS = {2: 0.99, 4: 1.605, 8: 3.01, 16: 7.245, 32: 17.17, 64: 40.4, 128: 96.16,
                    256: 217.1, 512: 420.9, 1024: 1075.0, 2048: 2405.0, 4096: 6200.0,
                    8192: 12920.0, 16384: 29490.0, 32768: 70270.0, 65536: 208300.0,
                    131072: 647700.0, 262144: 1355000.0}


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def libbench2_time(run, min_seconds=0.01, samples=8):
    """libbench2's timing rule: double the repetitions until one timing takes at least
    10 ms, then the minimum of 8 timings; seconds per repetition."""
    reps = 1
    while True:
        t0 = time.perf_counter()
        run(reps)
        if time.perf_counter() - t0 >= min_seconds:
            break
        reps *= 2
    best = float("inf")
    for _ in range(samples):
        t0 = time.perf_counter()
        run(reps)
        best = min(best, time.perf_counter() - t0)
    return best / reps


def cpu_baseline_sweep(sizes, plans_radices, budget_s):
    """The oracle -- the C restatement of the reference's refTransform with the GPU plans'
    radices -- on one pinned core, timed per size with libbench2's rule on a call repeated in
    C (as libbench2 drives the reference) and scaled to the sweep's batch of 2^24 / N
    transforms per size. Each call transforms four inputs at once, one per lane of a
    four-double AVX2 vector (the scalar oracle's operations lane by lane, so its bits; 1.0 to
    3.9 times the scalar path's rate). The reference's own published per-size times (another
    host) are given beside it, aggregated the same way."""
    sys.path.insert(0, ROOT)
    import numpy as np

    import oracle

    oracle.build()
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    per, total_b, total_t, ref_t = [], 0.0, 0.0, 0.0
    t_start = time.perf_counter()
    try:
        for N in sizes:
            rp = oracle.RefPlan(True, N, plans_radices[N])
            x = np.ascontiguousarray(oracle.lcg_input(N, seed=oracle.LCG_SEED ^ N, howmany=4))
            out = np.empty_like(x)
            t = libbench2_time(lambda reps: rp.repeat_x4(x, out, reps)) / 4
            batch = (1 << 24) // N
            total_t += t * batch
            total_b += 32.0 * N * batch
            ref_t += REF_PUBLISHED_NS[N] * 1e-9 * batch
            per.append({"N": N, "ns": round(t * 1e9, 2), "ref_published_ns": REF_PUBLISHED_NS[N]})
            if time.perf_counter() - t_start > budget_s:
                break
    finally:
        os.sched_setaffinity(0, old)
    done = len(per)
    return {
        "value": round(total_b / total_t / 1e9, 3),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle refTransform (C: gcc {' '.join(oracle.CFLAGS)}) with the GPU plans' "
                  f"radices, four transforms per call (one per lane of an AVX2 vector; the "
                  f"scalar oracle's bits), per size for {done} of the {len(sizes)} sweep sizes, "
                  f"pinned to core {core}, libbench2's rule (repetitions doubled to >= 10 ms, "
                  f"min of 8), scaled to batch 2^24/N; host {host_cpu_model()}, "
                  f"{os.cpu_count()} logical CPUs",
        "reference_published": {
            "value": round(total_b / ref_t / 1e9, 3), "unit": "GB/s",
            "host": "reference assembly engine i12, Emerald Rapids 2.1 GHz, 1 core (published)",
            "aggregated": "its per-size times of one transform scaled to batch 2^24/N, as above"},
        "per_size": per,
    }


# ------------------------------------------------------------------------------ workloads


def bench_sweep(args, torch, lf):
    sizes = [1 << k for k in range(1, 19)]
    if args.sizes:
        sizes = [n for n in sizes if n in args.sizes]
    total = 1 << 24
    x = uniform_complex(torch, (total,), 20261203)
    y = torch.empty_like(x)
    flags = lf.MEASURE if args.measure else lf.ESTIMATE
    plans = {N: lf.Plan(N, lf.FORWARD, total // N, flags=flags) for N in sizes}
    views = {N: (x.view(total // N, N), y.view(total // N, N)) for N in sizes}

    def step():
        for N in sizes:
            xi, yi = views[N]
            plans[N].call(xi, yi)

    t = time_steps(torch, step, args.steps, args.warmup)
    # per-size timing (events around back-to-back batched transforms of one size)
    per = []
    for N in sizes:
        xi, yi = views[N]
        for _ in range(max(1, args.warmup)):
            plans[N].call(xi, yi)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(3, args.steps)
        e0.record()
        for _ in range(reps):
            plans[N].call(xi, yi)
        e1.record()
        torch.cuda.synchronize()
        ts = e0.elapsed_time(e1) / 1e3 / reps
        gbs = 32.0 * total / ts / 1e9
        per.append({"N": N, "radices": plans[N].radices, "stages": plans[N].stages,
                    "us": round(ts * 1e6, 2), "GB/s": round(gbs, 1),
                    "frac": round(gbs / HBM_PEAK_GBS, 4)})
    bytes_step = 32.0 * total * len(sizes)
    value = bytes_step / t / 1e9
    # HBM traffic per step from the committed rocprofv3 PMC summary (FETCH_SIZE, WRITE_SIZE,
    # each scaled by a plain copy kernel), used only when its plans are this run's plans
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_pow2_sweep.json")
    if os.path.exists(tpath) and len(sizes) == 18:
        with open(tpath) as f:
            prof = json.load(f)
        same = all(prof["per_size"].get(str(q["N"]), {}).get("radices") == q["radices"]
                   for q in per)
        if same:
            traffic = round(prof["step_traffic_bytes"])
            traffic_src = ("stored rocprofv3 PMC summary of this sweep's plans "
                           "(profiles/traffic_pow2_sweep.json), not measured in this run")
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched power-of-two sweep: complex FP64 forward, out of place, "
                               "N=2^1..2^18, batch 2^24/N per size (256 MB in + 256 MB out), "
                               "uniform[-0.5,0.5) re/im; one step = all 18 sizes",
                   "sizes": len(sizes), "elements_per_size": total,
                   "planner": "measure" if args.measure else "estimate"},
        "roofline": {"bound": "hbm", "achieved": round(value, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(value / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "every FFT kernel of one sweep step: algorithmic 32N bytes per "
                               "transform over the timed step (the value's own clock)"},
        "per_size": per,
        "per_size_sum_us": round(sum(q["us"] for q in per), 2),
    }
    if len(sizes) == 18:
        res["dominant_kernel"] = dominant_kernel(torch, plans[65536], views[65536], args)
    if not args.no_extras and len(sizes) == 18:
        res["ragged"] = ragged_summary(args, torch, lf, value)
        res["autograd"] = autograd_summary(args, torch, lf)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_sweep(sizes, {N: plans[N].radices for N in sizes},
                                                 args.cpu_seconds)
    return res


def dominant_kernel(torch, plan, view, args):
    """The kernel with the largest share of a sweep step: the register stage kernel of 256
    points at 16 sub-problems per workgroup (both stages of 2^16, the second of 2^15, the
    first of 2^17). Timed live through the 2^16 plan, whose call is two launches of it; each
    launch carries half of the batch's algorithmic bytes (32 N per transform)."""
    xi, yi = view
    for _ in range(max(1, args.warmup)):
        plan.call(xi, yi)
    reps = max(5, args.steps)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.call(xi, yi)
    e1.record()
    torch.cuda.synchronize()
    launch = e0.elapsed_time(e1) / 1e3 / reps / 2
    alg = 32.0 * (1 << 24) / 2
    return {"kernel": "mfft::stage_reg_kernel<true, 256, 8, 16, 8, 8, 4>", "bound": "hbm",
            "launch_us": round(launch * 1e6, 2), "algorithmic_bytes_per_launch": alg,
            "achieved": round(alg / launch / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / launch / 1e9 / HBM_PEAK_GBS, 4),
            "note": "a stage of a two-stage transform: each launch reads and writes the whole "
                    "batch once, so the HBM bytes per launch are twice the algorithmic share"}


def ragged_summary(args, torch, lf, sweep_gbs):
    r = bench_ragged(args, torch, lf)
    return {"GB/s": r["value"], "frac": r["roofline"]["frac"], "ms": r["ms_per_step"],
            "launches": 2, "transforms": 10000, "ragged_vs_uniform": r["ragged_vs_uniform"],
            "uniform_same_sizes_GB/s": r["uniform_same_sizes"]["GB/s"],
            "ragged_vs_sweep": round(r["value"] / sweep_gbs, 4),
            "ragged_vs_per_size_launches": r["ragged_vs_per_size"]}


def autograd_summary(args, torch, lf):
    r = bench_autograd(args, torch, lf)
    c = r["config"]
    return {"fwd_bwd_over_fwd": r["value"], "graph_fwd_ms": c["graph_fwd_ms"],
            "graph_fwd_bwd_ms": c["graph_fwd_bwd_ms"], "eager_ratio": c["eager_ratio"],
            "shape": [512, 4096]}


def bench_single(args, torch, lf):
    """Config 1: one N=1024 transform, out of place (latency-bound on a GPU). Timed two ways:
    `reps` eager calls from Python (each a ctypes call and a kernel launch), and the same
    `reps` launches replayed from one HIP graph (what a launch-bound caller would capture);
    the value is the graph replay's time per transform."""
    N = 1024
    x = uniform_complex(torch, (N,), 1)
    y = torch.empty_like(x)
    p = lf.Plan(N, lf.FORWARD, 1)
    reps = 200

    def step():
        for _ in range(reps):
            p.call(x, y)

    t_eager = time_steps(torch, step, args.steps, args.warmup) / reps
    g = capture_graph(torch, step)
    t = time_steps(torch, g.replay, args.steps, args.warmup) / reps
    return {"metric": METRIC, "value": round(32.0 * N / t / 1e9, 4), "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 6),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "single transform N=1024, forward, out of place; "
                                   f"{reps} transforms replayed from one HIP graph per step",
                       "us_per_transform": round(t * 1e6, 3),
                       "us_per_transform_eager_calls": round(t_eager * 1e6, 3),
                       "radices": p.radices}}


def bench_mixed(args, torch, lf):
    """Config 3: batched 13-smooth sizes, each moving >= 256 MB."""
    sizes = [6, 60, 360, 1000, 2100, 10368, 75600, 165375]
    per, tot_b, tot_t = [], 0.0, 0.0
    for N in sizes:
        hm = -(-(1 << 24) // N)
        x = uniform_complex(torch, (hm, N), N)
        y = torch.empty_like(x)
        p = lf.Plan(N, lf.FORWARD, hm, flags=lf.MEASURE if args.measure else lf.ESTIMATE)
        t = time_steps(torch, lambda: p.call(x, y), args.steps, args.warmup)
        b = 32.0 * N * hm
        per.append({"N": N, "batch": hm, "radices": p.radices, "stages": p.stages,
                    "us": round(t * 1e6, 2), "GB/s": round(b / t / 1e9, 1),
                    "frac": round(b / t / 1e9 / HBM_PEAK_GBS, 4)})
        tot_b += b
        tot_t += t
        del x, y
    v = tot_b / tot_t / 1e9
    return {"metric": METRIC, "value": round(v, 2), "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(tot_t * 1e3, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "batched 13-smooth sizes {6,60,360,1000,2100,10368,75600,165375}, "
                                   "batch ceil(2^24/N)"},
            "roofline": {"bound": "hbm", "achieved": round(v, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(v / HBM_PEAK_GBS, 4), "traffic": None},
            "per_size": per}


def bench_ragged(args, torch, lf):
    """Config 4: 10,000 transforms of log-uniform 13-smooth sizes in [8, 65536], packed back
    to back, run as two persistent launches from size-sorted queues (stage A and single-stage
    tiles, then stage B), captured in one HIP graph. Compared with per-size plain launches
    (the same transforms regrouped by size, one batched uniform launch per distinct size) and
    with each distinct size's own uniform throughput."""
    sizes = lf.workloads.ragged_sizes()
    total = sum(sizes)
    x = uniform_complex(torch, (total,), 44)
    y = torch.empty_like(x)
    t0 = time.perf_counter()
    rp = lf.RaggedPlan(sizes, lf.FORWARD)
    plan_s = time.perf_counter() - t0
    rp.call(x, y)
    rp.status()
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        rp.call(x, y)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=s):
            rp.call(x, y)
    torch.cuda.current_stream().wait_stream(s)
    t_rag = time_steps(torch, graph.replay, args.steps, args.warmup)
    rp.status()
    b = 32.0 * total
    # per-size plain launches on the regrouped layout
    from collections import Counter

    cnt = Counter(sizes)
    plans = {n: lf.Plan(n, lf.FORWARD, c) for n, c in cnt.items()}
    offs, off = {}, 0
    for n, c in sorted(cnt.items()):
        offs[n] = off
        off += n * c

    def per_size():
        for n, c in cnt.items():
            o = offs[n]
            plans[n].call(x[o:o + n * c], y[o:o + n * c])

    t_ps = time_steps(torch, per_size, args.steps, args.warmup)
    # uniform-size throughput of the same size mix: every distinct size as a plain batched
    # launch of >= 64 MB (its own throughput), combined by bytes (time-weighted)
    t_uni = 0.0
    cap = 1 << 22  # elements per uniform batch (64 MB)
    for n, c in cnt.items():
        hm = max(1, cap // n)
        p = lf.Plan(n, lf.FORWARD, hm)
        xi = x[:n * hm].view(hm, n)
        yi = y[:n * hm].view(hm, n)
        tn = time_steps(torch, lambda: p.call(xi, yi), 2, 1)
        t_uni += tn * (c / hm)
        del p
    uni = b / t_uni / 1e9
    return {"metric": METRIC, "value": round(b / t_rag / 1e9, 2), "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_rag * 1e3, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "ragged batch: 10,000 complex FP64 forward transforms, sizes "
                                   "log-uniform over the 13-smooth integers in [8, 65536] (PCG64 "
                                   "seed 20261203), packed back to back, two persistent launches "
                                   "in one HIP graph", "sum_N": total, "distinct_sizes": len(cnt),
                       "plan_seconds": round(plan_s, 2)},
            "per_size_plain": {"GB/s": round(b / t_ps / 1e9, 2), "ms": round(t_ps * 1e3, 4),
                               "launches": len(cnt)},
            "ragged_vs_per_size": round(t_ps / t_rag, 4),
            "uniform_same_sizes": {"GB/s": round(uni, 2),
                                   "note": "each distinct size as a uniform batch of >= 64 MB, "
                                           "combined by bytes"},
            "ragged_vs_uniform": round(b / t_rag / 1e9 / uni, 4),
            "roofline": {"bound": "hbm", "achieved": round(b / t_rag / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(b / t_rag / 1e9 / HBM_PEAK_GBS, 4),
                         "traffic": None}}


def capture_graph(torch, fn, warm=3):
    """Capture fn() in a HIP graph (after warm-up runs on a side stream)."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(warm):
            fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    torch.cuda.synchronize()
    return g


def bench_autograd(args, torch, lf):
    """Config 5: loss = sum(|w * FFT(x)|^2), x (512, 4096) complex128; forward alone vs
    forward + backward (the hand-written adjoint) through the autograd Function. Both steps
    are timed as captured HIP graphs (GPU cost) and eagerly (including Python and
    autograd-engine launch overhead)."""
    B, N = 512, 4096
    x = uniform_complex(torch, (B, N), 55).requires_grad_(True)
    g = torch.Generator(device="cuda:0").manual_seed(56)
    w = torch.rand(N, dtype=torch.float64, device="cuda:0", generator=g)

    def fwd():
        with torch.no_grad():
            lf.weighted_power_loss(x, w)

    def fwd_bwd():
        x.grad = None
        loss = lf.weighted_power_loss(x, w)
        loss.backward()

    steps = args.steps * 4
    t_f = time_steps(torch, fwd, steps, args.warmup)
    t_fb = time_steps(torch, fwd_bwd, steps, args.warmup)
    # GPU cost: the same two steps, autograd engine included, captured in HIP graphs; K steps
    # per graph, so that a replay (about 30 us of GPU work per step) is not bound by the
    # host's graph launch. No step keeps its graph alive into the next one (the capture must
    # not inherit an eager step's gradient accumulator, which belongs to the default stream).
    K = 16

    def k_fwd():
        for _ in range(K):
            fwd()

    def k_fwd_bwd():
        for _ in range(K):
            fwd_bwd()

    gf = capture_graph(torch, k_fwd)
    gfb = capture_graph(torch, k_fwd_bwd)
    tg_f = time_steps(torch, gf.replay, steps, args.warmup) / K
    tg_fb = time_steps(torch, gfb.replay, steps, args.warmup) / K
    b = 32.0 * B * N
    return {"metric": METRIC, "value": round(tg_fb / tg_f, 4),
            "unit": "x (fwd+bwd time / fwd time, HIP graphs)",
            "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": round(tg_fb * 1e3, 4), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "differentiable FFT: loss = sum(|w*FFT(x)|^2), x (512, 4096) "
                                   "complex128, w ~ U[0,1)",
                       "graph_fwd_ms": round(tg_f * 1e3, 4), "graph_fwd_bwd_ms": round(tg_fb * 1e3, 4),
                       "eager_fwd_ms": round(t_f * 1e3, 4), "eager_fwd_bwd_ms": round(t_fb * 1e3, 4),
                       "eager_ratio": round(t_fb / t_f, 4),
                       "fwd_fft_GB/s": round(b / tg_f / 1e9, 1)}}


def bench_real(args, torch, lf):
    """Real-data forward transforms (r2c) of 2^4 .. 2^18 points, batch 2^24/N: the
    reference's rdft2 problems, computed through the complex path. Bytes are the ideal
    8N in + 16(N/2+1) out per transform."""
    sizes = args.sizes or [2 ** k for k in range(4, 19)]
    per, tot_b, tot_t = [], 0.0, 0.0
    for N in sizes:
        hm = (1 << 24) // N
        g = torch.Generator("cuda:0").manual_seed(N)
        x = torch.rand((hm, N), dtype=torch.float64, device="cuda:0", generator=g) - 0.5
        y = torch.empty((hm, N // 2 + 1), dtype=torch.complex128, device="cuda:0")
        p = lf.RealPlan(N, lf.FORWARD, hm)
        t = time_steps(torch, lambda: p.call(x, y), args.steps, args.warmup)
        b = (8.0 * N + 16.0 * (N // 2 + 1)) * hm
        per.append({"N": N, "batch": hm, "radices": p.radices, "us": round(t * 1e6, 2),
                    "GB/s": round(b / t / 1e9, 1)})
        tot_b += b
        tot_t += t
        del x, y
    v = tot_b / tot_t / 1e9
    return {"metric": "real-data forward transforms (r2c) GB/s", "value": round(v, 2),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(tot_t * 1e3, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": ("r2c FP64, N = 2^4..2^18" if not args.sizes else
                                    "r2c FP64, N in " + ", ".join(str(n) for n in sizes))
                                   + ", batch 2^24/N, out of place"},
            "per_size": per}


def bench_2d(args, torch, lf):
    """2-D complex forward transforms ocfNxN (N = 64 .. 4096), 2^24 elements per step."""
    sizes = args.sizes or [64, 128, 256, 512, 1024, 2048, 4096]
    per, tot_b, tot_t = [], 0.0, 0.0
    for N in sizes:
        hm = max(1, (1 << 24) // (N * N))
        x = uniform_complex(torch, (hm, N, N), N)
        y = torch.empty_like(x)
        p = lf.Plan2D(N, N, lf.FORWARD, hm)
        t = time_steps(torch, lambda: p.call(x, y), args.steps, args.warmup)
        b = 32.0 * N * N * hm
        col = lf.Plan(N, lf.FORWARD, N, istride=N, idist=1, ostride=N, odist=1)  # axis-0 part
        per.append({"N": N, "batch": hm, "us": round(t * 1e6, 2), "GB/s": round(b / t / 1e9, 1),
                    "axis0_radices": col.radices, "axis0_stages": col.stages,
                    "axis1_radices": p.radices})
        del col
        tot_b += b
        tot_t += t
        del x, y
    v = tot_b / tot_t / 1e9
    return {"metric": "2-D complex transforms GB/s", "value": round(v, 2), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(tot_t * 1e3, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "ocfNxN: 2-D complex FP64 forward, N = 64..4096, "
                                   "2^24 elements per size, out of place"},
            "per_size": per}


def bench_copy(args, torch, lf):
    """A plain device copy of 2^24 complex128 (256 MB read + 256 MB written): the practical
    HBM ceiling, and the reference launch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    total = 1 << 24
    x = uniform_complex(torch, (total,), 3)
    y = torch.empty_like(x)
    t = time_steps(torch, lambda: y.copy_(x), args.steps * 4, args.warmup)
    b = 32.0 * total
    return {"metric": "device copy GB/s", "value": round(b / t / 1e9, 1), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps * 4, "warmup": args.warmup,
            "ms_per_step": round(t * 1e3, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "torch copy_ of 2^24 complex128 (256 MB in, 256 MB out)"},
            "frac_of_peak": round(b / t / 1e9 / HBM_PEAK_GBS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="sweep",
                    choices=["sweep", "single", "mixed", "ragged", "autograd", "copy", "real",
                             "2d"])
    ap.add_argument("--sizes", type=int, nargs="*", help="restrict the sweep to these sizes")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="sweep only: leave out the ragged and autograd parts of the metric")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--measure", action="store_true",
                    help="plan with MFFT_MEASURE (time candidate schedules before the run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (a batch shards "
              "with no communication; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch

    lf = load_package()
    fn = {"sweep": bench_sweep, "single": bench_single, "mixed": bench_mixed,
          "ragged": bench_ragged, "autograd": bench_autograd, "copy": bench_copy,
          "real": bench_real, "2d": bench_2d}[args.config]
    res = fn(args, torch, lf)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
