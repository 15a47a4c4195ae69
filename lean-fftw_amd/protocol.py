"""The benchmark protocol's statistics: per-size medians over process runs, the geometric
mean G of per-size time ratios against a reference implementation, and its bootstrap 95%
confidence interval (the reference's tools/bench.py:112-156; docs/BENCHMARK_PROTOCOL.md).

Input: the protocol's results rows, {"impl", "n", "run", "time", ...} (one libbench2 `-s`
measurement per process run; rows whose time is null are skipped), e.g. the results.json of a
protocol run, or rows made from bench2.py's output for the GPU implementation.

    python protocol.py analyze RESULTS.json --ref lean-i12 \
        --combine fftw=fftw-measure,fftw-patient,fftw-exhaustive

prints the per-size medians and one line per compared implementation:
"REF / IMPL: G = ...  95% CI [..., ...]". G = exp(mean over sizes of log(med_ref / med_impl))
(below 1: REF is faster); a combined implementation's median per size is the minimum of its
members' medians (e.g. FFTW's best planner effort). The interval resamples each
(implementation, size)'s runs with replacement B = 10,000 times (numpy's PCG64, seed 1),
takes medians (a combined implementation: the minimum of its members' resampled medians),
and reads the 2.5th and 97.5th percentiles of the resampled G. The resampling draws come in
the reference's order -- the reference's resampled medians of every size first, then each
compared implementation's, sizes ascending, a combined one's members in the order given -- so
that the published intervals are reproduced to the last digit.
"""

from __future__ import annotations

import argparse
import json
import math
import sys


def medians(rows):
    """{impl: {n: [times]}} and {impl: {n: median}}."""
    import numpy as np

    runs: dict = {}
    for r in rows:
        if r.get("time") is None:
            continue
        runs.setdefault(r["impl"], {}).setdefault(int(r["n"]), []).append(float(r["time"]))
    med = {i: {n: float(np.median(v)) for n, v in by.items()} for i, by in runs.items()}
    return runs, med


def analyze(rows, ref=None, combine=None, B=10000, seed=1):
    """Medians, and (with `ref`) G and its bootstrap 95% interval for every other
    implementation measured at every size. `combine` = {name: [member impls]}."""
    import numpy as np

    runs, med = medians(rows)
    sizes = sorted({n for by in runs.values() for n in by})
    combine = {c: [m for m in ms if m in runs] for c, ms in (combine or {}).items()}
    for c, ms in combine.items():
        if ms:
            med[c] = {n: min(med[m][n] for m in ms if n in med[m]) for n in sizes}
    out = {"sizes": sizes, "medians": med}
    if not ref:
        return out
    rng = np.random.default_rng(seed)

    def resampled(impl, n):
        if impl in combine:
            return np.min([resampled(m, n) for m in combine[impl] if n in runs[m]], axis=0)
        x = np.asarray(runs[impl][n])
        pick = rng.integers(0, len(x), size=(B, len(x)))
        return np.median(x[pick], axis=1)

    ref_boot = {n: resampled(ref, n) for n in sizes}
    vs = {}
    for c in sorted(med):
        if c == ref or not all(n in med[c] for n in sizes):
            continue
        if c not in combine and not all(n in runs[c] for n in sizes):
            continue
        ratio = {n: med[ref][n] / med[c][n] for n in sizes}
        g = math.exp(sum(math.log(v) for v in ratio.values()) / len(sizes))
        acc = np.zeros(B)
        for n in sizes:
            acc += np.log(ref_boot[n] / resampled(c, n))
        gs = np.exp(acc / len(sizes))
        vs[c] = {"G": g, "ci95": [float(np.percentile(gs, 2.5)), float(np.percentile(gs, 97.5))],
                 "ratios": ratio}
    out["vs"] = vs
    return out


def report(out, ref) -> str:
    names = list(out["medians"])
    lines = ["N".rjust(8) + "".join(i[:14].rjust(15) for i in names)]
    for n in out["sizes"]:
        lines.append(str(n).rjust(8) + "".join(
            f"{out['medians'][i].get(n, float('nan')):.4g}".rjust(15) for i in names))
    for c, v in out.get("vs", {}).items():
        lines.append(f"{ref} / {c}: G = {v['G']:.4f}  95% CI [{v['ci95'][0]:.4f}, {v['ci95'][1]:.4f}]")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sp = ap.add_subparsers(dest="cmd", required=True)
    z = sp.add_parser("analyze")
    z.add_argument("results", help="results.json: a list of result rows")
    z.add_argument("--ref")
    z.add_argument("--combine", action="append", default=[], help="NAME=impl1,impl2,...")
    z.add_argument("--json", help="also write the analysis here")
    a = ap.parse_args(argv)
    rows = json.load(open(a.results))
    comb = {}
    for c in a.combine:
        k, v = c.split("=", 1)
        comb[k] = v.split(",")
    out = analyze(rows, a.ref, comb)
    print(report(out, a.ref))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
