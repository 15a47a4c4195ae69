"""Synthetic workload definitions (sizes only; no data).

ragged_sizes(): the mixed-size batch of the ragged benchmark -- `count` transform sizes
drawn log-uniformly from the 13-smooth integers in [lo, hi]: u ~ U[ln lo, ln hi], then the
smooth integer nearest to exp(u) in log space, from numpy's PCG64 with the given seed.
"""

from __future__ import annotations

import bisect
import math


def smooth13(n: int) -> bool:
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


def smooth_sizes(lo: int, hi: int) -> list[int]:
    return [n for n in range(lo, hi + 1) if smooth13(n)]


def ragged_sizes(count: int = 10000, lo: int = 8, hi: int = 65536, seed: int = 20261203):
    import numpy as np

    cand = smooth_sizes(lo, hi)
    logs = [math.log(c) for c in cand]
    u = np.random.Generator(np.random.PCG64(seed)).uniform(math.log(lo), math.log(hi), count)
    out = []
    for v in u:
        i = bisect.bisect_left(logs, v)
        if i == 0:
            out.append(cand[0])
        elif i == len(cand):
            out.append(cand[-1])
        else:
            out.append(cand[i] if logs[i] - v < v - logs[i - 1] else cand[i - 1])
    return out
