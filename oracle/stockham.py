"""Oracle: the Stockham schedule over any arithmetic (test infrastructure only).

stockhamWith (LeanFFTW/Gen/Stockham.lean:101-108) with passWith/bflyInputs (:77-98),
vectorised with numpy over the butterflies of a pass. The arithmetic is pluggable, as
the reference's `Arith` (LeanFFTW/Gen/Emit.lean:31-44); two instances are provided:

  * DepthArith -- depthArith of scripts/ErrorBound.lean:24-33 (rounding depth a and
    constant depth b, exact integers);
  * MagArith   -- magF of scripts/ErrorBound.lean:36-45 (magnitude bound on an all-ones
    input, binary64 with |constants| rounded up by 2^-50).

The reference published (a, b, magMax) for 36 plans (results/repro-2026-11-04/
bounds.jsonl); reproducing them checks the shape of every butterfly DAG of those plans,
including the odd radices and Good-Thomas kernels that no stored vector pins.
"""

from __future__ import annotations

import numpy as np

from . import symgen, trig


class DepthArith:
    def inp(self, n):
        return (np.zeros(n, np.int64), np.ones(n, np.int64))

    def cst(self, c, n):
        return (np.zeros(n, np.int64), np.ones(n, np.int64))

    def zero(self, n):
        return (np.zeros(n, np.int64), np.zeros(n, np.int64))

    def add(self, a, b):
        return (np.maximum(a[0], b[0]) + 1, np.maximum(a[1], b[1]))

    sub = add

    def mul(self, a, b):
        return (a[0] + b[0] + 1, a[1] + b[1])

    def neg(self, a):
        return a

    def fma(self, a, b, c):
        return (np.maximum(a[0] + b[0], c[0]) + 1, np.maximum(a[1] + b[1], c[1]))

    fms = fma
    fnma = fma


class MagArith:
    ELL = 2.0 ** -1022

    def inp(self, n):
        return np.ones(n)

    def cst(self, c, n):
        return np.full(n, abs(trig.cst_value(c)) * (1 + 2.0 ** -50) + self.ELL)

    def zero(self, n):
        return np.zeros(n)

    def add(self, a, b):
        return (a + b) + self.ELL

    sub = add

    def mul(self, a, b):
        return a * b + self.ELL

    def neg(self, a):
        return a

    def fma(self, a, b, c):
        return ((a * b) + c) + self.ELL

    fms = fma

    def fnma(self, a, b, c):
        return (c + (a * b)) + self.ELL


def _gather(v, idx):
    if isinstance(v, tuple):
        return tuple(x[idx] for x in v)
    return v[idx]


def _eval_kernel(A, ops, outs, n_in, slots, n):
    env = list(slots)
    for op in ops:
        k = op[0]
        if k == "zero":
            env.append(A.zero(n))
        elif k == "cst":
            env.append(A.cst(op[1], n))
        elif k == "neg":
            env.append(A.neg(env[op[1]]))
        elif k in ("add", "sub", "mul"):
            env.append(getattr(A, k)(env[op[1]], env[op[2]]))
        else:
            env.append(getattr(A, k)(env[op[1]], env[op[2]], env[op[3]]))
    return [env[s] for s in outs]


def _scatter(dst, idx, v):
    if isinstance(dst, tuple):
        for d, x in zip(dst, v):
            d[idx] = x
    else:
        dst[idx] = v


def _alloc_like(v, n):
    if isinstance(v, tuple):
        return tuple(np.empty(n, x.dtype) for x in v)
    return np.empty(n, v.dtype)


def stockham_with(A, fwd: bool, N: int, rs, z, tw_of):
    """z: 2N values (interleaved) in A's representation; tw_of(t): pass-t table values."""
    for t, R in enumerate(rs):
        ns = trig.pass_ns(rs, t)
        twd = t >= 1
        n_in = symgen.n_inputs(R, twd)
        ops, outs = symgen.cse_kernel(n_in, symgen.butterfly(fwd, R, twd))
        NR = N // R
        j = np.arange(NR)
        kk = j % ns
        b = j // ns
        slots = []
        for r in range(R):
            slots.append(_gather(z, 2 * (j + r * NR)))
            slots.append(_gather(z, 2 * (j + r * NR) + 1))
        if twd:
            tab = tw_of(t)
            for r in range(R - 1):
                slots.append(_gather(tab, 2 * r * ns + kk))
                slots.append(_gather(tab, (2 * r + 1) * ns + kk))
        ys = _eval_kernel(A, ops, outs, n_in, slots, NR)
        znew = _alloc_like(z, 2 * N)
        for r in range(R):
            c = b * ns * R + kk + r * ns
            _scatter(znew, 2 * c, ys[2 * r])
            _scatter(znew, 2 * c + 1, ys[2 * r + 1])
        z = znew
    return z


def error_bound_terms(N: int, rs) -> tuple[int, int, float]:
    """(a, b, magMax) of scripts/ErrorBound.lean:49-59 for the forward plan rs."""
    D = DepthArith()
    M = MagArith()

    def tw_size(t):
        return 2 * (rs[t] - 1) * trig.pass_ns(rs, t)

    d = stockham_with(D, True, N, rs, D.inp(2 * N), lambda t: D.inp(tw_size(t)))
    m = stockham_with(M, True, N, rs, M.inp(2 * N), lambda t: M.inp(tw_size(t)))
    return int(d[0].max()), int(d[1].max()), float(m.max())
