"""Oracle: exact-integer twiddle factors (test infrastructure only).

This module is part of the CPU oracle. Only the tests, the smoke check and the
CPU-baseline leg of the benchmark may import it; the shipped library has its own
implementation of the same routine in C++ and never calls this one.

Restatement of the reference's twiddle routine, LeanFFTW/Core/Trig.lean:
cos and sin of 2*pi*k/n are summed as Taylor series in 160-bit fixed point after an
exact octant reduction, then rounded once to binary64 (nearest, ties to even).

Pinned by tests/test_oracle_trig.py against
  * results/repro-2026-11-04/kernel-constants.txt (237 exact bit patterns of
    cos/sin/tan kernel constants), stored as tests/golden/kernel_constants.json;
  * results/repro-2026-11-04/twiddles.jsonl (max |error| of whole base tables,
    3 significant digits), stored as tests/golden/twiddle_tau.json;
  * mpmath at 200 bits (agreement to <= 1 ulp, exact on the closed forms).
"""

from __future__ import annotations

import math
import struct
from functools import lru_cache

# LeanFFTW/Core/Trig.lean:25 (precBits), :28-33 (piDigits, piScale, piFixed)
PREC_BITS = 160
PI_DIGITS = 3141592653589793238462643383279502884197169399375105820974944
PI_SCALE = 10 ** 60
PI_FIXED = PI_DIGITS * (1 << PREC_BITS) // PI_SCALE
ONE = 1 << PREC_BITS


def _fmul(a: int, b: int) -> int:
    """Trig.lean:36 -- floor(a*b / 2^precBits) on naturals."""
    return (a * b) >> PREC_BITS


def angle_fixed(num: int, den: int) -> int:
    """Trig.lean:39 -- theta * 2^precBits for theta = 2*pi*num/den."""
    return 2 * PI_FIXED * num // den


def cos_sin_fixed(theta: int) -> tuple[int, int]:
    """Trig.lean:43-62 -- 39 Taylor terms past the first, positive and negative
    terms accumulated separately, each term truncated by fmul and an integer division."""
    t2 = _fmul(theta, theta)
    c_pos, c_neg = ONE, 0
    s_pos, s_neg = theta, 0
    tc, ts = ONE, theta
    for j in range(1, 40):
        tc = _fmul(tc, t2) // ((2 * j - 1) * (2 * j))
        ts = _fmul(ts, t2) // ((2 * j) * (2 * j + 1))
        if j % 2 == 1:
            c_neg += tc
            s_neg += ts
        else:
            c_pos += tc
            s_pos += ts
    # Lean's Nat subtraction truncates at zero; both differences are >= 0 on [0, pi/4]
    return max(c_pos - c_neg, 0), max(s_pos - s_neg, 0)


def bits_to_float(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b & 0xFFFFFFFFFFFFFFFF))[0]


def float_to_bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def fixed_to_float(v: int) -> float:
    """Trig.lean:67-86 -- round v * 2^-precBits to binary64, nearest, ties to even."""
    if v == 0:
        return 0.0
    lg = v.bit_length() - 1
    if lg <= 52:
        q = v << (52 - lg)
    else:
        sh = lg - 52
        q = v >> sh
        r = v - (q << sh)
        half = 1 << (sh - 1)
        if r > half or (r == half and q % 2 == 1):
            q += 1
    if q == 1 << 53:
        q, lg = 1 << 52, lg + 1
    e = lg - PREC_BITS + 1023
    if e <= 0:
        raise ValueError("twiddle below the normal range")
    return bits_to_float((e << 52) | (q - (1 << 52)))


def tan_2pi(num: int, den: int) -> float:
    """Trig.lean:90-92 -- fixed-point quotient s/c of the two series, rounded once."""
    c, s = cos_sin_fixed(angle_fixed(num, den))
    return fixed_to_float(s * ONE // c)


def neg_z(x: float) -> float:
    """Trig.lean:95 -- negation that never yields -0.0."""
    return 0.0 if x == 0.0 else -x


@lru_cache(maxsize=None)
def small_at(n: int, num: int) -> tuple[float, float]:
    """Trig.lean:98-100 -- (cos, sin)(2*pi*num/(8n))."""
    c, s = cos_sin_fixed(angle_fixed(num, 8 * n))
    return fixed_to_float(c), fixed_to_float(s)


def _reduce(small, k: int, n: int) -> tuple[float, float]:
    """Trig.lean:104-122 -- exact octant reduction and the symmetry map per octant."""
    if n == 0:
        return 1.0, 0.0
    k %= n
    oct_, rem = divmod(8 * k, n)
    c, s = small(rem) if oct_ % 2 == 0 else small(n - rem)
    return {
        0: (c, s),
        1: (s, c),
        2: (neg_z(s), c),
        3: (neg_z(c), s),
        4: (neg_z(c), neg_z(s)),
        5: (neg_z(s), neg_z(c)),
        6: (s, neg_z(c)),
        7: (c, neg_z(s)),
    }[oct_]


def cos_sin_2pi(k: int, n: int) -> tuple[float, float]:
    """Trig.lean:126 -- (cos(2*pi*k/n), sin(2*pi*k/n)) as binary64."""
    if n == 0:
        return 1.0, 0.0
    return _reduce(lambda num: small_at(n, num), k, n)


def cos_sin_2pi_table(n: int) -> list[tuple[float, float]]:
    """Trig.lean:131-134 -- m -> cosSin2Pi m n for m < n, evaluating only the
    n/g + 1 reduced angles that occur (g = gcd(8, n))."""
    g = math.gcd(8, n)
    memo = []
    for i in range(n // g + 1):
        c, s = cos_sin_fixed(angle_fixed(g * i, 8 * n))
        memo.append((fixed_to_float(c), fixed_to_float(s)))
    return [_reduce(lambda num: memo[num // g], m, n) for m in range(n)]


def cst_value(c: tuple) -> float:
    """SExpr.lean:41-45 (Cst.toFloat): cos/sin via cosSin2Pi, sqrthalf = cos(2*pi/8),
    tan via tan2Pi."""
    kind = c[0]
    if kind == "cos":
        return cos_sin_2pi(c[1], c[2])[0]
    if kind == "sin":
        return cos_sin_2pi(c[1], c[2])[1]
    if kind == "sqrthalf":
        return cos_sin_2pi(1, 8)[0]
    if kind == "tan":
        return tan_2pi(c[1], c[2])
    raise ValueError(c)


def pass_twiddles(tab, fwd: bool, N: int, R: int, ns: int) -> list[float]:
    """FastPlan.lean:27-37 -- for r in [1, R): ns real parts of w^(r*k*N/(ns*R)),
    then ns imaginary parts (negZ'd when forward)."""
    step = N // (ns * R)
    out: list[float] = []
    for r in range(1, R):
        idx = [r * k * step for k in range(ns)]
        out.extend(tab[i][0] if i < N else 1.0 for i in idx)
        for i in idx:
            s = tab[i][1] if i < N else 0.0
            out.append(neg_z(s) if fwd else s)
    return out


def pass_ns(rs, t: int) -> int:
    """FastPlan.lean:22 -- product of the radices before pass t."""
    p = 1
    for r in rs[:t]:
        p *= r
    return p
