"""Device-code generator for the radix-R butterflies of the Stockham passes.

Emits csrc/bfly_gen.h: one straight-line HIP function per (direction, radix, twiddled),
computing the same binary64 operation DAG as the reference's butterflyK
(LeanFFTW/Gen/Stockham.lean:65-66), so that a pass on the GPU is bit-identical to the
reference schedule on the same twiddle tables:

  * radix 2^k: the generic decimation-in-time exec/mkPlan run on symbolic values
    (LeanFFTW/Core/Alg.lean:69-87, Gen/Stockham.lean:49);
  * other radices: Good-Thomas over the prime-power blocks (Gen/SymDFT.lean:276-285),
    mixed radix inside a block (:246-255), paired odd-length leaves (:188-235);
  * constant rotations as i^q times cos*(a -+ tan*b) (Gen/Sym.lean:48-74), twiddle inputs
    multiplied by fms/fma (Gen/Sym.lean:94-97), smart constructors (Gen/SExpr.lean:116-171),
    then FMA contraction (Gen/SExpr.lean:174-176).

Constants are exact hexadecimal literals of the reference's integer trigonometry
(LeanFFTW/Core/Trig.lean). The generated code must be compiled with
-ffp-contract=off: every fused operation is an explicit fma().

Run:  python lean-fftw_amd/codegen.py   (also run by the build)
"""

from __future__ import annotations

import math
import os
import struct
import sys

RADICES = (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16)

# ------------------------------------------------------------------ constants
# Fixed-point cos/sin/tan of 2*pi*num/den (LeanFFTW/Core/Trig.lean:25-134).

_P = 160
_PI = 3141592653589793238462643383279502884197169399375105820974944 * (1 << _P) // 10 ** 60


def _series(theta: int):
    t2 = theta * theta >> _P
    cp, cn, sp, sn, tc, ts = 1 << _P, 0, theta, 0, 1 << _P, theta
    for j in range(1, 40):
        tc = (tc * t2 >> _P) // ((2 * j - 1) * 2 * j)
        ts = (ts * t2 >> _P) // (2 * j * (2 * j + 1))
        if j & 1:
            cn, sn = cn + tc, sn + ts
        else:
            cp, sp = cp + tc, sp + ts
    return cp - cn, sp - sn


def _round(v: int) -> float:
    if v == 0:
        return 0.0
    e = v.bit_length() - 1
    if e > 52:
        sh = e - 52
        q, r = v >> sh, v & ((1 << sh) - 1)
        h = 1 << (sh - 1)
        q += 1 if (r > h or (r == h and q & 1)) else 0
        if q >> 53:
            q, e = q >> 1, e + 1
    else:
        q = v << (52 - e)
    return struct.unpack("<d", struct.pack("<Q", ((e - _P + 1023) << 52) | (q & ((1 << 52) - 1))))[0]


def _negz(x):
    return 0.0 if x == 0.0 else -x


def cos_sin(k: int, n: int):
    k %= n
    o, rem = divmod(8 * k, n)
    num = rem if o % 2 == 0 else n - rem
    c, s = (_round(v) for v in _series(2 * _PI * num // (8 * n)))
    return [(c, s), (s, c), (_negz(s), c), (_negz(c), s), (_negz(c), _negz(s)),
            (_negz(s), _negz(c)), (s, _negz(c)), (c, _negz(s))][o]


def const_value(c) -> float:
    kind = c[0]
    if kind == "sqrthalf":
        return cos_sin(1, 8)[0]
    if kind == "tan":
        cc, ss = _series(2 * _PI * c[1] // c[2])
        return _round(ss * (1 << _P) // cc)
    cs = cos_sin(c[1], c[2])
    return cs[0] if kind == "cos" else cs[1]


# ------------------------------------------------------------------ node table
# A node is an int id; self.op[id] is (kind, operands...). Identical nodes share an id.


class Dag:
    def __init__(self, n_in: int):
        self.op: list = []
        self.ids: dict = {}
        self.inputs = [self._node(("in", i)) for i in range(n_in)]
        self.zero = self._node(("zero",))
        self._contracted: dict = {}

    def _node(self, key) -> int:
        i = self.ids.get(key)
        if i is None:
            i = len(self.op)
            self.op.append(key)
            self.ids[key] = i
        return i

    def kind(self, i):
        return self.op[i][0]

    # raw constructors
    def c(self, const):
        return self._node(("cst", const))

    def raw(self, kind, *xs):
        return self._node((kind,) + xs)

    # smart constructors (SExpr.lean:116-171)
    def neg(self, a):
        if a == self.zero:
            return a
        if self.kind(a) == "neg":
            return self.op[a][1]
        return self.raw("neg", a)

    def add(self, a, b):
        if a == self.zero:
            return b
        if b == self.zero:
            return a
        if self.kind(b) == "neg":
            return self.raw("sub", a, self.op[b][1])
        if self.kind(a) == "neg":
            return self.raw("sub", b, self.op[a][1])
        return self.raw("add", a, b)

    def sub(self, a, b):
        if b == self.zero:
            return a
        if a == self.zero:
            return self.neg(b)
        if self.kind(b) == "neg":
            return self.raw("add", a, self.op[b][1])
        if self.kind(a) == "neg":
            return self.neg(self.raw("add", self.op[a][1], b))
        return self.raw("sub", a, b)

    def mulc(self, const, a):
        if a == self.zero:
            return a
        if self.kind(a) == "neg":
            return self.neg(self.raw("mul", self.c(const), self.op[a][1]))
        return self.raw("mul", self.c(const), a)

    def sub_tb(self, t, b, a):  # a - t*b
        return a if b == self.zero else self.raw("fnma", self.c(t), b, a)

    def add_ta(self, t, a, b):  # t*a + b
        return b if a == self.zero else self.raw("fma", self.c(t), a, b)

    def ta_sub(self, t, a, b):  # t*a - b
        return self.neg(b) if a == self.zero else self.raw("fms", self.c(t), a, b)

    # contraction (SExpr.lean:174-176), memoised per node
    def contract(self, i):
        r = self._contracted.get(i)
        if r is not None:
            return r
        k = self.op[i]
        if k[0] in ("add", "sub"):
            a, b = self.contract(k[1]), self.contract(k[2])
            if self.kind(a) == "mul":
                r = self.raw("fma" if k[0] == "add" else "fms", self.op[a][1], self.op[a][2], b)
            elif self.kind(b) == "mul":
                r = self.raw("fma" if k[0] == "add" else "fnma", self.op[b][1], self.op[b][2], a)
            else:
                r = self.raw(k[0], a, b)
        elif k[0] in ("mul", "neg", "fma", "fms", "fnma"):
            r = self.raw(k[0], *[self.contract(x) for x in k[1:]])
        else:
            r = i
        self._contracted[i] = r
        return r


# ------------------------------------------------------------------ complex layer
# A complex value is ("w", fwd, n, k) for the twiddle exp(-+2 pi i k/n) or (re, im).


class Cx:
    def __init__(self, dag: Dag):
        self.d = dag
        self.czero = (dag.zero, dag.zero)

    def ipow(self, q, re, im):
        d = self.d
        return [(re, im), (d.neg(im), re), (d.neg(re), d.neg(im)), (im, d.neg(re))][q % 4]

    def rot_first_quadrant(self, k, n, a, b):
        d = self.d
        if k == 0:
            return a, b
        if 8 * k == n:
            return d.mulc(("sqrthalf",), d.sub(a, b)), d.mulc(("sqrthalf",), d.add(a, b))
        if 8 * k < n:
            cc, tt = ("cos", k, n), ("tan", k, n)
            return d.mulc(cc, d.sub_tb(tt, b, a)), d.mulc(cc, d.add_ta(tt, a, b))
        k2 = n - 4 * k
        cc, tt = ("cos", k2, 4 * n), ("tan", k2, 4 * n)
        return d.mulc(cc, d.ta_sub(tt, a, b)), d.mulc(cc, d.add_ta(tt, b, a))

    def rotate(self, fwd, n, k, a, b):
        k %= n
        if fwd:
            k = (n - k) % n
        q = 4 * k // n
        a, b = self.rot_first_quadrant(4 * k - q * n, 4 * n, a, b)
        return self.ipow(q, a, b)

    def split(self, x):
        if x[0] == "w":
            return self.rotate(x[1], x[2], x[3], self.d.c(("cos", 0, 1)), self.d.zero)
        return x

    def plus(self, x, y):
        (a, b), (c, e) = self.split(x), self.split(y)
        return (self.d.add(a, c), self.d.add(b, e))

    def times(self, x, y):
        if x[0] == "w" and y[0] != "w":
            return self.rotate(x[1], x[2], x[3], *y)
        if y[0] == "w" and x[0] != "w":
            return self.rotate(y[1], y[2], y[3], *x)
        if x[0] == "w":
            return self.rotate(x[1], x[2], x[3], *self.split(y))
        return self.general(x, y)

    def general(self, x, y):
        d = self.d
        (a, b), (c, e) = self.split(x), self.split(y)
        return (d.raw("fms", a, c, d.raw("mul", b, e)), d.raw("fma", a, e, d.raw("mul", b, c)))

    # ---- Alg.lean exec over a DIT factor list (2^k radices)
    def dit(self, fwd, fs, x):
        if len(fs) <= 1:
            p = fs[0] if fs else 1
            return self.direct(fwd, p, x)
        r = fs[0]
        m = math.prod(fs[1:])
        n = r * m
        ys = [self.dit(fwd, fs[1:], lambda j2, j1=j1: x(j1 + r * j2)) for j1 in range(r)]
        out = []
        for k in range(n):
            acc = self.czero
            for j1 in range(r):
                acc = self.plus(acc, self.times(("w", fwd, n, j1 * k % n), ys[j1][k % m]))
            out.append(acc)
        return out

    def direct(self, fwd, p, x):
        out = []
        for k in range(p):
            acc = self.czero
            for j in range(p):
                acc = self.plus(acc, self.times(("w", fwd, p, j * k % p), x(j)))
            out.append(acc)
        return out

    # ---- SymDFT.lean
    @staticmethod
    def _cos_sign(m, n):
        m = m if 2 * m <= n else n - m
        return (False, ("cos", m, n)) if 4 * m <= n else (True, ("cos", n - 2 * m, 2 * n))

    @staticmethod
    def _sin_half(m, n):
        return ("sin", m, n) if 4 * m <= n else ("sin", n - 2 * m, 2 * n)

    def _sin_sign(self, m, n):
        return (False, self._sin_half(m, n)) if 2 * m <= n else (True, self._sin_half(n - m, n))

    def _acc(self, init, terms):
        d = self.d
        for ng, c, a in terms:
            init = d.sub_tb(c, a, init) if ng else d.add_ta(c, a, init)
        return init

    def _sum(self, terms):
        d = self.d
        if not terms:
            return d.zero
        ng, c, a = terms[0]
        first = d.neg(d.mulc(c, a)) if ng else d.mulc(c, a)
        return self._acc(first, terms[1:])

    def odd_prime(self, fwd, p, x):
        d = self.d
        xs = [self.split(x(i)) for i in range(p)]
        h = p // 2
        sums = [(d.add(xs[i + 1][0], xs[p - i - 1][0]), d.add(xs[i + 1][1], xs[p - i - 1][1]))
                for i in range(h)]
        difs = [(d.sub(xs[i + 1][0], xs[p - i - 1][0]), d.sub(xs[i + 1][1], xs[p - i - 1][1]))
                for i in range(h)]

        def a_part(k, part):
            return self._acc(xs[0][part], [self._cos_sign((i + 1) * k % p, p) + (sums[i][part],)
                                           for i in range(h)])

        def s_part(k, part):
            return self._sum([self._sin_sign((i + 1) * k % p, p) + (difs[i][part],)
                              for i in range(h)])

        def plus_i(k):   # A_k + i S_k
            return (d.sub(a_part(k, 0), s_part(k, 1)), d.add(a_part(k, 1), s_part(k, 0)))

        def minus_i(k):  # A_k - i S_k
            return (d.add(a_part(k, 0), s_part(k, 1)), d.sub(a_part(k, 1), s_part(k, 0)))

        re, im = xs[0]
        for i in range(h):
            re = d.add(re, sums[i][0])
        for i in range(h):
            im = d.add(im, sums[i][1])
        out = [(re, im)]
        for k in range(1, p):
            if k <= h:
                out.append(minus_i(k) if fwd else plus_i(k))
            else:
                out.append(plus_i(p - k) if fwd else minus_i(p - k))
        return out

    def prime_leaf(self, fwd, p, x):
        if p % 2 == 1 and p >= 3:
            return self.odd_prime(fwd, p, x)
        return self.direct(fwd, p, x)

    def mixed(self, fwd, fs, x):
        if len(fs) <= 1:
            return self.prime_leaf(fwd, fs[0] if fs else 1, x)
        r, m = fs[0], math.prod(fs[1:])
        ys = [self.mixed(fwd, fs[1:], lambda j2, j1=j1: x(j1 + r * j2)) for j1 in range(r)]
        cols = [self.prime_leaf(fwd, r, lambda j1, k1=k1: self.times(("w", fwd, r * m, j1 * k1),
                                                                      ys[j1][k1]))
                for k1 in range(m)]
        return [cols[k % m][k // m] for k in range(r * m)]

    def good_thomas(self, fwd, blocks, x):
        if len(blocks) <= 1:
            return self.mixed(fwd, factor(blocks[0]) if blocks else [], x)
        b, m = blocks[0], math.prod(blocks[1:])
        ys = [self.good_thomas(fwd, blocks[1:], lambda j2, j1=j1: x((m * j1 + b * j2) % (b * m)))
              for j1 in range(b)]
        cols = [self.mixed(fwd, factor(b), lambda j1, k2=k2: ys[j1][k2]) for k2 in range(m)]
        return [cols[k % m][k % b] for k in range(b * m)]


def factor(n: int) -> list[int]:
    fs, p = [], 2
    while n > 1:
        if p * p > n:
            fs.append(n)
            break
        while n % p == 0:
            fs.append(p)
            n //= p
        p += 1
    return fs


def blocks_of(n: int) -> list[int]:
    fs = factor(n)
    order = sorted(set(fs), key=fs.index)
    return [p ** fs.count(p) for p in order]


def butterfly_dag(fwd: bool, R: int, tw: bool):
    n_in = 4 * R - 2 if tw else 2 * R
    d = Dag(n_in)
    cx = Cx(d)
    I = d.inputs

    def x(r):
        v = (I[2 * r], I[2 * r + 1])
        if tw and r >= 1:
            w = (I[2 * R + 2 * (r - 1)], I[2 * R + 2 * (r - 1) + 1])
            return cx.general(w, v)
        return v

    if R & (R - 1) == 0:
        ys = cx.dit(fwd, factor(R), x)
    else:
        ys = cx.good_thomas(fwd, blocks_of(R), x)
    outs = []
    for y in ys:
        re, im = cx.split(y)
        outs += [d.contract(re), d.contract(im)]
    return d, outs


# ------------------------------------------------------------------ emission


def op_count(fwd: bool, R: int, tw: bool) -> int:
    """Arithmetic operations reachable from the outputs (constants and inputs are free)."""
    d, outs = butterfly_dag(fwd, R, tw)
    seen, stack = set(), list(outs)
    while stack:
        i = stack.pop()
        if i in seen or d.op[i][0] in ("in", "cst", "zero"):
            continue
        seen.add(i)
        stack.extend(d.op[i][1:])
    return len(seen)


def _lit(v: float) -> str:
    return "0.0" if v == 0.0 else v.hex()


def op_kinds(fwd: bool, R: int, tw: bool) -> tuple:
    """(additions, multiplications, fused multiply-adds) reachable from the outputs: add, sub
    and neg count as additions, fma, fms and fnma as fused operations (the kinds of the
    reference's Weights, LeanFFTW/Gen/OpCount.lean:23-49)."""
    d, outs = butterfly_dag(fwd, R, tw)
    seen, stack = set(), list(outs)
    n = {"add": 0, "mul": 0, "fma": 0}
    kind = {"add": "add", "sub": "add", "neg": "add", "mul": "mul", "fma": "fma", "fms": "fma",
            "fnma": "fma"}
    while stack:
        i = stack.pop()
        if i in seen or d.op[i][0] in ("in", "cst", "zero"):
            continue
        seen.add(i)
        n[kind[d.op[i][0]]] += 1
        stack.extend(d.op[i][1:])
    return n["add"], n["mul"], n["fma"]


def emit_butterfly(fwd: bool, R: int, tw: bool) -> str:
    d, outs = butterfly_dag(fwd, R, tw)
    # reachable nodes in id order (ids are created after their operands: topological)
    need = set()
    stack = list(outs)
    while stack:
        i = stack.pop()
        if i in need:
            continue
        need.add(i)
        stack.extend(x for x in d.op[i][1:] if isinstance(x, int) and d.op[i][0] not in ("in", "cst"))
    names = {}
    body = []
    for i in sorted(need):
        k = d.op[i]
        if k[0] == "in":
            s = k[1]
            names[i] = f"x[{s // 2}].{'xy'[s % 2]}" if s < 2 * R else f"w[{(s - 2 * R) // 2}].{'xy'[s % 2]}"
            continue
        if k[0] == "cst":
            names[i] = _lit(const_value(k[1]))
            continue
        if k[0] == "zero":
            names[i] = "0.0"
            continue
        a = [names[j] for j in k[1:]]
        expr = {
            "add": lambda: f"{a[0]} + {a[1]}",
            "sub": lambda: f"{a[0]} - {a[1]}",
            "mul": lambda: f"{a[0]} * {a[1]}",
            "neg": lambda: f"-{a[0]}",
            "fma": lambda: f"fma({a[0]}, {a[1]}, {a[2]})",
            "fms": lambda: f"fma({a[0]}, {a[1]}, -{a[2]})",
            "fnma": lambda: f"fma(-{a[0]}, {a[1]}, {a[2]})",
        }[k[0]]()
        names[i] = f"t{i}"
        body.append(f"  const double t{i} = {expr};")
    sig = "double2 (&x)[%d], const double2 (&w)[%d]" % (R, R - 1) if tw else "double2 (&x)[%d]" % R
    head = f"template <> struct Bfly<{'true' if fwd else 'false'}, {R}, {'true' if tw else 'false'}> {{\n" \
           f"  static __device__ __forceinline__ void run({sig}) {{"
    tail = [f"    x[{r}] = make_double2({names[outs[2 * r]]}, {names[outs[2 * r + 1]]});"
            for r in range(R)]
    return "\n".join([head] + ["  " + b for b in body] + tail + ["  }", "};"])


def emit_all(radices=RADICES) -> str:
    out = [
        "// Generated by lean-fftw_amd/codegen.py from the reference's butterfly algorithm",
        "// (LeanFFTW/Gen/Stockham.lean:65-66). Do not edit. Compile with -ffp-contract=off.",
        "#pragma once",
        "#include <hip/hip_runtime.h>",
        "namespace mfft {",
        "// Bfly<FWD, R, TW>::run(x[, w]): x holds the R inputs and receives the R outputs;",
        "// w holds the twiddles w_1..w_{R-1} of a twiddled pass.",
        "template <bool FWD, int R, bool TW> struct Bfly;",
    ]
    for R in radices:
        for fwd in (True, False):
            for tw in (False, True):
                out.append(emit_butterfly(fwd, R, tw))
    ops = [op_count(True, R, True) if R in radices else 0 for R in range(17)]
    out.append("// arithmetic operations of the twiddled forward butterfly of each radix")
    out.append("constexpr int kBflyOps[17] = {%s};" % ", ".join(map(str, ops)))
    out.append("// (additions, multiplications, fused multiply-adds) of butterfly [forward][twiddled][R]")
    rows = []
    for fwd in (False, True):
        for tw in (False, True):
            cells = ["{%d, %d, %d}" % (op_kinds(fwd, R, tw) if R in radices else (0, 0, 0))
                     for R in range(17)]
            rows.append("{" + ", ".join(cells) + "}")
    out.append("constexpr int kBflyOpKinds[2][2][17][3] = {{%s, %s}, {%s, %s}};" % tuple(rows))
    out.append("}  // namespace mfft")
    return "\n".join(out) + "\n"


def main(path=None):
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "bfly_gen.h")
    text = emit_all()
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


if __name__ == "__main__":
    print(main(sys.argv[1] if len(sys.argv) > 1 else None))
