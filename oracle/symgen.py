"""Oracle: symbolic butterflies of the Stockham passes (test infrastructure only).

Part of the CPU oracle: only the tests, the smoke check and the CPU-baseline leg of
the benchmark may use it. The shipped library generates its own device code.

Restatement of the reference's kernel generator:
  * expressions and smart constructors  -- LeanFFTW/Gen/SExpr.lean:58-176
  * symbolic complex values, rotations   -- LeanFFTW/Gen/Sym.lean:24-118
  * generic DIT exec/mkPlan (2^k radix)  -- LeanFFTW/Core/Alg.lean:64-106
  * odd leaves, mixed radix, Good-Thomas -- LeanFFTW/Gen/SymDFT.lean:151-285
  * butterfly / butterflyK inputs        -- LeanFFTW/Gen/Stockham.lean:36-66
The binary64 value of every output is fixed by the contracted expression tree
(SExpr.evalF, SExpr.lean:77-87); CSE (Emit.lean:120-175) only shares identical
subtrees and orders commutative operands, which does not change any value, so the
oracle evaluates the contracted trees directly.

Expressions are hash-consed: ``E`` objects are unique per structure, so identity
is structural equality and memoised evaluation is linear in the DAG size.
"""

from __future__ import annotations

from functools import lru_cache

# --------------------------------------------------------------------------------------
# Expressions (SExpr.lean:58-72)
# --------------------------------------------------------------------------------------

_INTERN: dict = {}


class E:
    __slots__ = ("op", "a", "uid", "__weakref__")

    def __repr__(self):  # pragma: no cover - debugging aid
        return f"E{self.uid}:{self.op}"


def _mk(op: str, *a) -> E:
    key = (op,) + tuple(x.uid if isinstance(x, E) else x for x in a)
    e = _INTERN.get(key)
    if e is None:
        e = E()
        e.op = op
        e.a = a
        e.uid = len(_INTERN)
        _INTERN[key] = e
    return e


def inp(i: int) -> E:
    return _mk("in", i)


ZERO = _mk("zero")


def cst(c: tuple) -> E:
    return _mk("cst", c)


def add(a, b):
    return _mk("add", a, b)


def sub(a, b):
    return _mk("sub", a, b)


def mul(a, b):
    return _mk("mul", a, b)


def neg(a):
    return _mk("neg", a)


def fma(a, b, c):
    return _mk("fma", a, b, c)


def fms(a, b, c):
    return _mk("fms", a, b, c)


def fnma(a, b, c):
    return _mk("fnma", a, b, c)


# Smart constructors, SExpr.lean:116-171


def mk_neg(a: E) -> E:
    if a.op == "zero":
        return ZERO
    if a.op == "neg":
        return a.a[0]
    return neg(a)


def mk_add(a: E, b: E) -> E:
    if a.op == "zero":
        return b
    if b.op == "zero":
        return a
    if b.op == "neg":
        return sub(a, b.a[0])
    if a.op == "neg":
        return sub(b, a.a[0])
    return add(a, b)


def mk_sub(a: E, b: E) -> E:
    if b.op == "zero":
        return a
    if a.op == "zero":
        return mk_neg(b)
    if b.op == "neg":
        return add(a, b.a[0])
    if a.op == "neg":
        return mk_neg(add(a.a[0], b))
    return sub(a, b)


def mk_mul_c(c: tuple, a: E) -> E:
    if a.op == "zero":
        return ZERO
    if a.op == "neg":
        return mk_neg(mul(cst(c), a.a[0]))
    return mul(cst(c), a)


def mk_fnma_c(t: tuple, b: E, a: E) -> E:
    """a - t*b"""
    if b.op == "zero":
        return a
    return fnma(cst(t), b, a)


def mk_fma_c(t: tuple, a: E, b: E) -> E:
    """t*a + b"""
    if a.op == "zero":
        return b
    return fma(cst(t), a, b)


def mk_fms_c(t: tuple, a: E, b: E) -> E:
    """t*a - b"""
    if a.op == "zero":
        return mk_neg(b)
    return fms(cst(t), a, b)


# FMA contraction, SExpr.lean:174-176 (contractAdd/contractSub/contract)

_CONTRACT: dict = {}


def contract(e: E) -> E:
    r = _CONTRACT.get(e.uid)
    if r is not None:
        return r
    op = e.op
    if op == "add":
        a, b = contract(e.a[0]), contract(e.a[1])
        if a.op == "mul":
            r = fma(a.a[0], a.a[1], b)
        elif b.op == "mul":
            r = fma(b.a[0], b.a[1], a)
        else:
            r = add(a, b)
    elif op == "sub":
        a, b = contract(e.a[0]), contract(e.a[1])
        if a.op == "mul":
            r = fms(a.a[0], a.a[1], b)
        elif b.op == "mul":
            r = fnma(b.a[0], b.a[1], a)
        else:
            r = sub(a, b)
    elif op == "mul":
        r = mul(contract(e.a[0]), contract(e.a[1]))
    elif op == "neg":
        r = neg(contract(e.a[0]))
    elif op in ("fma", "fms", "fnma"):
        r = _mk(op, contract(e.a[0]), contract(e.a[1]), contract(e.a[2]))
    else:
        r = e
    _CONTRACT[e.uid] = r
    return r


# --------------------------------------------------------------------------------------
# Symbolic complex values (Sym.lean:24-118)
# --------------------------------------------------------------------------------------
# ("tw", fwd, n, k)  -- the twiddle exp(-+2*pi*i*k/n)
# ("val", re, im)


def val(re: E, im: E):
    return ("val", re, im)


SZERO = val(ZERO, ZERO)
ONE_CST = ("cos", 0, 1)


def mul_ipow(q: int, re: E, im: E):
    q %= 4
    if q == 0:
        return re, im
    if q == 1:
        return mk_neg(im), re
    if q == 2:
        return mk_neg(re), mk_neg(im)
    return im, mk_neg(re)


def rot_q(k: int, n: int, a: E, b: E):
    """Sym.lean:48-63 -- multiply a+ib by exp(i*2*pi*k/n), 0 <= 2*pi*k/n < pi/2."""
    if k == 0:
        return a, b
    if 8 * k == n:
        return mk_mul_c(("sqrthalf",), mk_sub(a, b)), mk_mul_c(("sqrthalf",), mk_add(a, b))
    if 8 * k < n:
        c, t = ("cos", k, n), ("tan", k, n)
        return mk_mul_c(c, mk_fnma_c(t, b, a)), mk_mul_c(c, mk_fma_c(t, a, b))
    k2 = n - 4 * k
    c, t = ("cos", k2, 4 * n), ("tan", k2, 4 * n)
    return mk_mul_c(c, mk_fms_c(t, a, b)), mk_mul_c(c, mk_fma_c(t, b, a))


def rot(fwd: bool, n: int, k: int, a: E, b: E):
    """Sym.lean:66-74"""
    k %= n
    if fwd:
        k = (n - k) % n
    q = 4 * k // n
    k1 = 4 * k - q * n
    a2, b2 = rot_q(k1, 4 * n, a, b)
    return mul_ipow(q, a2, b2)


def parts(x):
    if x[0] == "val":
        return x[1], x[2]
    _, fwd, n, k = x
    return rot(fwd, n, k, cst(ONE_CST), ZERO)


def sc_add(x, y):
    a, b = parts(x)
    c, d = parts(y)
    return val(mk_add(a, c), mk_add(b, d))


def sc_sub(x, y):
    a, b = parts(x)
    c, d = parts(y)
    return val(mk_sub(a, c), mk_sub(b, d))


def mul_gen(x, y):
    a, b = parts(x)
    c, d = parts(y)
    return val(fms(a, c, mul(b, d)), fma(a, d, mul(b, c)))


def sc_mul(x, y):
    if x[0] == "tw" and y[0] == "val":
        return val(*rot(x[1], x[2], x[3], y[1], y[2]))
    if x[0] == "val" and y[0] == "tw":
        return val(*rot(y[1], y[2], y[3], x[1], x[2]))
    if x[0] == "tw" and y[0] == "tw":
        a, b = parts(y)
        return val(*rot(x[1], x[2], x[3], a, b))
    return mul_gen(x, y)


def sc_tw(fwd: bool, n: int, k: int):
    return ("tw", fwd, n, k)


def _get(v, i):
    return v[i] if i < len(v) else SZERO


# --------------------------------------------------------------------------------------
# Generic algorithms over the symbolic ops (Alg.lean:64-106)
# --------------------------------------------------------------------------------------


def prime_factors(n: int) -> list[int]:
    out, d = [], 2
    while n > 1:
        if d * d > n:
            out.append(n)
            break
        if n % d == 0:
            out.append(d)
            n //= d
        else:
            d += 1
    return out


def _sum_range(n, f):
    acc = SZERO
    for j in range(n):
        acc = sc_add(acc, f(j))
    return acc


def mk_plan(fwd: bool, fs: list[int]):
    """Alg.lean:69-75 -- DIT tree, the last factor a direct leaf."""
    if not fs:
        return ("direct", 1, [sc_tw(fwd, 1, i) for i in range(1)])
    if len(fs) == 1:
        p = fs[0]
        return ("direct", p, [sc_tw(fwd, p, i) for i in range(p)])
    sub_ = mk_plan(fwd, fs[1:])
    n = 1
    for f in fs:
        n *= f
    return ("ct", fs[0], [sc_tw(fwd, n, i) for i in range(n)], sub_)


def plan_size(p) -> int:
    return p[1] if p[0] == "direct" else p[1] * plan_size(p[3])


def alg_exec(plan, x):
    """Alg.lean:78-87"""
    if plan[0] == "direct":
        _, n, w = plan
        return [_sum_range(n, lambda j, k=k: sc_mul(_get(w, j * k % n), x(j))) for k in range(n)]
    _, r, w, sub_ = plan
    m = plan_size(sub_)
    n = r * m
    ys = [alg_exec(sub_, lambda j2, j1=j1: x(j1 + r * j2)) for j1 in range(r)]
    return [
        _sum_range(r, lambda j1, k=k: sc_mul(_get(w, j1 * k % n), _get(ys[j1], k % m)))
        for k in range(n)
    ]


# --------------------------------------------------------------------------------------
# Odd leaves, mixed radix, Good-Thomas (SymDFT.lean:151-285)
# --------------------------------------------------------------------------------------


def cos_h(m, n):
    return (False, ("cos", m, n)) if 4 * m <= n else (True, ("cos", n - 2 * m, 2 * n))


def cos_q(m, n):
    return cos_h(m if 2 * m <= n else n - m, n)


def sin_h(m, n):
    return ("sin", m, n) if 4 * m <= n else ("sin", n - 2 * m, 2 * n)


def sin_q(m, n):
    return (False, sin_h(m, n)) if 2 * m <= n else (True, sin_h(n - m, n))


def sfma(ng, c, a, acc):
    return mk_fnma_c(c, a, acc) if ng else mk_fma_c(c, a, acc)


def smul(ng, c, a):
    return mk_neg(mk_mul_c(c, a)) if ng else mk_mul_c(c, a)


def lacc(init, ts):
    acc = init
    for ng, c, a in ts:
        acc = sfma(ng, c, a, acc)
    return acc


def lsum(ts):
    if not ts:
        return ZERO
    ng, c, a = ts[0]
    return lacc(smul(ng, c, a), ts[1:])


def _pair_a(p, x, i):
    u, v = parts(x(i + 1)), parts(x(p - (i + 1)))
    return mk_add(u[0], v[0]), mk_add(u[1], v[1])


def _pair_b(p, x, i):
    u, v = parts(x(i + 1)), parts(x(p - (i + 1)))
    return mk_sub(u[0], v[0]), mk_sub(u[1], v[1])


def _c_terms(p, k, x, part):
    out = []
    for i in range(p // 2):
        ng, c = cos_q((i + 1) * k % p, p)
        out.append((ng, c, _pair_a(p, x, i)[part]))
    return out


def _s_terms(p, k, x, part):
    out = []
    for i in range(p // 2):
        ng, c = sin_q((i + 1) * k % p, p)
        out.append((ng, c, _pair_b(p, x, i)[part]))
    return out


def _odd_a(p, k, x):
    x0 = parts(x(0))
    return lacc(x0[0], _c_terms(p, k, x, 0)), lacc(x0[1], _c_terms(p, k, x, 1))


def _odd_s(p, k, x):
    return lsum(_s_terms(p, k, x, 0)), lsum(_s_terms(p, k, x, 1))


def _odd_plus(p, k, x):
    a, s = _odd_a(p, k, x), _odd_s(p, k, x)
    return val(mk_sub(a[0], s[1]), mk_add(a[1], s[0]))


def _odd_minus(p, k, x):
    a, s = _odd_a(p, k, x), _odd_s(p, k, x)
    return val(mk_add(a[0], s[1]), mk_sub(a[1], s[0]))


def _odd_x0(p, x):
    x0 = parts(x(0))
    re, im = x0
    for i in range(p // 2):
        re = mk_add(re, _pair_a(p, x, i)[0])
    for i in range(p // 2):
        im = mk_add(im, _pair_a(p, x, i)[1])
    return val(re, im)


def odd_leaf(fwd, p, x):
    """SymDFT.lean:231-235"""
    out = []
    for k in range(p):
        if k == 0:
            out.append(_odd_x0(p, x))
        elif k <= p // 2:
            out.append(_odd_minus(p, k, x) if fwd else _odd_plus(p, k, x))
        else:
            out.append(_odd_plus(p, p - k, x) if fwd else _odd_minus(p, p - k, x))
    return out


def leaf_sym(fwd, p, x):
    """SymDFT.lean:239-241"""
    if p % 2 == 1 and p >= 3:
        return odd_leaf(fwd, p, x)
    return alg_exec(("direct", p, [sc_tw(fwd, p, i) for i in range(p)]), x)


def sym_dft(fwd, fs, x):
    """SymDFT.lean:246-255"""
    if not fs:
        return leaf_sym(fwd, 1, x)
    if len(fs) == 1:
        return leaf_sym(fwd, fs[0], x)
    r, rest = fs[0], fs[1:]
    m = 1
    for f in rest:
        m *= f
    ys = [sym_dft(fwd, rest, lambda j2, j1=j1: x(j1 + r * j2)) for j1 in range(r)]
    cols = [
        leaf_sym(fwd, r, lambda j1, k1=k1: sc_mul(sc_tw(fwd, r * m, j1 * k1), _get(ys[j1], k1)))
        for k1 in range(m)
    ]
    return [_get(cols[k % m], k // m) for k in range(r * m)]


def pp_blocks(n):
    """SymDFT.lean:261-263 -- one prime power per distinct prime, in order."""
    fs = prime_factors(n)
    seen = []
    for p in fs:
        if p not in seen:
            seen.append(p)
    return [p ** fs.count(p) for p in seen]


def pfa_dft(fwd, bs, x):
    """SymDFT.lean:276-285"""
    if not bs:
        return sym_dft(fwd, [], x)
    if len(bs) == 1:
        return sym_dft(fwd, prime_factors(bs[0]), x)
    b, rest = bs[0], bs[1:]
    m = 1
    for f in rest:
        m *= f
    ys = [pfa_dft(fwd, rest, lambda j2, j1=j1: x((m * j1 + b * j2) % (b * m))) for j1 in range(b)]
    cols = [
        sym_dft(fwd, prime_factors(b), lambda j1, k2=k2: _get(ys[j1], k2)) for k2 in range(m)
    ]
    return [_get(cols[k % m], k % b) for k in range(b * m)]


# --------------------------------------------------------------------------------------
# Butterflies (Stockham.lean:36-66)
# --------------------------------------------------------------------------------------


def bfly_in(R: int, twiddled: bool, r: int):
    xr = val(inp(2 * r), inp(2 * r + 1))
    if twiddled and r >= 1:
        w = val(inp(2 * R + 2 * (r - 1)), inp(2 * R + 2 * (r - 1) + 1))
        return mul_gen(w, xr)
    return xr


def bfly_sym(fwd: bool, R: int, twiddled: bool):
    x = lambda r: bfly_in(R, twiddled, r)  # noqa: E731
    if R & (R - 1) == 0:
        return alg_exec(mk_plan(fwd, prime_factors(R)), x)
    # PfaOK (SymDFT.lean:266-267) holds for every ppBlocks split
    return pfa_dft(fwd, pp_blocks(R), x)


@lru_cache(maxsize=None)
def butterfly(fwd: bool, R: int, twiddled: bool) -> tuple:
    """Stockham.lean:56-59 -- the 2R contracted output expressions re y0, im y0, ..."""
    y = bfly_sym(fwd, R, twiddled)
    outs = []
    for i in range(2 * R):
        p = parts(_get(y, i // 2))
        outs.append(contract(p[i % 2]))
    return tuple(outs)


def n_inputs(R: int, twiddled: bool) -> int:
    """Stockham.lean:62"""
    return 4 * R - 2 if twiddled else 2 * R


# --------------------------------------------------------------------------------------
# CSE'd kernel (Emit.lean:120-175) -- used for operation counts and C emission
# --------------------------------------------------------------------------------------


def cse_kernel(n_in: int, outs) -> tuple[list, list]:
    """Slots 0..n_in-1 are inputs; every other slot is one operation. Commutative
    operands are ordered by slot (ord2), identical operations share a slot."""
    table: dict = {}
    ops: list = []
    memo: dict = {}

    def node(op):
        s = table.get(op)
        if s is None:
            s = n_in + len(ops)
            table[op] = s
            ops.append(op)
        return s

    def ord2(x, y):
        return (x, y) if x <= y else (y, x)

    def go(e: E) -> int:
        if e.uid in memo:
            return memo[e.uid]
        o = e.op
        if o == "in":
            s = e.a[0]
        elif o == "zero":
            s = node(("zero",))
        elif o == "cst":
            s = node(("cst", e.a[0]))
        elif o in ("add", "mul"):
            a, b = go(e.a[0]), go(e.a[1])
            s = node((o,) + ord2(a, b))
        elif o == "sub":
            a, b = go(e.a[0]), go(e.a[1])
            s = node(("sub", a, b))
        elif o == "neg":
            s = node(("neg", go(e.a[0])))
        else:
            a, b = go(e.a[0]), go(e.a[1])
            c = go(e.a[2])
            s = node((o,) + ord2(a, b) + (c,))
        memo[e.uid] = s
        return s

    out_slots = [go(e) for e in outs]
    return ops, out_slots


def kernel_counts(fwd: bool, R: int, twiddled: bool) -> tuple[int, int, int]:
    """Emit.lean:178-183 -- (adds/subs/negs, muls, fmas) of butterflyK."""
    ops, _ = cse_kernel(n_inputs(R, twiddled), butterfly(fwd, R, twiddled))
    a = m = f = 0
    for op in ops:
        if op[0] in ("add", "sub", "neg"):
            a += 1
        elif op[0] == "mul":
            m += 1
        elif op[0] in ("fma", "fms", "fnma"):
            f += 1
    return a, m, f


def kernel_constants(fwd: bool, R: int, twiddled: bool) -> set:
    ops, _ = cse_kernel(n_inputs(R, twiddled), butterfly(fwd, R, twiddled))
    return {op[1] for op in ops if op[0] == "cst"}
