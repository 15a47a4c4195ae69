"""Pins the CPU oracle to the reference's own fixtures and to mathematics.

Each check is chosen so that a plausible mistake in the restatement fails it:
  * twiddle routine: 237 exact constant bit patterns (kernel-constants.txt), per-size
    table error tau (twiddles.jsonl), closed forms, mpmath correct rounding;
  * 2^k butterflies: vectors of the reference's printed kernels (Native/Kernels.lean);
  * every radix: operation counts published in docs/ITERATIONS.md, the constant set and
    its order of appearance, and the error-bound depths (a, b, magMax) of 36 plans
    (bounds.jsonl) -- the DAG shape of odd-radix and Good-Thomas butterflies;
  * schedule: closed forms (impulse, constant, tone), a quad-precision DFT, accuracy within
    2x of the reference's published L2 error;
  * the quad-precision DFT itself: mpmath at 200 bits (brute force, small N).
"""

import json
import math
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN


def f64(b):
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


# ---------------------------------------------------------------- twiddles


def test_kernel_constants_exact():
    from oracle import trig

    consts = json.load(open(os.path.join(GOLDEN, "kernel_constants.json")))
    assert len(consts) == 237
    for nm, k, n, b in consts:
        c = ("sqrthalf",) if (nm, k, n) == ("cos", 1, 8) else (nm, k, n)
        assert bits(trig.cst_value(c)) == b, (nm, k, n)


def test_twiddle_closed_forms():
    from oracle import trig

    assert trig.cos_sin_2pi(0, 7) == (1.0, 0.0)
    assert trig.cos_sin_2pi(1, 4) == (0.0, 1.0)
    assert trig.cos_sin_2pi(2, 4) == (-1.0, 0.0)
    assert trig.cos_sin_2pi(3, 4) == (0.0, -1.0)
    c, s = trig.cos_sin_2pi(1, 8)
    assert c == s == math.sqrt(0.5)
    # exact symmetries of the octant map (Trig.lean:114-122)
    for n in (12, 60, 1000, 2100):
        for k in range(1, n):
            c, s = trig.cos_sin_2pi(k, n)
            c2, s2 = trig.cos_sin_2pi(n - k, n)
            assert c == c2 and (s == -s2 or s == s2 == 0.0)
    # negZ: forward tables never hold -0.0
    assert bits(trig.neg_z(0.0)) == 0 and bits(trig.neg_z(-0.0)) == 0


def test_twiddles_vs_mpmath():
    import mpmath

    from oracle import trig

    mpmath.mp.prec = 200
    worst = 0.0
    for n in (3, 5, 7, 11, 13, 48, 360, 1000):
        for k in range(n):
            c, s = trig.cos_sin_2pi(k, n)
            tc = mpmath.cos(2 * mpmath.pi * k / n)
            ts = mpmath.sin(2 * mpmath.pi * k / n)
            for got, ref in ((c, tc), (s, ts)):
                if abs(ref) < mpmath.mpf(2) ** -150:  # exact zeros of the octant map
                    assert got == 0.0
                    continue
                err = abs(mpmath.mpf(got) - ref)
                worst = max(worst, float(err / math.ulp(float(ref))))
    # correctly rounded up to near-ties: never off by more than half an ulp (+ slack)
    assert worst <= 0.5 + 1e-9, worst


@pytest.mark.parametrize("n", [8, 16, 64, 1000, 4096, 10368])
def test_table_tau(n):
    """twiddles.jsonl: max |table - long-double cos/sin| per N, 3 significant digits."""
    from oracle import trig

    tau = dict((a, b) for a, b in json.load(open(os.path.join(GOLDEN, "twiddle_tau.json"))))
    if n not in tau:
        pytest.skip("size not published")
    tab = np.array(trig.cos_sin_2pi_table(n), dtype=np.longdouble)
    m = np.arange(n, dtype=np.longdouble)
    ang = 2 * np.pi * m / np.longdouble(n)
    ang = np.longdouble(2) * np.longdouble("3.14159265358979323846264338327950288") * m / n
    err = max(np.max(np.abs(tab[:, 0] - np.cos(ang))), np.max(np.abs(tab[:, 1] - np.sin(ang))))
    assert float(f"{float(err):.2e}") == pytest.approx(tau[n], rel=0.011), (float(err), tau[n])


def test_table_matches_pointwise():
    from oracle import trig

    for n in (6, 60, 360, 1001):
        tab = trig.cos_sin_2pi_table(n)
        assert all(tab[m] == trig.cos_sin_2pi(m, n) for m in range(n))


def test_pass_twiddle_layout():
    """FastPlan.lean:27-37: entry (r, k=0) is 1 + 0i; forward negates the imaginary part."""
    from oracle import trig

    N, rs = 120, [4, 6, 5]
    tab = trig.cos_sin_2pi_table(N)
    for t in (1, 2):
        R, ns = rs[t], trig.pass_ns(rs, t)
        f = trig.pass_twiddles(tab, True, N, R, ns)
        b = trig.pass_twiddles(tab, False, N, R, ns)
        assert len(f) == 2 * (R - 1) * ns
        for r in range(1, R):
            re = f[2 * (r - 1) * ns:(2 * r - 1) * ns]
            im = f[(2 * r - 1) * ns:2 * r * ns]
            assert re[0] == 1.0 and bits(im[0]) == 0
            for k in range(ns):
                c, s = trig.cos_sin_2pi(r * k * N // (ns * R), N)
                assert re[k] == c and im[k] == trig.neg_z(s)
                assert b[(2 * r - 1) * ns + k] == s


# ---------------------------------------------------------------- butterflies


def test_native_kernel_vectors(oracle):
    d = json.load(open(os.path.join(GOLDEN, "native_bfly.json")))
    assert len(d["cases"]) == 64
    for c in d["cases"]:
        s = np.array([f64(b) for b in c["in"]])
        y = oracle.butterfly_eval(c["fwd"], c["R"], c["tw"], s)
        assert [bits(v) for v in y] == c["out"], (c["R"], c["tw"], c["fwd"])


def test_op_counts():
    """docs/ITERATIONS.md:61-62, :226-229, :285-287."""
    from oracle.symgen import kernel_counts

    untw = {3: 12, 5: 36, 7: 66, 11: 150, 13: 204, 6: 36, 10: 92, 12: 96, 15: 168}
    for R, n in untw.items():
        assert sum(kernel_counts(True, R, False)) == n, R
    assert kernel_counts(True, 16, False) == (100, 0, 48)
    assert kernel_counts(True, 16, True) == (100, 30, 78)
    assert kernel_counts(True, 32, False) == (212, 0, 176)
    assert kernel_counts(True, 32, True) == (212, 62, 238)
    assert sum(kernel_counts(True, 8, True)) == 80


def test_constant_set_and_order():
    """scripts/ErrorBound.lean:64-83 dump order: R = 2..64 smooth, fwd then bwd,
    twiddled then not, kernel op order."""
    from oracle.symgen import cse_kernel, butterfly, n_inputs

    def smooth(n):
        for p in (2, 3, 5, 7, 11, 13):
            while n % p == 0:
                n //= p
        return n == 1

    seen = []
    for R in range(2, 65):
        if not smooth(R):
            continue
        for fwd in (True, False):
            for tw in (True, False):
                ops, _ = cse_kernel(n_inputs(R, tw), butterfly(fwd, R, tw))
                for op in ops:
                    if op[0] == "cst":
                        key = ("cos", 1, 8) if op[1][0] == "sqrthalf" else op[1]
                        if key not in seen:
                            seen.append(key)
    ref = [tuple(c[:3]) for c in json.load(open(os.path.join(GOLDEN, "kernel_constants.json")))]
    assert seen == ref


def test_error_bound_depths():
    from oracle.stockham import error_bound_terms

    for d in json.load(open(os.path.join(GOLDEN, "bounds.json"))):
        if d["N"] > 70000:
            continue  # the two largest run in test_error_bound_depths_large
        a, b, m = error_bound_terms(d["N"], d["R"])
        assert (a, b) == (d["a"], d["b"]), d
        assert abs(m - d["magMax"]) <= 5e-7 * max(1.0, d["magMax"]), d


@pytest.mark.slow
def test_error_bound_depths_large():
    from oracle.stockham import error_bound_terms

    for d in json.load(open(os.path.join(GOLDEN, "bounds.json"))):
        if d["N"] <= 70000:
            continue
        a, b, m = error_bound_terms(d["N"], d["R"])
        assert (a, b) == (d["a"], d["b"]), d
        assert abs(m - d["magMax"]) <= 5e-7 * max(1.0, d["magMax"]), d


def test_python_and_c_evaluators_agree(oracle):
    """The generated C (gcc -ffp-contract=off) against an exact-rational evaluation of the
    same trees (fma rounded once from the exact value), on every radix."""
    from fractions import Fraction

    from oracle import symgen, trig

    def ev(e, env, memo):
        if e.uid in memo:
            return memo[e.uid]
        o = e.op
        if o == "in":
            v = env[e.a[0]]
        elif o == "zero":
            v = 0.0
        elif o == "cst":
            v = trig.cst_value(e.a[0])
        elif o == "neg":
            v = -ev(e.a[0], env, memo)
        elif o in ("add", "sub", "mul"):
            a, b = ev(e.a[0], env, memo), ev(e.a[1], env, memo)
            v = a + b if o == "add" else a - b if o == "sub" else a * b
        else:
            a, b, c = (ev(x, env, memo) for x in e.a)
            if o == "fms":
                c = -c
            if o == "fnma":
                a = -a
            v = float(Fraction(a) * Fraction(b) + Fraction(c))
        memo[e.uid] = v
        return v

    rng = np.random.default_rng(5)
    for R in (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16):
        for fwd in (True, False):
            for tw in (False, True):
                s = rng.uniform(-1, 1, symgen.n_inputs(R, tw))
                y = oracle.butterfly_eval(fwd, R, tw, s)
                memo = {}
                ref = [ev(e, list(s), memo) for e in symgen.butterfly(fwd, R, tw)]
                assert [bits(v) for v in y] == [bits(v) for v in ref], (R, fwd, tw)


# ---------------------------------------------------------------- schedule


PLANS = [
    (2, [2]), (8, [8]), (16, [4, 4]), (64, [8, 8]), (1024, [8, 4, 4, 8]), (1024, [16, 16, 4]),
    (6, [6]), (60, [6, 10]), (60, [4, 15]), (360, [8, 9, 5]), (1000, [8, 5, 5, 5]),
    (2100, [12, 7, 5, 5]), (10368, [16, 6, 9, 12]), (210, [15, 7, 2]), (143, [11, 13]),
]


@pytest.mark.parametrize("N,rs", PLANS)
def test_closed_forms(oracle, N, rs):
    imp = np.zeros(N, np.complex128)
    imp[0] = 1
    for fwd in (True, False):
        assert np.array_equal(oracle.ref_transform(fwd, rs, imp), np.ones(N))
    const = np.ones(N, np.complex128)
    y = oracle.ref_transform(True, rs, const)
    assert abs(y[0] - N) <= 1e-12 * N and np.max(np.abs(y[1:])) <= 1e-12 * N
    k0 = 3 % N
    tone = np.exp(2j * np.pi * k0 * np.arange(N) / N)
    y = oracle.ref_transform(True, rs, tone)
    ref = np.zeros(N)
    ref[k0] = N
    assert np.max(np.abs(y - ref)) <= 1e-12 * N


@pytest.mark.parametrize("N,rs", PLANS)
def test_against_quad_dft(oracle, N, rs):
    x = oracle.lcg_input(N, seed=oracle.LCG_SEED ^ N, howmany=3)
    for fwd in (True, False):
        y = oracle.ref_transform(fwd, rs, x)
        err = oracle.rel_l2(fwd, x, y)
        assert np.all(err < 3e-16 * max(1.0, math.log2(N)) ** 0.5 + 1e-16), err
        # linearity and the inverse (backward(forward(x)) = N x)
    y = oracle.ref_transform(True, rs, x)
    z = oracle.ref_transform(False, rs, y) / N
    assert np.max(np.abs(z - x)) < 1e-14


def test_quad_dft_vs_mpmath(oracle):
    """The binary128 DFT against a 200-bit brute-force DFT: the prime factors 2, 3, 5, 7,
    11 and 13, one- to four-level recursions (143 = 11 x 13, 210 = 2 x 3 x 5 x 7,
    390 = 2 x 3 x 5 x 13), both directions."""
    import mpmath

    mpmath.mp.prec = 200
    for N, fwd in ((6, True), (16, True), (35, True), (143, True), (143, False), (210, True),
                   (390, False)):
        x = oracle.lcg_input(N, seed=99 + N)
        hi, lo = oracle.mp_dft(fwd, x)
        sgn = -1 if fwd else 1
        w = [mpmath.expjpi(sgn * mpmath.mpf(2 * m) / N) for m in range(N)]
        xs = [mpmath.mpc(v.real, v.imag) for v in x]
        for k in range(N):
            acc = mpmath.fsum(xs[j] * w[(j * k) % N] for j in range(N))
            got = mpmath.mpc(hi[k].real, hi[k].imag) + mpmath.mpc(lo[k].real, lo[k].imag)
            assert abs(got - acc) < mpmath.mpf(2) ** -100, (N, fwd, k)


def test_rel_l2_reads_back_a_known_error(oracle):
    """rel_l2 = ||y - Y||_2 / ||Y||_2 pinned from both sides.

    Closed forms that need no DFT value: the impulse transforms to all ones, so y = ones
    reads 0, y = 0 reads 1, y = -Y reads 2, y = (1 + e) ones reads e, one entry off by d
    reads |d| / sqrt(N). Then y = Y + delta with a chosen ||delta|| / ||Y|| from 1e-6 to
    1e-15 reads back the ratio computed independently at 200 bits from the stored doubles
    (a missing square root, a 1/N, a dropped denominator or a sign error fails)."""
    import mpmath

    for N in (8, 60, 1000):
        imp = np.zeros(N, np.complex128)
        imp[0] = 1
        x = imp[None, :]
        ones = np.ones((1, N), np.complex128)
        assert oracle.rel_l2(True, x, ones)[0] == 0.0
        assert oracle.rel_l2(True, x, np.zeros_like(ones))[0] == 1.0
        assert oracle.rel_l2(True, x, -ones)[0] == 2.0
        e = 2.0 ** -30
        assert oracle.rel_l2(True, x, ones * (1 + e))[0] == e
        y = ones.copy()
        y[0, N // 3] += 0.5j
        assert abs(oracle.rel_l2(False, x, y)[0] - 0.5 / math.sqrt(N)) <= 1e-16
    mpmath.mp.prec = 200
    rng = np.random.default_rng(11)
    for N in (16, 60, 143):
        x = oracle.lcg_input(N, seed=31 + N)
        hi, lo = oracle.mp_dft(True, x)
        for eps in (1e-6, 1e-10, 1e-14, 1e-15):
            d = rng.standard_normal(N) + 1j * rng.standard_normal(N)
            d *= eps * np.linalg.norm(hi) / np.linalg.norm(d)
            y = hi + d
            got = oracle.rel_l2(True, x[None, :], y[None, :])[0]
            num = den = mpmath.mpf(0)
            for k in range(N):
                yr = mpmath.mpf(hi[k].real) + mpmath.mpf(lo[k].real)
                yi = mpmath.mpf(hi[k].imag) + mpmath.mpf(lo[k].imag)
                dr, di = mpmath.mpf(y[k].real) - yr, mpmath.mpf(y[k].imag) - yi
                num += dr * dr + di * di
                den += yr * yr + yi * yi
            want = float(mpmath.sqrt(num / den))
            assert abs(got - want) <= 1e-9 * want, (N, eps, got, want)
            # storing y = hi + delta in binary64 moves it by at most 2^-53 |y| per component
            assert abs(want - eps) <= 0.15 * eps, (N, eps, want)
        x = oracle.lcg_input(N, seed=99 + N)
        hi, lo = oracle.mp_dft(True, x)
        for k in range(N):
            acc = mpmath.mpc(0)
            for j in range(N):
                acc += mpmath.mpc(x[j].real, x[j].imag) * mpmath.exp(-2j * mpmath.pi * j * k / N)
            got = mpmath.mpc(hi[k].real, hi[k].imag) + mpmath.mpc(lo[k].real, lo[k].imag)
            assert abs(got - acc) < mpmath.mpf(2) ** -100


def test_oracle_accuracy_within_reference(oracle):
    """The restated refTransform is as accurate as the reference's published L2
    (forward, max over rounds) -- within 2x at each size checked."""
    acc = json.load(open(os.path.join(GOLDEN, "accuracy_ref.json")))
    cases = [(1024, [8, 4, 4, 8], acc["pow2"]["1024"]["fwd_l2"]),
             (4096, [8, 16, 8, 4], acc["pow2"]["4096"]["fwd_l2"]),
             (1000, [8, 5, 5, 5], acc["xcheck"]["1000"]["lean_l2"]),
             (10368, [16, 6, 9, 12], acc["xcheck"]["10368"]["lean_l2"])]
    for N, rs, ref in cases:
        x = oracle.lcg_input(N, seed=7 ^ N, howmany=4)
        err = oracle.rel_l2(True, x, oracle.ref_transform(True, rs, x)).max()
        assert err <= 2 * ref, (N, err, ref)


def test_lcg_matches_reference_recurrence(oracle):
    M = (1 << 64) - 1
    s = oracle.LCG_SEED
    vals = []
    for _ in range(10):
        s = (s * 6364136223846793005 + 1442695040888963407) & M
        vals.append((s >> 11) / 9007199254740992.0 - 0.5)
    assert np.array_equal(oracle.lcg_input(5).view(np.float64), np.array(vals))
    big = oracle.lcg_input(5000, howmany=3)
    s = oracle.LCG_SEED
    for i in range(2 * 5000 * 3):
        s = (s * 6364136223846793005 + 1442695040888963407) & M
    assert big.view(np.float64).reshape(-1)[-1] == (s >> 11) / 9007199254740992.0 - 0.5


def test_composite_maps_vs_numpy(oracle):
    """The compositions the real-data and 2-D transforms are checked against (embedding,
    first n/2 + 1 outputs, Hermitian extension, real part, axis 0 then axis 1; the generic
    engine's LineOp.apply and alongAxis, Problem.lean:142-182) evaluated with the oracle
    agree with numpy's independent rfft / irfft / fft2 (pocketfft)."""
    rng = np.random.default_rng(4)
    for n, rs in ((8, [8]), (15, [15]), (1000, [8, 5, 5, 5]), (1024, [8, 4, 4, 8])):
        x = rng.standard_normal((3, n))
        half = oracle.ref_transform(True, rs, x.astype(np.complex128))[:, : n // 2 + 1]
        ref = np.fft.rfft(x, axis=1)
        assert np.linalg.norm(half - ref) / np.linalg.norm(ref) < 1e-14
        h = ref + 1j * rng.standard_normal(ref.shape) * (np.arange(ref.shape[1]) == 0)
        k = np.arange(n)
        ext = np.where(k <= n // 2, h[:, np.minimum(k, n // 2)],
                       np.conj(h[:, np.minimum((n - k) % n, n // 2)]))
        back = oracle.ref_transform(False, rs, ext).real
        assert np.linalg.norm(back - n * x) / np.linalg.norm(n * x) < 1e-14
    n0, n1 = 12, 16
    x = rng.standard_normal((2, n0, n1)) + 1j * rng.standard_normal((2, n0, n1))
    cols = np.ascontiguousarray(x.transpose(0, 2, 1)).reshape(-1, n0)
    a = oracle.ref_transform(True, [12], cols).reshape(2, n1, n0).transpose(0, 2, 1)
    b = oracle.ref_transform(True, [16], np.ascontiguousarray(a).reshape(-1, n1)).reshape(x.shape)
    ref = np.fft.fft2(x)
    assert np.linalg.norm(b - ref) / np.linalg.norm(ref) < 1e-14


def test_four_lane_path_matches_scalar(oracle):
    """The CPU baseline's four-lane path (four transforms per call, one per lane of a
    four-double vector, each butterfly's operations applied lane by lane) returns the scalar
    oracle's bits: every radix the plans use (2..16 incl. 11 and 13), twiddled and untwiddled
    passes, both directions, one- to seven-pass plans."""
    rng = np.random.default_rng(11)
    plans = [(2, [2]), (6, [6]), (143, [11, 13]), (1024, [16, 8, 8]), (2048, [16, 8, 16]),
             (4096, [8, 8, 8, 8]), (10368, [12, 8, 12, 9]), (2100, [15, 14, 10]),
             (75600, [10, 9, 3, 10, 7, 4]), (165375, [5, 5, 5, 3, 9, 7, 7])]
    for N, rs in plans:
        for fwd in (True, False):
            x = (rng.random((4, N)) - 0.5) + 1j * (rng.random((4, N)) - 0.5)
            rp = oracle.RefPlan(fwd, N, rs)
            a = np.empty_like(x)
            rp.execute(x, a)
            b = np.full_like(x, np.nan)
            rp.execute_x4(x, b)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (N, rs, fwd)
            c = np.full_like(x, np.nan)
            rp.repeat_x4(x, c, 2)
            assert np.array_equal(a.view(np.uint64), c.view(np.uint64)), (N, rs, fwd)
