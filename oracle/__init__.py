"""The CPU oracle of the FP64 complex FFT hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import, call,
link or execute anything under oracle/, and only as the checker (or the timed CPU
baseline), never as the thing measured or shipped. The product library
(lean-fftw_amd/) has its own twiddle routine, butterflies and kernels and must fail
loudly when its HIP build is missing.

Contents, each restating a reference file (citations in the modules):
  trig.py         exact-integer twiddles          LeanFFTW/Core/Trig.lean, FastPlan.lean
  symgen.py       symbolic butterflies            LeanFFTW/Gen/{SExpr,Sym,SymDFT,Stockham,Emit}.lean
  cgen.py         C emission of those butterflies LeanFFTW/Gen/SExpr.lean:47-52 (cLiteral)
  refstockham.c   refTransform pass loop          LeanFFTW/Gen/Stockham.lean:77-115
  mpfft.c         binary128 DFT (accuracy oracle; replaces libbench2's mp FFT)
  stockham.py     Python drivers, generic-arith Stockham (error-bound depths)

Parity status (what pins each part) is recorded in the module docstrings and the tests.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
import threading
from functools import lru_cache

import numpy as np

from . import trig

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liboracle.so")

_lock = threading.Lock()
_lib = None

# -O3 for the timed CPU baseline; x86-64-v3 (AVX2) rather than -march=native because the
# library is built once and also runs on other hosts; no contraction and no fast-math, so
# every butterfly rounds exactly as written
CFLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math"]


def build(force: bool = False) -> str:
    """Generate the butterflies and compile liboracle.so (gcc, -ffp-contract=off)."""
    from . import cgen

    os.makedirs(BUILD, exist_ok=True)
    gen = os.path.join(BUILD, "bfly_gen.c")
    vgen = os.path.join(BUILD, "vbfly_gen.c")
    for path, src in ((gen, cgen.emit_all()), (vgen, cgen.emit_vector_all())):
        old = open(path).read() if os.path.exists(path) else None
        if old != src:
            with open(path, "w") as f:
                f.write(src)
    srcs = [gen, vgen, os.path.join(HERE, "refstockham.c"), os.path.join(HERE, "vstockham.c"),
            os.path.join(HERE, "mpfft.c")]
    if not force and os.path.exists(LIB):
        mt = os.path.getmtime(LIB)
        if all(os.path.getmtime(s) <= mt for s in srcs):
            return LIB
    tmp = LIB + ".tmp"
    cmd = [
        "gcc", *CFLAGS,
        "-std=gnu11", "-shared", "-fPIC", "-o", tmp, *srcs, "-lquadmath", "-lm",
    ]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB):
                build()
            L = ctypes.CDLL(LIB)
            dp = ctypes.POINTER(ctypes.c_double)
            L.oracle_ref_transform.argtypes = [
                ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp,
                ctypes.c_int64, dp, dp, dp,
            ]
            L.oracle_ref_transform.restype = ctypes.c_int
            L.oracle_ref_repeat.argtypes = L.oracle_ref_transform.argtypes
            L.oracle_ref_repeat.restype = ctypes.c_int
            L.oracle_ref_transform_x4.argtypes = [
                ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp,
                dp, dp, dp,
            ]
            L.oracle_ref_transform_x4.restype = ctypes.c_int
            L.oracle_ref_repeat_x4.argtypes = L.oracle_ref_transform.argtypes
            L.oracle_ref_repeat_x4.restype = ctypes.c_int
            L.oracle_butterfly.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
            L.oracle_butterfly.restype = ctypes.c_int
            L.oracle_mp_dft.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, dp, dp, dp]
            L.oracle_mp_dft.restype = ctypes.c_int
            L.oracle_rel_l2.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, dp, dp, dp]
            L.oracle_rel_l2.restype = ctypes.c_int
            _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ------------------------------------------------------------------------------------
# Twiddle tables (FastPlan.lean:22-47), cached per size
# ------------------------------------------------------------------------------------


@lru_cache(maxsize=64)
def base_table(N: int) -> tuple:
    return tuple(trig.cos_sin_2pi_table(N))


def plan_twiddles(fwd: bool, N: int, rs) -> np.ndarray:
    """planTwiddles (FastPlan.lean:40-47): pass tables of passes 1..s-1 concatenated."""
    tab = base_table(N)
    out: list[float] = []
    for t in range(1, len(rs)):
        out.extend(trig.pass_twiddles(tab, fwd, N, rs[t], trig.pass_ns(rs, t)))
    return np.array(out, dtype=np.float64)


# ------------------------------------------------------------------------------------
# Reference transform
# ------------------------------------------------------------------------------------


def as_interleaved(x) -> np.ndarray:
    """complex128 (..., N) -> contiguous float64 view (..., 2N)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    return x.view(np.float64)


def ref_transform(fwd: bool, rs, x) -> np.ndarray:
    """refTransform fwd N rs on each row of x (complex128, shape (howmany, N) or (N,))."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    shape = x.shape
    N = shape[-1]
    rs = [int(r) for r in rs]
    if math.prod(rs) != N:
        raise ValueError(f"plan {rs} does not multiply to {N}")
    howmany = int(np.prod(shape[:-1])) if len(shape) > 1 else 1
    xin = x.reshape(howmany, N).view(np.float64)
    out = np.empty((howmany, 2 * N), dtype=np.float64)
    work = np.empty(4 * N, dtype=np.float64)
    tw = plan_twiddles(fwd, N, rs)
    if tw.size == 0:
        tw = np.zeros(1)
    rsa = (ctypes.c_int * len(rs))(*rs)
    rc = lib().oracle_ref_transform(
        1 if fwd else 0, N, len(rs), rsa, _dp(tw), howmany, _dp(xin), _dp(out), _dp(work)
    )
    if rc != 0:
        raise ValueError(f"oracle cannot run plan {rs}")
    return out.view(np.complex128).reshape(shape)


def ref_transform_items(args) -> list:
    """[ref_transform(fwd, rs, x) for (rs, x) in items], args = (fwd, items): the unit of
    work of ref_transform_parallel's worker processes."""
    fwd, items = args
    return [ref_transform(fwd, rs, x) for rs, x in items]


def ref_transform_parallel(fwd: bool, items, workers: int = 8) -> list:
    """ref_transform over many (radices, x) pairs in `workers` spawned processes (the
    exact-integer twiddle tables of many distinct sizes dominate the time). Spawned, not
    forked, so that it is safe in a process that has initialised a GPU."""
    import multiprocessing as mp

    items = list(items)
    if workers <= 1 or len(items) < 2:
        return ref_transform_items((fwd, items))
    # group by size so each process builds a size's tables once; largest sizes first
    order = sorted(range(len(items)), key=lambda i: -len(items[i][1]))
    chunks = [[] for _ in range(workers)]
    load = [0] * workers
    for i in order:
        w = load.index(min(load))
        chunks[w].append(i)
        load[w] += len(items[i][1])
    from concurrent.futures import ProcessPoolExecutor

    # an executor, not a Pool: a worker that dies raises BrokenProcessPool instead of being
    # replaced forever
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        parts = list(ex.map(ref_transform_items, [(fwd, [items[i] for i in c]) for c in chunks]))
    out = [None] * len(items)
    for c, part in zip(chunks, parts):
        for i, y in zip(c, part):
            out[i] = y
    return out


def butterfly_eval(fwd: bool, R: int, tw: bool, s: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(s, dtype=np.float64)
    y = np.empty(2 * R, dtype=np.float64)
    if lib().oracle_butterfly(1 if fwd else 0, R, 1 if tw else 0, _dp(s), _dp(y)) != 0:
        raise ValueError(R)
    return y


# ------------------------------------------------------------------------------------
# Accuracy oracle (binary128)
# ------------------------------------------------------------------------------------


def mp_dft(fwd: bool, x) -> tuple[np.ndarray, np.ndarray]:
    """Binary128 DFT of each row, as (hi, lo) complex128 arrays with hi + lo ~ exact."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    shape = x.shape
    N = shape[-1]
    howmany = int(np.prod(shape[:-1])) if len(shape) > 1 else 1
    xin = x.reshape(howmany, N).view(np.float64)
    hi = np.empty((howmany, 2 * N))
    lo = np.empty((howmany, 2 * N))
    if lib().oracle_mp_dft(-1 if fwd else 1, N, howmany, _dp(xin), _dp(hi), _dp(lo)) != 0:
        raise MemoryError
    return hi.view(np.complex128).reshape(shape), lo.view(np.complex128).reshape(shape)


def rel_l2(fwd: bool, x, y) -> np.ndarray:
    """Relative L2 error of each row of y against the binary128 DFT of x."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.ascontiguousarray(y, dtype=np.complex128)
    N = x.shape[-1]
    howmany = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    err = np.empty(howmany)
    rc = lib().oracle_rel_l2(
        -1 if fwd else 1, N, howmany, _dp(x.reshape(howmany, N).view(np.float64)),
        _dp(y.reshape(howmany, N).view(np.float64)), _dp(err),
    )
    if rc != 0:
        raise MemoryError
    return err


# ------------------------------------------------------------------------------------
# Inputs: the reference's LCG (scripts/NativeBench.lean:31-37)
# ------------------------------------------------------------------------------------

LCG_SEED = 0x9E3779B97F4A7C15
_MUL = 6364136223846793005
_INC = 1442695040888963407


def lcg_input(N: int, seed: int = LCG_SEED, howmany: int = 1) -> np.ndarray:
    """2*N*howmany values (s >> 11) / 2^53 - 0.5 of the LCG
    s <- s*6364136223846793005 + 1442695040888963407 mod 2^64, read as complex values
    (re, im interleaved), shape (N,) or (howmany, N). Vectorised by jumping ahead."""
    total = 2 * N * howmany
    M = (1 << 64) - 1
    B = min(total, 4096)
    head = np.empty(B, dtype=np.uint64)
    s = seed & M
    for i in range(B):
        s = (s * _MUL + _INC) & M
        head[i] = s
    # affine map of B steps: s -> A s + C
    A, C = 1, 0
    for _ in range(B):
        A, C = (A * _MUL) & M, (C * _MUL + _INC) & M
    nblk = -(-total // B)
    vals = np.empty(nblk * B, dtype=np.uint64)
    cur = head
    Au, Cu = np.uint64(A), np.uint64(C)
    with np.errstate(over="ignore"):
        for b in range(nblk):
            vals[b * B:(b + 1) * B] = cur
            cur = cur * Au + Cu
    out = (vals[:total] >> np.uint64(11)).astype(np.float64) / 9007199254740992.0 - 0.5
    c = out.view(np.complex128)
    return c if howmany == 1 else c.reshape(howmany, N)


class RefPlan:
    """refTransform with precomputed tables: execute() runs only the compiled pass loop
    (for timing the oracle as the CPU baseline; tables are planning, as in the reference)."""

    def __init__(self, fwd: bool, N: int, rs):
        self.fwd, self.N, self.rs = bool(fwd), int(N), [int(r) for r in rs]
        if math.prod(self.rs) != self.N:
            raise ValueError("plan does not multiply to N")
        tw = plan_twiddles(self.fwd, self.N, self.rs)
        self.tw = tw if tw.size else np.zeros(1)
        self.rsa = (ctypes.c_int * len(self.rs))(*self.rs)
        self.work = np.empty(4 * self.N)

    def repeat(self, x: np.ndarray, out: np.ndarray, reps: int) -> None:
        """The one-transform plan `reps` times on the same x (shape (N,)) into out."""
        rc = lib().oracle_ref_repeat(
            1 if self.fwd else 0, self.N, len(self.rs), self.rsa, _dp(self.tw), int(reps),
            _dp(x.reshape(-1).view(np.float64)), _dp(out.reshape(-1).view(np.float64)),
            _dp(self.work),
        )
        if rc != 0:
            raise ValueError("oracle cannot run this plan")

    def _work_x4(self) -> np.ndarray:
        """24 N doubles on a 32-byte boundary (the four-lane path's vectors)."""
        if getattr(self, "_w4", None) is None:
            raw = np.empty(24 * self.N + 4)
            off = (-raw.ctypes.data % 32) // 8
            self._w4 = raw[off:off + 24 * self.N]
        return self._w4

    def execute_x4(self, x: np.ndarray, out: np.ndarray) -> None:
        """The four-lane path (one transform per lane of a four-double vector): x, out
        complex128 of shape (4, N), C-contiguous; the scalar path's bits."""
        rc = lib().oracle_ref_transform_x4(
            1 if self.fwd else 0, self.N, len(self.rs), self.rsa, _dp(self.tw),
            _dp(x.reshape(-1).view(np.float64)), _dp(out.reshape(-1).view(np.float64)),
            _dp(self._work_x4()),
        )
        if rc != 0:
            raise ValueError("oracle cannot run this plan")

    def repeat_x4(self, x: np.ndarray, out: np.ndarray, reps: int) -> None:
        """The four-transform call `reps` times on the same x (shape (4, N)) into out."""
        rc = lib().oracle_ref_repeat_x4(
            1 if self.fwd else 0, self.N, len(self.rs), self.rsa, _dp(self.tw), int(reps),
            _dp(x.reshape(-1).view(np.float64)), _dp(out.reshape(-1).view(np.float64)),
            _dp(self._work_x4()),
        )
        if rc != 0:
            raise ValueError("oracle cannot run this plan")

    def execute(self, x: np.ndarray, out: np.ndarray) -> None:
        """x, out: complex128 arrays of shape (howmany, N), C-contiguous."""
        howmany = x.size // self.N
        rc = lib().oracle_ref_transform(
            1 if self.fwd else 0, self.N, len(self.rs), self.rsa, _dp(self.tw), howmany,
            _dp(x.reshape(-1).view(np.float64)), _dp(out.reshape(-1).view(np.float64)),
            _dp(self.work),
        )
        if rc != 0:
            raise ValueError("oracle cannot run this plan")
