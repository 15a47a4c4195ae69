"""mfft: MI355X (gfx950) FP64 complex FFT -- Python binding of the C ABI (include/mfft.h).

The shipped compute path is libmfft.so (HIP kernels for gfx950, built in-tree by
build.py). This module loads it with ctypes and wraps it for torch tensors; torch is
used only for device memory and streams. There is no CPU fallback: without the built
library every entry point raises.

Mirrors the reference's plan/execute surface:
  plan_measure(N, fwd)           <- lfa_plan_measure (LeanFFTW/Asm/Planner.lean:291-302)
  Plan.call(inp, out)            <- lfa_call (LeanFFTW/Asm/Kernels.lean:1203-1204)
  Plan.exec(inp) / exec_into     <- Native.exec / execInto (LeanFFTW/Native/Exec.lean:90-95)
  plan_problem(spec) / exec_spec <- lfftw_generic_plan / exec (LeanFFTW/FFI.lean:50-58)
  set_threads(n)                 <- lfa_set_threads (LeanFFTW/Asm/Planner.lean:305-306)
"""

from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmfft.so")

FORWARD = -1
BACKWARD = 1
ESTIMATE = 0
MEASURE = 1

_lock = threading.Lock()
_lib = None


class MfftError(RuntimeError):
    pass


def library_path() -> str:
    return LIB_PATH


def lib():
    """Load libmfft.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise MfftError(
                    f"{LIB_PATH} is missing: build it with `python lean-fftw_amd/build.py` "
                    "(there is no CPU fallback)"
                )
            L = ctypes.CDLL(LIB_PATH)
            vp, i64, dp = ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)
            sig = {
                "mfft_plan_many": (vp, [i64, ctypes.c_int, i64, i64, i64, i64, i64, ctypes.c_uint]),
                "mfft_plan_dft_1d": (vp, [i64, ctypes.c_int, i64, ctypes.c_uint]),
                "mfft_plan_spec": (vp, [dp, ctypes.c_size_t, ctypes.c_uint]),
                "mfft_plan_r2c": (vp, [i64, i64, ctypes.c_uint]),
                "mfft_plan_c2r": (vp, [i64, i64, ctypes.c_uint]),
                "mfft_plan_dft_2d": (vp, [i64, i64, ctypes.c_int, i64, ctypes.c_uint]),
                "mfft_plan_kind": (ctypes.c_int, [vp]),
                "mfft_plan_with_radices": (vp, [i64, ctypes.c_int, i64, ctypes.POINTER(ctypes.c_int),
                                                ctypes.c_int, ctypes.c_int]),
                "mfft_plan_with_stages": (vp, [i64, ctypes.c_int, i64, ctypes.POINTER(ctypes.c_int),
                                               ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
                "mfft_execute": (ctypes.c_int, [vp, vp, vp, vp]),
                "mfft_execute_host": (ctypes.c_int, [vp, dp, dp]),
                "mfft_execute_static": (ctypes.c_int, [vp, vp, vp, vp]),
                "mfft_execute_spec": (ctypes.c_int, [vp, vp, vp, vp]),
                "mfft_execute_weighted": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
                "mfft_execute_weighted_sumsq": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp]),
                "mfft_execute_weighted_scaled": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_double,
                                                                vp, vp]),
                "mfft_plan_partials": (i64, [vp]),
                "mfft_destroy_plan": (None, [vp]),
                "mfft_plan_radices": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
                "mfft_plan_stages": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
                "mfft_plan_schedule": (ctypes.c_int, [i64, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                                      ctypes.POINTER(ctypes.c_int)]),
                "mfft_plan_ragged": (vp, [i64, ctypes.POINTER(ctypes.c_int64), ctypes.c_int, ctypes.c_uint]),
                "mfft_execute_ragged": (ctypes.c_int, [vp, vp, vp, vp]),
                "mfft_ragged_status": (ctypes.c_int, [vp, vp]),
                "mfft_ragged_class_plan": (ctypes.c_int, [vp, i64, ctypes.POINTER(ctypes.c_int),
                                                          ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
                "mfft_ragged_queue": (i64, [i64, ctypes.POINTER(ctypes.c_int64), i64,
                                            ctypes.POINTER(ctypes.c_int64), i64,
                                            ctypes.POINTER(ctypes.c_int64), i64,
                                            ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_int64), i64]),
                "mfft_destroy_ragged": (None, [vp]),
                "mfft_cos_sin_2pi": (None, [ctypes.c_uint64, ctypes.c_uint64, dp, dp]),
                "mfft_tan_2pi": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64]),
                "mfft_pass_twiddles": (ctypes.c_int, [ctypes.c_int, i64, ctypes.c_int, i64, dp]),
                "mfft_stockham_cost": (i64, [ctypes.c_int, i64, ctypes.POINTER(ctypes.c_int),
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
                "mfft_schedule_cost": (ctypes.c_double, [ctypes.c_int, i64, ctypes.POINTER(ctypes.c_int),
                                                         ctypes.c_int, ctypes.c_int]),
                "mfft_set_threads": (ctypes.c_int, [ctypes.c_size_t]),
                "mfft_set_wisdom_path": (ctypes.c_int, [ctypes.c_char_p]),
                "mfft_error_string": (ctypes.c_char_p, [ctypes.c_int]),
                "mfft_version": (ctypes.c_char_p, []),
            }
            for name, (res, args) in sig.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise MfftError(f"{what}: {lib().mfft_error_string(rc).decode()} ({rc})")


def _stream_ptr(stream):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


# Plans released while a HIP graph is being captured: their device memory is freed at the
# next release outside a capture (hipFree is not allowed during a capture).
_deferred: list = []


def _capturing() -> bool:
    try:
        import torch

        return bool(torch.cuda.is_available() and torch.cuda.is_current_stream_capturing())
    except Exception:  # pragma: no cover - interpreter shutdown
        return False


def _release(destroy, h) -> None:
    if _capturing():
        _deferred.append((destroy, h))
        return
    while _deferred:
        d, old = _deferred.pop()
        d(old)
    destroy(h)


# ------------------------------------------------------------------------------ twiddles


def cos_sin_2pi(k: int, n: int) -> tuple[float, float]:
    c, s = ctypes.c_double(), ctypes.c_double()
    lib().mfft_cos_sin_2pi(k, n, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


def tan_2pi(num: int, den: int) -> float:
    return lib().mfft_tan_2pi(num, den)


def pass_twiddles(fwd: bool, N: int, R: int, ns: int):
    import numpy as np

    out = np.empty(2 * (R - 1) * ns, dtype=np.float64)
    _check(lib().mfft_pass_twiddles(1 if fwd else 0, N, R, ns,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
           "mfft_pass_twiddles")
    return out


def schedule(n: int):
    """(radices, split) the planner chooses for size n (host only, allocates nothing)."""
    buf = (ctypes.c_int * 64)()
    split = ctypes.c_int()
    k = lib().mfft_plan_schedule(n, buf, 64, ctypes.byref(split))
    if k < 0:
        raise MfftError(f"unsupported size {n}")
    return list(buf[:k]), split.value


def stockham_cost(fwd: bool, n: int, radices, weights) -> int:
    """The reference's operation count of a schedule (stockhamCost) with weights
    (add, mul, fma, load, store)."""
    rs = (ctypes.c_int * len(radices))(*radices)
    w = (ctypes.c_int64 * 5)(*weights)
    v = lib().mfft_stockham_cost(1 if fwd else 0, n, rs, len(radices), w)
    if v < 0:
        raise MfftError(f"invalid schedule {radices} for {n}")
    return int(v)


def schedule_cost(fwd: bool, n: int, radices, split: int = 0) -> float:
    """The measuring planner's modelled GPU time of a schedule (ns per transform)."""
    rs = (ctypes.c_int * len(radices))(*radices)
    v = lib().mfft_schedule_cost(1 if fwd else 0, n, rs, len(radices), split)
    if v < 0:
        raise MfftError(f"invalid schedule {radices} / {split} for {n}")
    return float(v)


def set_threads(n: int) -> None:
    _check(lib().mfft_set_threads(n), "mfft_set_threads")


def set_wisdom_path(path) -> None:
    """JSON-lines file where MEASURE plans look up and append their timed schedules.
    None or "" falls back to the MFFT_WISDOM environment variable (no cache when unset)."""
    _check(lib().mfft_set_wisdom_path((path or "").encode()), "mfft_set_wisdom_path")


# ------------------------------------------------------------------------------ plans


class Plan:
    """A batched 1-D complex FP64 transform on the current device.

    Plan(n, sign=FORWARD, howmany=1, radices=None, split=0, istride=1, idist=None,
    ostride=1, odist=None, flags=ESTIMATE, passes=None). `radices` forces the Stockham radix
    sequence (the reference's `rs`); `split` the number of passes of the first kernel stage, or
    `passes` the passes of each of up to three stages.
    """

    def __init__(self, n, sign=FORWARD, howmany=1, radices=None, split=0, istride=1,
                 idist=None, ostride=1, odist=None, flags=ESTIMATE, _handle=None, passes=None):
        L = lib()
        self.n, self.sign, self.howmany = int(n), int(sign), int(howmany)
        self.istride, self.ostride = int(istride), int(ostride)
        self.idist = int(n if idist is None else idist)
        self.odist = int(n if odist is None else odist)
        if _handle is not None:
            h = _handle
        elif radices is not None:
            if (self.istride, self.ostride, self.idist, self.odist) != (1, 1, self.n, self.n):
                raise ValueError("forced radices are for contiguous batches")
            arr = (ctypes.c_int * len(radices))(*radices)
            if passes is not None:
                pa = (ctypes.c_int * len(passes))(*passes)
                h = L.mfft_plan_with_stages(self.n, self.sign, self.howmany, arr, len(radices), pa,
                                            len(passes))
            else:
                h = L.mfft_plan_with_radices(self.n, self.sign, self.howmany, arr, len(radices),
                                             int(split))
        else:
            h = L.mfft_plan_many(self.n, self.sign, self.howmany, self.istride, self.idist,
                                 self.ostride, self.odist, flags)
        if not h:
            raise MfftError(f"unsupported problem: n={n} sign={sign} howmany={howmany} radices={radices}")
        self._h = ctypes.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _release(_lib.mfft_destroy_plan, h)
            self._h = None

    @property
    def radices(self) -> list[int]:
        buf = (ctypes.c_int * 32)()
        n = lib().mfft_plan_radices(self._h, buf, 32)
        return list(buf[:n])

    @property
    def stages(self) -> list[int]:
        buf = (ctypes.c_int * 8)()
        n = lib().mfft_plan_stages(self._h, buf, 8)
        return list(buf[:n])

    def _check_tensor(self, t, what):
        import torch

        if not isinstance(t, torch.Tensor) or t.dtype != torch.complex128 or not t.is_cuda:
            raise TypeError(f"{what} must be a complex128 CUDA tensor")
        span = (self.n - 1) * (self.istride if what == "input" else self.ostride) + (
            self.howmany - 1) * (self.idist if what == "input" else self.odist) + 1
        if t.numel() < span:
            raise ValueError(f"{what} holds {t.numel()} elements, the plan spans {span}")

    def call(self, inp, out, stream=None):
        """out <- transform(inp) on device tensors (inp may be out); asynchronous."""
        self._check_tensor(inp, "input")
        self._check_tensor(out, "output")
        _check(lib().mfft_execute(self._h, ctypes.c_void_p(inp.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)),
               "mfft_execute")
        return out

    def call_static(self, inp, out, stream=None):
        """As call, with the library's static device scratch as the two-stage intermediate
        (mfft_execute_static: allocates nothing; calls overlapping on one GPU race)."""
        self._check_tensor(inp, "input")
        self._check_tensor(out, "output")
        _check(lib().mfft_execute_static(self._h, ctypes.c_void_p(inp.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)),
               "mfft_execute_static")
        return out

    def call_weighted(self, inp, out, pre=None, post=None, stream=None):
        """out <- diag(post) F diag(pre) inp, with real float64 weights of length n."""
        self._check_tensor(inp, "input")
        self._check_tensor(out, "output")
        import torch

        for w in (pre, post):
            if w is not None and (w.dtype != torch.float64 or not w.is_cuda or w.numel() < self.n):
                raise TypeError("weights must be float64 CUDA tensors of length n")
        ptr = lambda w: ctypes.c_void_p(w.data_ptr()) if w is not None else None  # noqa: E731
        _check(lib().mfft_execute_weighted(self._h, ctypes.c_void_p(inp.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), ptr(pre), ptr(post),
                                           _stream_ptr(stream)), "mfft_execute_weighted")
        return out

    def call_weighted_scaled(self, inp, out, pre, scale, factor=1.0, post=None, stream=None):
        """out <- diag(post) F diag((factor * scale[0]) * pre) inp: the input weights times a
        scalar that stays on the device (a float64 CUDA tensor of one element), the product
        formed in the kernel."""
        self._check_tensor(inp, "input")
        self._check_tensor(out, "output")
        import torch

        for w in (pre, post):
            if w is not None and (w.dtype != torch.float64 or not w.is_cuda or w.numel() < self.n):
                raise TypeError("weights must be float64 CUDA tensors of length n")
        if pre is None:
            raise TypeError("call_weighted_scaled needs input weights")
        if scale.dtype != torch.float64 or not scale.is_cuda or scale.numel() != 1:
            raise TypeError("scale must be a float64 CUDA tensor of one element")
        ptr = lambda w: ctypes.c_void_p(w.data_ptr()) if w is not None else None  # noqa: E731
        _check(lib().mfft_execute_weighted_scaled(self._h, ctypes.c_void_p(inp.data_ptr()),
                                                  ctypes.c_void_p(out.data_ptr()), ptr(pre),
                                                  ptr(scale), float(factor), ptr(post),
                                                  _stream_ptr(stream)),
               "mfft_execute_weighted_scaled")
        return out

    @property
    def partials(self) -> int:
        return int(lib().mfft_plan_partials(self._h))

    def call_weighted_sumsq(self, inp, out, partials, pre=None, post=None, stream=None):
        """As call_weighted; partials (float64, length self.partials) receives per-workgroup
        sums of |out|^2, whose total is sum |out|^2."""
        import torch

        self._check_tensor(inp, "input")
        self._check_tensor(out, "output")
        if partials.dtype != torch.float64 or not partials.is_cuda or partials.numel() < self.partials:
            raise TypeError("partials must be a float64 CUDA tensor of length plan.partials")
        ptr = lambda w: ctypes.c_void_p(w.data_ptr()) if w is not None else None  # noqa: E731
        _check(lib().mfft_execute_weighted_sumsq(self._h, ctypes.c_void_p(inp.data_ptr()),
                                                 ctypes.c_void_p(out.data_ptr()), ptr(pre),
                                                 ptr(post), ctypes.c_void_p(partials.data_ptr()),
                                                 _stream_ptr(stream)), "mfft_execute_weighted_sumsq")
        return out

    def exec(self, inp, stream=None):
        import torch

        out = torch.empty_like(inp)
        return self.call(inp, out, stream)

    exec_into = call

    def execute_host(self, inp, out):
        """Host numpy arrays (complex128): copy in, transform, copy out."""
        import numpy as np

        assert inp.dtype == np.complex128 and out.dtype == np.complex128
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib().mfft_execute_host(self._h, inp.ctypes.data_as(dp), out.ctypes.data_as(dp)),
               "mfft_execute_host")
        return out


def plan_measure(N: int, fwd: bool, howmany: int = 1) -> Plan:
    """lfa_plan_measure without touching user arrays: a plan of N points, one direction."""
    return Plan(N, FORWARD if fwd else BACKWARD, howmany, flags=MEASURE)


class SpecPlan(Plan):
    """A plan from the reference's generic problem spec (LeanFFTW/FFI.lean:13-21)."""

    def __init__(self, spec, flags=ESTIMATE):
        import numpy as np

        arr = np.ascontiguousarray(spec, dtype=np.float64)
        h = lib().mfft_plan_spec(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), arr.size, flags)
        if not h:
            raise MfftError("unsupported problem spec")
        self._h = ctypes.c_void_p(h)
        self.n = int(spec[8])
        self.sign = -1 if spec[1] < 0 else 1

    def call(self, inp_doubles, out_doubles, stream=None):
        _check(lib().mfft_execute_spec(self._h, ctypes.c_void_p(inp_doubles.data_ptr()),
                                       ctypes.c_void_p(out_doubles.data_ptr()), _stream_ptr(stream)),
               "mfft_execute_spec")
        return out_doubles


def plan_problem(spec) -> SpecPlan | None:
    try:
        return SpecPlan(spec)
    except MfftError:
        return None


class _CompositePlan:
    """Shared handle logic of the composite plans (r2c, c2r, 2-D)."""

    _h = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _release(_lib.mfft_destroy_plan, h)
            self._h = None

    @property
    def radices(self) -> list[int]:
        buf = (ctypes.c_int * 32)()
        n = lib().mfft_plan_radices(self._h, buf, 32)
        return list(buf[:n])

    @staticmethod
    def _check(t, dtype, numel, what):
        if t.dtype != dtype or not t.is_cuda or not t.is_contiguous():
            raise TypeError(f"{what} must be a contiguous {dtype} CUDA tensor")
        if t.numel() < numel:
            raise ValueError(f"{what} holds {t.numel()} elements, the plan needs {numel}")

    def _exec(self, inp, out, stream):
        _check(lib().mfft_execute(self._h, ctypes.c_void_p(inp.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)),
               "mfft_execute")
        return out


class RealPlan(_CompositePlan):
    """Batched real-data transforms of n points (the reference's rdft2, rank 1):
    forward (r2c) float64 [howmany, n] -> complex128 [howmany, n//2 + 1]; backward (c2r,
    unnormalised) complex128 [howmany, n//2 + 1] -> float64 [howmany, n]. Input and output
    are distinct tensors."""

    def __init__(self, n, sign=FORWARD, howmany=1, flags=ESTIMATE):
        L = lib()
        self.n, self.sign, self.howmany = int(n), int(sign), int(howmany)
        self.nh = self.n // 2 + 1
        f = L.mfft_plan_r2c if self.sign == FORWARD else L.mfft_plan_c2r
        h = f(self.n, self.howmany, flags)
        if not h:
            raise MfftError(f"unsupported real problem: n={n} sign={sign} howmany={howmany}")
        self._h = ctypes.c_void_p(h)

    def call(self, inp, out, stream=None):
        import torch

        if inp.data_ptr() == out.data_ptr():
            raise ValueError("real transforms are out of place")
        fwd = self.sign == FORWARD
        self._check(inp, torch.float64 if fwd else torch.complex128,
                    self.howmany * (self.n if fwd else self.nh), "input")
        self._check(out, torch.complex128 if fwd else torch.float64,
                    self.howmany * (self.nh if fwd else self.n), "output")
        return self._exec(inp, out, stream)

    def exec(self, inp, stream=None):
        import torch

        if self.sign == FORWARD:
            out = torch.empty((self.howmany, self.nh), dtype=torch.complex128, device=inp.device)
        else:
            out = torch.empty((self.howmany, self.n), dtype=torch.float64, device=inp.device)
        return self.call(inp, out, stream)


class Plan2D(_CompositePlan):
    """2-D complex transforms of howmany contiguous row-major n0 x n1 complex128 arrays:
    the DFT along axis 0, then along axis 1 (inp may be out)."""

    def __init__(self, n0, n1, sign=FORWARD, howmany=1, flags=ESTIMATE):
        self.n0, self.n1, self.sign, self.howmany = int(n0), int(n1), int(sign), int(howmany)
        h = lib().mfft_plan_dft_2d(self.n0, self.n1, self.sign, self.howmany, flags)
        if not h:
            raise MfftError(f"unsupported 2-D problem: {n0} x {n1} sign={sign} howmany={howmany}")
        self._h = ctypes.c_void_p(h)

    def call(self, inp, out, stream=None):
        import torch

        size = self.howmany * self.n0 * self.n1
        self._check(inp, torch.complex128, size, "input")
        self._check(out, torch.complex128, size, "output")
        return self._exec(inp, out, stream)

    def exec(self, inp, stream=None):
        import torch

        return self.call(inp, torch.empty_like(inp), stream)


class RaggedPlan:
    """Transforms of mixed sizes packed back to back: per call, two persistent launches (every
    stage-A and single-stage tile, then every stage-B tile) from size-sorted queues."""

    def __init__(self, sizes, sign=FORWARD, flags=ESTIMATE):
        import numpy as np

        self.sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        self.total = int(self.sizes.sum())
        self.sign = sign
        h = lib().mfft_plan_ragged(len(self.sizes), self.sizes.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int64)), sign, flags)
        if not h:
            raise MfftError("unsupported ragged batch")
        self._h = ctypes.c_void_p(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _release(_lib.mfft_destroy_ragged, h)
            self._h = None

    def call(self, inp, out, stream=None):
        import torch

        for t in (inp, out):
            if t.dtype != torch.complex128 or not t.is_cuda or t.numel() < self.total:
                raise TypeError("ragged tensors must be complex128 CUDA tensors of the total size")
        _check(lib().mfft_execute_ragged(self._h, ctypes.c_void_p(inp.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), _stream_ptr(stream)),
               "mfft_execute_ragged")
        return out

    def status(self, stream=None):
        """Synchronise the stream; raises if the launches failed."""
        _check(lib().mfft_ragged_status(self._h, _stream_ptr(stream)), "ragged launch")

    def class_plan(self, n: int):
        """(radices, split) of the schedule used for the batch's transforms of size n."""
        buf = (ctypes.c_int * 64)()
        split = ctypes.c_int()
        k = lib().mfft_ragged_class_plan(self._h, n, buf, 64, ctypes.byref(split))
        if k < 0:
            raise KeyError(n)
        return list(buf[:k]), split.value


# ------------------------------------------------------------------------------ torch-level API

_plan_cache: dict = {}


def stream_key():
    """(device, stream) of the current stream: plan caches are per stream, since a plan's
    intermediate must not be shared by executions that overlap on two streams."""
    import torch

    s = torch.cuda.current_stream()
    return (s.device.index, s.cuda_stream)


def _cached_plan(n, sign, howmany, cache=None):
    """The plan of (n, sign, howmany) for the current stream, created on first use.

    During a HIP graph capture no plan is created: making one uploads its tables with
    synchronous copies, which a capture forbids. The capture takes the plan made for this
    shape before it on another stream of the same device (torch captures on a stream of its
    own), so the graph's replays share that plan's intermediate with its executions on that
    stream and must not overlap them; with no such plan it raises MfftError before anything
    is recorded."""
    cache = _plan_cache if cache is None else cache
    sk = stream_key()
    key = (n, sign, howmany, sk)
    p = cache.get(key)
    if p is not None:
        return p
    if _capturing():
        for (kn, ks, kh, (dev, _)), q in cache.items():
            if (kn, ks, kh, dev) == (n, sign, howmany, sk[0]):
                return q
        raise MfftError(f"no plan of {n} points x {howmany} (sign {sign}) exists on device "
                        f"{sk[0]}: plans cannot be created during a HIP graph capture, so run "
                        "the step once before capturing it")
    p = Plan(n, sign, howmany)
    cache[key] = p
    return p


def fft(x, sign: int = FORWARD):
    """Unnormalised transform along the last dimension of a complex128 CUDA tensor."""
    x = x.contiguous()
    n = x.shape[-1]
    howmany = x.numel() // n if n else 0
    out = x.new_empty(x.shape)
    if howmany:
        _cached_plan(n, sign, howmany).call(x, out)
    return out


def ifft_unnormalized(x):
    return fft(x, BACKWARD)


def ragged_queue(sizes, shapes, nqueue=1):
    """Host only: the work queues of a ragged batch's two launches (mfft_ragged_queue).
    `shapes` lists (n, stages, P, Q, G_A, G_B) per class in ascending n. Returns
    (tiles, n_first, qoff, offs): tiles an (ntiles, 5) int64 array of
    (class, stage, count, base, sub0), the first n_first of them the first launch's; queue q
    of the first launch is tiles[qoff[q]:qoff[q + 1]], of the second
    tiles[qoff[nqueue + 1 + q]:qoff[nqueue + 2 + q]]."""
    import numpy as np

    sz = np.ascontiguousarray(sizes, dtype=np.int64)
    sh = np.ascontiguousarray(shapes, dtype=np.int64).reshape(-1, 6)
    p64 = ctypes.POINTER(ctypes.c_int64)
    nf = ctypes.c_int64()
    n = lib().mfft_ragged_queue(len(sz), sz.ctypes.data_as(p64), len(sh), sh.ctypes.data_as(p64),
                                nqueue, None, 0, ctypes.byref(nf), None, None, 0)
    if n < 0:
        raise MfftError("inconsistent ragged queue description")
    tiles = np.zeros((max(n, 1), 5), dtype=np.int64)
    qoff = np.zeros(2 * (nqueue + 1), dtype=np.int64)
    offs = np.zeros(max(len(sz), 1), dtype=np.int64)
    lib().mfft_ragged_queue(len(sz), sz.ctypes.data_as(p64), len(sh), sh.ctypes.data_as(p64),
                            nqueue, tiles.ctypes.data_as(p64), n, ctypes.byref(nf),
                            qoff.ctypes.data_as(p64), offs.ctypes.data_as(p64), len(sz))
    return tiles[:n], int(nf.value), qoff, offs


from . import workloads  # noqa: E402,F401
from .autograd import FFTFunction, WeightedPowerLoss, fft_autograd, weighted_power_loss  # noqa: E402,F401
