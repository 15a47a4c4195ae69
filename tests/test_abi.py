"""CPU checks of the shipped library: it loads, exports every entry point include/mfft.h
declares, and its host-side twiddle routine and device-code generator agree bit for bit
with the oracle. No GPU compute is called here."""

import ctypes
import importlib.util
import os
import re
import struct

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mfft.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mfft_\w+)\s*\(", text)))


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def test_library_exports_every_declared_symbol(lf):
    L = lf.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_gfx950_hip():
    path = os.path.join(ROOT, "lean-fftw_amd", "_build", "libmfft.so")
    data = open(path, "rb").read()
    assert b"gfx950" in data
    assert b"stage_kernel" in data and b"ragged_kernel" in data


def test_host_twiddles_match_oracle(lf):
    from oracle import trig

    for n in (1, 2, 3, 8, 12, 60, 1000, 2100, 65536, 165375):
        for k in sorted({0, 1, n // 8, n // 4, n // 3, n // 2, n - 1, 7 % n, (5 * n) // 7}):
            assert lf.cos_sin_2pi(k, n) == trig.cos_sin_2pi(k, n), (k, n)
    for num, den in ((1, 36), (2, 36), (16, 400), (1, 9), (3, 52)):  # num/den <= 1/8
        assert bits(lf.tan_2pi(num, den)) == bits(trig.tan_2pi(num, den))


@pytest.mark.parametrize("N,R,ns", [(16, 4, 4), (1024, 8, 128), (360, 5, 8), (2100, 7, 300),
                                    (165375, 7, 23625)])
def test_pass_tables_match_oracle(lf, N, R, ns):
    from oracle import trig

    tab = trig.cos_sin_2pi_table(N)
    for fwd in (True, False):
        ref = np.array(trig.pass_twiddles(tab, fwd, N, R, ns))
        got = lf.pass_twiddles(fwd, N, R, ns)
        assert np.array_equal(ref.view(np.uint64), got.view(np.uint64))


def test_codegen_matches_oracle_dags():
    """The product's generated butterflies (lean-fftw_amd/codegen.py) compute the same
    binary64 DAG as the oracle: exact evaluation of both on random inputs, every radix."""
    from fractions import Fraction

    import oracle

    spec = importlib.util.spec_from_file_location("mfft_codegen",
                                                  os.path.join(ROOT, "lean-fftw_amd", "codegen.py"))
    cg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cg)

    def ev(d, outs, s):
        v = {}
        for i, k in enumerate(d.op):
            o = k[0]
            if o == "in":
                r = s[k[1]]
            elif o == "zero":
                r = 0.0
            elif o == "cst":
                r = cg.const_value(k[1])
            elif o == "neg":
                r = -v[k[1]]
            elif o == "add":
                r = v[k[1]] + v[k[2]]
            elif o == "sub":
                r = v[k[1]] - v[k[2]]
            elif o == "mul":
                r = v[k[1]] * v[k[2]]
            else:
                a, b, c = v[k[1]], v[k[2]], v[k[3]]
                c = -c if o == "fms" else c
                a = -a if o == "fnma" else a
                r = float(Fraction(a) * Fraction(b) + Fraction(c))
            v[i] = r
        return [v[o] for o in outs]

    rng = np.random.default_rng(3)
    for R in cg.RADICES:
        for fwd in (True, False):
            for tw in (False, True):
                d, outs = cg.butterfly_dag(fwd, R, tw)
                s = rng.uniform(-1, 1, 4 * R - 2 if tw else 2 * R)
                got = ev(d, outs, list(s))
                ref = oracle.butterfly_eval(fwd, R, tw, s)
                assert [bits(x) for x in got] == [bits(x) for x in ref], (R, fwd, tw)


def test_generated_header_is_current():
    spec = importlib.util.spec_from_file_location("mfft_codegen2",
                                                  os.path.join(ROOT, "lean-fftw_amd", "codegen.py"))
    cg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(cg)
    path = os.path.join(ROOT, "lean-fftw_amd", "csrc", "bfly_gen.h")
    assert open(path).read() == cg.emit_all()


def test_generated_tile_tables_are_current(lf):
    """The committed tables of csrc/mfft_tile_gen*.hip are what lean-fftw_amd/tilegen.py writes
    for the planner as built: every shape's radices are the size's default schedule."""
    spec = importlib.util.spec_from_file_location("mfft_tilegen",
                                                  os.path.join(ROOT, "lean-fftw_amd", "tilegen.py"))
    tg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tg)
    ents = tg.entries(lf.lib())
    parts = [ents[i::tg.PARTS] for i in range(tg.PARTS)]
    for i, pe in enumerate(parts):
        path = os.path.join(ROOT, "lean-fftw_amd", "csrc", f"mfft_tile_gen{i}.hip")
        assert open(path).read() == tg.render(i, pe), path


def test_real_data_stage_tiles_match_the_planner(lf):
    """The hand table's tile variants with the real-data maps in their first / last passes
    (csrc/mfft_tile.hip, tile copies 3) are the two stages of the default schedules of 10368,
    75600 and 165375 points, with the lane orders of a two-stage plan (stage A reads columns,
    stage B reads and writes rows), each beside its plain twin of the same tile; so the
    real-data plans of those sizes fuse their maps rather than fall back to the composition."""
    import re

    src = open(os.path.join(ROOT, "lean-fftw_amd", "csrc", "mfft_tile.hip")).read()
    pat = r"MFFT_T\((\d+), \d+, RS_\(([\d, ]+)\), (\d), (\d), (\d), (\d+), (\d+),"
    table = [(int(m[0]), tuple(int(x) for x in m[1].split(",")), int(m[2]), int(m[3]), int(m[4]),
              int(m[5]), int(m[6])) for m in re.findall(pat, src)]
    mapped = {(L, rs, fi, fo, G, NT) for L, rs, fi, fo, io, G, NT in table if io == 3}
    plain = {(L, rs, fi, fo, G, NT) for L, rs, fi, fo, io, G, NT in table if io == 0}
    want = set()
    for n in (10368, 75600, 165375):
        rs, split = lf.schedule(n)
        assert split, n
        a, b = tuple(rs[:split]), tuple(rs[split:])
        P = 1
        for r in a:
            P *= r
        for stage, (L, rr, fi, fo) in (("A", (P, a, 1, 0)), ("B", (n // P, b, 1, 1))):
            hit = [t for t in mapped if t[:4] == (L, rr, fi, fo)]
            assert len(hit) == 1, (n, stage, L, rr)
            want.add(hit[0])
    assert mapped == want
    assert mapped <= plain


def test_missing_library_fails_loudly(lf, tmp_path, monkeypatch):
    monkeypatch.setattr(lf, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(lf, "_lib", None)
    with pytest.raises(lf.MfftError):
        lf.lib()


def test_planner_schedules_are_valid(lf):
    """Every schedule multiplies to N, uses radices 2..16 with prime factors <= 13, and its
    kernel stages fit the LDS tiles or, for N = 16384, the split register kernel's schedule
    16 x 8 | 16 x 8. No GPU needed: the planner query allocates nothing."""
    import math

    w = lf.workloads
    sizes = sorted(set([2 ** k for k in range(1, 19)] + w.smooth_sizes(2, 3000)[::7] +
                       [10368, 27000, 75600, 165375, 65520, 1 << 20]))
    for n in sizes:
        rs, split = lf.schedule(n)
        assert math.prod(rs) == n, (n, rs)
        assert all(2 <= r <= 16 and w.smooth13(r) for r in rs), (n, rs)
        if split:
            P = math.prod(rs[:split])
            assert P <= 2048 and n // P <= 2048, (n, rs, split)
        elif n > 8192:
            assert n == 16384 and rs == [16, 8, 16, 8], (n, rs)
    with pytest.raises(lf.MfftError):
        lf.schedule(17)


def _bench2():
    import importlib.util
    path = os.path.join(ROOT, "lean-fftw_amd", "bench2.py")
    spec = importlib.util.spec_from_file_location("mfft_bench2", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench2_problem_strings_and_output_format():
    """The libbench2-shaped command line: problem grammar, time format and mflops convention,
    and a speed line the reference's protocol runner parses (tools/bench.py:45-49)."""
    import re

    b = _bench2()
    p = b.parse_problem("ocf1024")
    assert p == {"name": "ocf1024", "inplace": False, "real": False, "forward": True,
                 "dims": [1024]}
    assert b.parse_problem("icb64x128")["dims"] == [64, 128]
    assert b.parse_problem("orb1000")["real"] and not b.parse_problem("orb1000")["forward"]
    for bad in ("ocf", "xcf8", "irf8", "orf4x4", "ocf0"):
        with pytest.raises(ValueError):
            b.parse_problem(bad)
    assert b.fmt_time(1.25e-6) == "1.25 us" and b.fmt_time(19e-6) == "19.00 us"
    assert b.fmt_time(3.5e-9) == "3.50 ns" and b.fmt_time(2.0) == "2.00 s"
    # libbench2's convention, e.g. ocf1024 at 1.25 us -> 40960 "mflops"
    assert abs(b.mflops(p, 1.25e-6) - 5 * 1024 * 10 / 1.25) < 1e-6
    line = f"Problem: ocf1024, setup: {b.fmt_time(19e-6)}, time: {b.fmt_time(1.25e-6)}, " \
           f"``mflops'': {b.mflops(p, 1.25e-6):.3f}"
    m = re.search(r"setup:\s*([0-9.e+-]+)\s*(ns|us|ms|s),\s*time:\s*([0-9.e+-]+)\s*(ns|us|ms|s)", line)
    assert m and float(m.group(3)) == 1.25
    assert b.main(["-s", "bogus"]) == 2


def test_ragged_queue_covers_every_tile_once(lf):
    """The host queue builder of ragged batches (mfft_ragged_queue, host only): every
    transform's tiles are queued exactly once, every stage-B tile in the second launch and
    every stage-A / single-stage tile in the first (so the launch boundary orders each
    transform's stages); each launch has `nqueue` queues, a class's tiles of a launch all in
    one queue, the queues' points balanced by longest-processing-time (no two queues differ by
    more than the heaviest class); within a queue the stage-A tiles come ahead of the
    single-stage ones, classes largest first, a transform's tiles in sub-problem order; tiles
    stay within their sub-problem range, and single-stage tiles list the offsets of
    transforms of their own size."""
    import numpy as np

    # class shapes: n, stages, P, Q, tile width of stage A (or the single stage), of stage B
    shapes = [(6, 1, 0, 0, 341, 0), (60, 1, 0, 0, 34, 0), (1000, 1, 0, 0, 2, 0),
              (4096, 2, 64, 64, 16, 8), (10368, 2, 96, 108, 21, 12), (65520, 2, 240, 273, 8, 8)]
    rng = np.random.default_rng(3)
    sizes = rng.choice([s[0] for s in shapes], size=400)
    counts = {n: int((sizes == n).sum()) for n in set(sizes.tolist())}
    for nq in (1, 3, 8):
        tiles, n_first, qoff, offs = lf.ragged_queue(sizes, shapes, nq)
        check_ragged_queues(shapes, sizes, counts, tiles, n_first, qoff, offs, nq)


def check_ragged_queues(shapes, sizes, counts, tiles, n_first, qoff, offs, nq):
    import numpy as np

    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    first, second = tiles[:n_first], tiles[n_first:]
    assert (second[:, 1] == 1).all() and (first[:, 1] == 0).all()
    assert qoff[0] == 0 and qoff[nq] == n_first and qoff[nq + 1] == n_first
    assert qoff[2 * nq + 1] == len(tiles) and (np.diff(qoff) >= 0).all()
    two = [i for i, s in enumerate(shapes) if s[1] == 2]
    for launch in (0, 1):
        home, load = {}, np.zeros(nq)
        for q in range(nq):
            part = tiles[qoff[launch * (nq + 1) + q]:qoff[launch * (nq + 1) + q + 1]]
            is_two = np.isin(part[:, 0], two)
            # stage-A tiles first, then single-stage tiles; largest class first within each
            assert not (np.diff(is_two.astype(int)) > 0).any()
            for run in (part[is_two], part[~is_two]):
                assert (np.diff(run[:, 0]) <= 0).all()
            for cls in set(part[:, 0].tolist()):
                assert home.setdefault(cls, q) == q  # a class's tiles in one queue
                load[q] += shapes[cls][0] * counts[shapes[cls][0]]
        heaviest = max(shapes[c][0] * counts[shapes[c][0]] for c in home) if home else 0
        assert load.max() - load.min() <= heaviest
    seen_single = []
    cover = {}
    for cls, stage, count, base, sub0 in tiles:
        n, ns, P, Q, g0, g1 = shapes[cls]
        if ns == 1:
            assert 1 <= count <= g0
            tr = offs[base:base + count]
            assert all(sizes[np.searchsorted(starts, o)] == n for o in tr)
            seen_single.extend(int(o) for o in tr)
        else:
            lim, width = (Q, g0) if stage == 0 else (P, g1)
            assert 0 <= sub0 < lim and 1 <= count <= width and sub0 + count <= lim
            assert base in set(starts[sizes == n])
            cover.setdefault((int(base), int(stage)), []).append((int(sub0), int(count)))
    single_starts = sorted(int(o) for o, n in zip(starts, sizes) if shapes_n_stages(shapes, n) == 1)
    assert sorted(seen_single) == single_starts
    for o, n in zip(starts, sizes):
        ns, P, Q = [(s[1], s[2], s[3]) for s in shapes if s[0] == n][0]
        if ns != 2:
            continue
        for stage, lim in ((0, Q), (1, P)):
            runs = cover[(int(o), stage)]
            assert [a for a, _ in runs] == sorted(a for a, _ in runs)  # sub-problem order
            pos = 0
            for a, c in runs:  # contiguous, no overlap, complete
                assert a == pos
                pos += c
            assert pos == lim


def test_ragged_queue_rejects_bad_input(lf):
    shapes = [(6, 1, 0, 0, 341, 0), (60, 1, 0, 0, 34, 0)]
    with pytest.raises(lf.MfftError):
        lf.ragged_queue([6, 7], shapes)  # a size without a class
    with pytest.raises(lf.MfftError):
        lf.ragged_queue([6, 60], shapes, nqueue=0)


def shapes_n_stages(shapes, n):
    return [s[1] for s in shapes if s[0] == n][0]


PLANS_COST = [(1024, [16, 8, 8]), (1024, [8, 4, 4, 8]), (1000, [10, 10, 10]), (2100, [15, 14, 10]),
              (1001, [13, 11, 7]), (75600, [10, 9, 3, 10, 7, 4]), (165375, [5, 5, 5, 3, 9, 7, 7]),
              (262144, [8, 8, 8, 8, 8, 8]), (6, [6]), (12, [12])]


@pytest.mark.parametrize("fwd", [True, False])
def test_stockham_cost_against_oracle_counts(lf, fwd):
    """mfft_stockham_cost (the reference's stockhamCost, LeanFFTW/Gen/OpCount.lean:55-69) from
    the product's generated butterflies equals the same sum over the oracle's butterfly DAGs
    (whose counts are pinned to the reference's published ones), for several weightings; the
    memory weights alone give the closed form 4R loads / stores per untwiddled and 6R - 2 per
    twiddled butterfly."""
    from oracle.symgen import kernel_counts

    for n, rs in PLANS_COST:
        for W in ((1, 1, 1, 0, 0), (1, 1, 2, 0, 0), (3, 5, 7, 11, 13)):
            want = 0
            for t, R in enumerate(rs):
                a, m, f = kernel_counts(fwd, R, t >= 1)
                loads = 2 * R + (2 * (R - 1) if t >= 1 else 0)
                want += (n // R) * (W[0] * a + W[1] * m + W[2] * f + W[3] * loads + W[4] * 2 * R)
            assert lf.stockham_cost(fwd, n, rs, W) == want, (n, rs, W)
        mem = sum((n // R) * (4 * R if t == 0 else 6 * R - 2) for t, R in enumerate(rs))
        assert lf.stockham_cost(fwd, n, rs, (0, 0, 0, 1, 1)) == mem
    with pytest.raises(lf.MfftError):
        lf.stockham_cost(True, 1024, [16, 8, 4], (1, 1, 1, 0, 0))


def test_schedule_cost_orders_stages_and_passes(lf):
    """The tuner's GPU model: a second kernel stage (another HBM round trip) costs more than
    any single-stage schedule of the size; a further LDS pass costs more than fewer passes of
    the same arithmetic class."""
    one = lf.schedule_cost(True, 4096, [8, 8, 8, 8], 0)
    two = lf.schedule_cost(True, 4096, [8, 8, 8, 8], 2)
    assert two > 1.4 * one
    assert lf.schedule_cost(True, 4096, [4, 4, 4, 4, 4, 4], 0) > lf.schedule_cost(True, 4096, [16, 16, 16], 0)
    # HBM dominates: 32 N bytes at 6.5 TB/s per stage
    assert 32 * 4096 / 6.5e3 < one < 2 * 32 * 4096 / 6.5e3


def test_planner_covers_the_reference_domain(lf):
    """Every 13-smooth size of the reference's fast engine's domain (2 <= N <= 2^24,
    bench-leanfftw.c:253-268) has a schedule: one, two or three kernel stages, each stage's
    sub-problem within the 2048-point LDS tile."""
    lim = 1 << 24
    nums = {1}
    for pr in (2, 3, 5, 7, 11, 13):
        grown = set()
        for n in nums:
            m = n
            while m * pr <= lim:
                m *= pr
                grown.add(m)
        nums |= grown
    nums = sorted(n for n in nums if n >= 2)
    assert len(nums) == 9597
    for n in nums:
        rs, split = lf.schedule(n)
        assert int(np.prod(rs)) == n
    with pytest.raises(lf.MfftError):
        lf.schedule(17)


def test_plan_cache_under_capture(lf, monkeypatch):
    """Outside a capture the torch-level cache makes one plan per (shape, stream); during a
    capture it creates none (a plan's table uploads are synchronous copies, which a capture
    forbids): it takes this shape's plan from another stream of the device, or raises."""
    made = []
    state = {"stream": (0, 111), "capturing": False}
    monkeypatch.setattr(lf, "Plan", lambda n, sign, howmany: made.append((n, sign, howmany)) or object())
    monkeypatch.setattr(lf, "stream_key", lambda: state["stream"])
    monkeypatch.setattr(lf, "_capturing", lambda: state["capturing"])
    cache = {}
    a = lf._cached_plan(4096, -1, 8, cache)
    assert lf._cached_plan(4096, -1, 8, cache) is a and len(made) == 1
    state["stream"] = (0, 222)
    b = lf._cached_plan(4096, -1, 8, cache)
    assert b is not a and len(made) == 2
    state.update(stream=(0, 333), capturing=True)
    assert lf._cached_plan(4096, -1, 8, cache) in (a, b)
    for shape in ((4096, 1, 8), (4096, -1, 16), (2048, -1, 8)):
        with pytest.raises(lf.MfftError, match="capture"):
            lf._cached_plan(*shape, cache)
    state["stream"] = (1, 333)  # another device's plans are not taken
    with pytest.raises(lf.MfftError):
        lf._cached_plan(4096, -1, 8, cache)
    assert len(made) == 2
