"""GPU parity of the fixed-schedule LDS-tile kernels (csrc/mfft_tile.hip).

Default plans whose stages have a compile-time tile schedule run these kernels once the batch
fills a tile, so each case takes a batch of at least one full tile plus a partial last one
(for two-stage plans the stages' sub-problem counts Q and P do that). Bar: bit-exact against
the oracle's refTransform for the plan's radices, both directions, out of place and in place.
"""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():  # pragma: no cover - CPU-only containers
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402

DEV = "cuda:0"

# (N, batch): single-stage sizes with tiles of G transforms (2: 1024, 4: 512, 8: 256,
# 16: 128, 6: 341, 32: 64, 60: 34, 360: 5, 1000: 2, 2100: 1) and two-stage sizes whose
# stages have tile schedules (10368 = 96 x 108, 75600 = 270 x 280, 165375 = 375 x 441; their
# second stages run XCD-remapped grids padded to a multiple of 8 tiles, over one or several
# transforms); 64 ...
# 2048 points run the tile kernels in place of the register-resident kernel, 4096 and the
# two-stage power-of-two sizes keep the register kernels
TILE_CASES = [(2, 1537), (4, 769), (8, 385), (16, 193), (6, 513), (32, 97), (60, 52),
              (360, 8), (1000, 3), (2100, 3), (10368, 2), (75600, 1), (165375, 1),
              (75600, 3), (165375, 2),
              (64, 49), (128, 25), (256, 13), (512, 7), (1024, 3), (2048, 2), (4096, 2),
              (32768, 1), (65536, 1), (131072, 1), (262144, 1)]


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


@pytest.mark.parametrize("N,hm", TILE_CASES)
@pytest.mark.parametrize("fwd", [True, False])
def test_tile_kernels_bit_exact(lf, oracle, N, hm, fwd):
    run_case(lf, oracle, N, hm, fwd, None)


# the radix-16 tile schedules with explicit radices: a tile of two and a single transform
RADIX16_CASES = [(1024, 3, [16, 8, 8]), (1024, 1, [16, 8, 8]), (2048, 2, [16, 8, 16]),
                 (2048, 1, [16, 8, 16])]


@pytest.mark.parametrize("N,hm,rs", RADIX16_CASES)
@pytest.mark.parametrize("fwd", [True, False])
def test_tile_radix16_schedules_bit_exact(lf, oracle, N, hm, rs, fwd):
    run_case(lf, oracle, N, hm, fwd, rs)


# two-stage stage shapes with two tiles in the table, the wider first: stages with fewer
# sub-problems than the wide tile holds take the narrow one (2304 = 96 x 24: 24 columns of 96
# points, tiles of 21 rather than 42; 2592 = 24 x 108: 24 residue classes of 108, tiles of 18
# rather than 36; 4410 = 10 x 441: 10 residue classes of 441, tiles of 8 rather than 16;
# 2800 = 10 x 280: 10 residue classes of 280, tiles of 8 rather than 12)
NARROW_STAGE_CASES = [(2304, [12, 8, 12, 2], 2), (2592, [12, 2, 12, 9], 2),
                      (4410, [10, 9, 7, 7], 1), (2800, [10, 10, 7, 4], 1)]


@pytest.mark.parametrize("N,rs,split", NARROW_STAGE_CASES)
@pytest.mark.parametrize("fwd", [True, False])
def test_narrow_stage_tiles_bit_exact(lf, oracle, N, rs, split, fwd):
    run_case(lf, oracle, N, 3, fwd, rs, split)


def run_case(lf, oracle, N, hm, fwd, rs, split=None):
    x = oracle.lcg_input(N, seed=0x7E1 ^ N ^ (fwd << 33), howmany=hm)
    kw = {} if split is None else {"split": split}
    plan = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, hm, radices=rs, **kw)
    ref = oracle.ref_transform(fwd, plan.radices, x)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to(DEV)
    out = torch.empty_like(xd)
    plan.call(xd, out)
    plan.call(xd, xd)  # in place
    torch.cuda.synchronize()
    assert bits_equal(out.cpu().numpy(), ref), (N, plan.radices, plan.stages, "out of place")
    assert bits_equal(xd.cpu().numpy(), ref), (N, plan.radices, plan.stages, "in place")


# columns of a row-major (N, hm) array: single-stage tiles of 32 adjacent transforms of 64 and
# of 32 points (rows 40 elements apart, and 300 apart: not whole 128-byte lines, so the launch
# is XCD-remapped over a grid padded from 10 to 16 tiles), both stages of the two-stage column
# plan of 4096 (64 x 64) points, and the single register stage of 2048-point columns (tiles of
# 4 columns, XCD-remapped)
COLUMN_TILE_CASES = [(64, 40), (64, 300), (32, 300), (4096, 40), (2048, 40)]
COLUMN_STAGES = {4096: [2, 2], 2048: [4]}


@pytest.mark.parametrize("N,hm", COLUMN_TILE_CASES)
@pytest.mark.parametrize("fwd", [True, False])
def test_column_tile_kernels_bit_exact(lf, oracle, N, hm, fwd):
    x = oracle.lcg_input(N, seed=0xC01 + N + hm, howmany=hm)
    xt = torch.from_numpy(np.ascontiguousarray(x.T)).to(DEV)
    plan = lf.Plan(N, lf.FORWARD if fwd else lf.BACKWARD, hm, istride=hm, idist=1, ostride=hm,
                   odist=1)
    assert plan.stages == COLUMN_STAGES.get(N, plan.stages), (N, plan.radices, plan.stages)
    ref = np.ascontiguousarray(oracle.ref_transform(fwd, plan.radices, x).T)
    out = torch.empty_like(xt)
    plan.call(xt, out)
    plan.call(xt, xt)  # in place
    torch.cuda.synchronize()
    assert bits_equal(out.cpu().numpy(), ref), (N, plan.radices, plan.stages, "out of place")
    assert bits_equal(xt.cpu().numpy(), ref), (N, plan.radices, plan.stages, "in place")


# real-data plans whose complex part is a staged tile kernel take the maps into its tile copies
# (a batch of at least one tile: 2: 1024, 4: 512, 8: 256, 16: 128, 6: 256, 32: 64, 64: 32
# transforms)
REAL_TILE_CASES = [(2, 1100), (4, 600), (8, 300), (16, 200), (6, 300), (32, 70), (64, 40)]


def _herm_ext(h, n):
    """hermExt (Problem.lean:138-140): h_k for k <= n/2, conj(h_{n-k}) above."""
    k = np.arange(n)
    return np.where(k <= n // 2, h[:, np.minimum(k, n // 2)],
                    np.conj(h[:, np.minimum((n - k) % n, n // 2)]))


@pytest.mark.parametrize("n,hm", REAL_TILE_CASES)
def test_real_data_maps_in_tile_copies_bit_exact(lf, oracle, n, hm):
    rng = np.random.default_rng(n + hm)
    x = rng.standard_normal((hm, n))
    plan = lf.RealPlan(n, lf.FORWARD, hm)
    got = plan.exec(torch.from_numpy(x).to(DEV))
    torch.cuda.synchronize()
    full = oracle.ref_transform(True, plan.radices, x.astype(np.complex128))
    assert bits_equal(got.cpu().numpy(), full[:, : n // 2 + 1]), (n, "r2c")
    h = rng.standard_normal((hm, n // 2 + 1)) + 1j * rng.standard_normal((hm, n // 2 + 1))
    plan = lf.RealPlan(n, lf.BACKWARD, hm)
    got = plan.exec(torch.from_numpy(h).to(DEV))
    torch.cuda.synchronize()
    full = oracle.ref_transform(False, plan.radices, _herm_ext(h, n))
    assert bits_equal(got.cpu().numpy(), np.ascontiguousarray(full.real)), (n, "c2r")


# the generated tables (csrc/mfft_tile_gen*.hip): every 13-smooth size up to 4096 points that is
# not a power of two, through its default plan, with one full tile and a partial one
def _generated_shapes():
    import os
    import re

    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "lean-fftw_amd", "csrc")
    out = []
    for i in range(8):
        with open(os.path.join(here, f"mfft_tile_gen{i}.hip")) as f:
            for m in re.finditer(r"MFFT_T\((\d+), \d+, RS_\(([\d, ]+)\), 0, 0, (\d), (\d+), (\d+),",
                                 f.read()):
                out.append((int(m.group(1)), [int(x) for x in m.group(2).split(",")],
                            int(m.group(3)), int(m.group(4))))
    return sorted(out)


GEN = _generated_shapes()
GEN_COMPLEX = [(n, rs, G) for n, rs, io, G in GEN if io != 2]
GEN_REAL = [(n, rs, G) for n, rs, io, G in GEN if io == 2]


@pytest.mark.parametrize("n,rs,G", GEN_COMPLEX, ids=[str(g[0]) for g in GEN_COMPLEX])
def test_generated_tile_kernels_bit_exact(lf, oracle, n, rs, G):
    hm = G + G // 2 + 1
    fwd = n % 2 == 0  # both directions over the set
    x = oracle.lcg_input(n, seed=0x6E7 ^ n, howmany=hm)
    plan = lf.Plan(n, lf.FORWARD if fwd else lf.BACKWARD, hm)
    assert plan.radices == rs, (n, plan.radices, rs)
    ref = oracle.ref_transform(fwd, rs, x)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to(DEV)
    out = torch.empty_like(xd)
    plan.call(xd, out)
    torch.cuda.synchronize()
    assert bits_equal(out.cpu().numpy(), ref), (n, rs)


@pytest.mark.parametrize("n,rs,G", GEN_REAL, ids=[str(g[0]) for g in GEN_REAL])
def test_generated_real_data_tiles_bit_exact(lf, oracle, n, rs, G):
    hm = G + G // 2 + 1
    rng = np.random.default_rng(n)
    x = rng.standard_normal((hm, n))
    plan = lf.RealPlan(n, lf.FORWARD, hm)
    assert plan.radices == rs, (n, plan.radices, rs)
    got = plan.exec(torch.from_numpy(x).to(DEV))
    torch.cuda.synchronize()
    full = oracle.ref_transform(True, rs, x.astype(np.complex128))
    assert bits_equal(got.cpu().numpy(), full[:, : n // 2 + 1]), (n, "r2c")
    h = rng.standard_normal((hm, n // 2 + 1)) + 1j * rng.standard_normal((hm, n // 2 + 1))
    plan = lf.RealPlan(n, lf.BACKWARD, hm)
    got = plan.exec(torch.from_numpy(h).to(DEV))
    torch.cuda.synchronize()
    full = oracle.ref_transform(False, rs, _herm_ext(h, n))
    assert bits_equal(got.cpu().numpy(), np.ascontiguousarray(full.real)), (n, "c2r")


# 2-D plans of small arrays run both axes in one pass through LDS (csrc/mfft_tile2d.hip)
def _oracle_2d(lf, oracle, fwd, x):
    """Axis 0 then axis 1 with the oracle, each with the radices of the plan's part."""
    hm, n0, n1 = x.shape
    sign = lf.FORWARD if fwd else lf.BACKWARD
    rs0 = lf.Plan(n0, sign, n1, istride=n1, idist=1, ostride=n1, odist=1).radices
    rs1 = lf.Plan(n1, sign, n0 * hm, istride=1, idist=n1, ostride=1, odist=n1).radices
    cols = np.ascontiguousarray(x.transpose(0, 2, 1)).reshape(hm * n1, n0)
    a = oracle.ref_transform(fwd, rs0, cols).reshape(hm, n1, n0).transpose(0, 2, 1)
    rows = np.ascontiguousarray(a).reshape(hm * n0, n1)
    return oracle.ref_transform(fwd, rs1, rows).reshape(hm, n0, n1)


@pytest.mark.parametrize("n0,n1,hm", [(64, 64, 3), (32, 32, 5), (32, 64, 2), (64, 32, 2),
                                      (32, 128, 2), (128, 32, 2), (16, 16, 4), (16, 32, 3),
                                      (32, 16, 3), (16, 64, 2), (64, 16, 2), (16, 128, 2),
                                      (128, 16, 2)])
@pytest.mark.parametrize("fwd", [True, False])
def test_fused_2d_small_arrays_bit_exact(lf, oracle, n0, n1, hm, fwd):
    rng = np.random.default_rng(n0 * 1000 + n1 + hm)
    x = rng.standard_normal((hm, n0, n1)) + 1j * rng.standard_normal((hm, n0, n1))
    plan = lf.Plan2D(n0, n1, lf.FORWARD if fwd else lf.BACKWARD, hm)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to(DEV)
    out = plan.exec(xd)
    plan.call(xd, xd)  # in place
    torch.cuda.synchronize()
    ref = _oracle_2d(lf, oracle, fwd, x)
    assert bits_equal(out.cpu().numpy(), ref), (n0, n1, "out of place")
    assert bits_equal(xd.cpu().numpy(), ref), (n0, n1, "in place")
