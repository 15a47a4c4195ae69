"""Generate the tables of fixed-schedule tile kernels for every 13-smooth size up to 4096 points
that is not a power of two (csrc/mfft_tile_gen0.hip ... mfft_tile_gen7.hip).

    python lean-fftw_amd/tilegen.py      # needs a built _build/libmfft.so (host calls only)

Each entry is the single-stage contiguous shape of the size's default plan: the radices come
from the library's own planner (mfft_plan_schedule, a host-only query), the tile holds
G = 2048 // n transforms (one above 2048 points) on 256 threads, sizes up to 16 points stage
the tile through LDS (as the planner does), and every size also gets the staged variant that
takes the real-data maps (the complex parts of r2c / c2r plans: the real input or Hermitian
half read, the half or real output written, in the tile copies). Shapes already in the
hand-tuned table of mfft_tile.hip are skipped. tests/test_abi.py checks that the
committed tables still match the planner.
"""

from __future__ import annotations

import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
PARTS = 8
STAGE_IO_MAX = 16   # the planner's staged-tile limit for plain complex plans
# shapes of the hand table (mfft_tile.hip): (n, io)
HAND = {(6, 1), (6, 2), (60, 0), (360, 0), (1000, 0), (2100, 0)}
# sizes whose tile measured slower than the stage kernel on MI355X (64 MB batches: 20, 25 and 36
# points 3-7% slower; the 472 generated complex shapes run a median 1.26x faster, up to 2.3x)
SLOWER = {(20, 0), (25, 0), (36, 0)}


def smooth13(n: int) -> bool:
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


def schedule(lib, n: int):
    buf = (ctypes.c_int * 32)()
    split = ctypes.c_int(0)
    k = lib.mfft_plan_schedule(ctypes.c_int64(n), buf, 32, ctypes.byref(split))
    if k <= 0:
        raise RuntimeError(f"no schedule for {n}")
    return list(buf[:k]), split.value


def entries(lib):
    """(n, radices, io, G, NT) of every generated shape, in size order."""
    out = []
    for n in range(3, 4097):
        if not smooth13(n) or (n & (n - 1)) == 0:
            continue
        rs, split = schedule(lib, n)
        if split:
            continue
        G = max(1, 2048 // n)
        ios = [1 if n <= STAGE_IO_MAX else 0, 2]
        for io in ios:
            if (n, io) not in HAND and (n, io) not in SLOWER:
                out.append((n, rs, io, G, 256))
    return out


def render(part: int, ents) -> str:
    lines = [
        f"// Generated by lean-fftw_amd/tilegen.py (part {part} of {PARTS}): fixed-schedule tile",
        "// kernels for the default single-stage plans of 13-smooth sizes that are not powers of two.",
        '#include "mfft_tile.h"',
        "",
        "namespace mfft {",
        "",
        f"static const TileSchedule kPart[] = {{",
    ]
    for n, rs, io, G, NT in ents:
        r = ", ".join(str(x) for x in rs)
        lines.append(f"    MFFT_T({n}, {len(rs)}, RS_({r}), 0, 0, {io}, {G}, {NT}, {r}),")
    if not ents:
        lines.append("    {},")
    lines += [
        "};",
        f"extern const TileSchedule *const kTileGenPart{part} = kPart;",
        f"extern const int kTileGenPart{part}Count = {len(ents)};",
        "",
        "}  // namespace mfft",
        "",
    ]
    return "\n".join(lines)


def main():
    lib = ctypes.CDLL(os.path.join(HERE, "_build", "libmfft.so"))
    lib.mfft_plan_schedule.restype = ctypes.c_int
    ents = entries(lib)
    # round-robin over the parts, so that each translation unit takes a similar mix of sizes
    parts = [ents[i::PARTS] for i in range(PARTS)]
    for i, pe in enumerate(parts):
        with open(os.path.join(CSRC, f"mfft_tile_gen{i}.hip"), "w") as f:
            f.write(render(i, pe))
    print(f"{len(ents)} shapes in {PARTS} parts")


if __name__ == "__main__":
    sys.exit(main())
