// Fixed-schedule LDS-tile stage kernels: the hand-tuned table of stage shapes, the lookup over
// it and the generated tables, and the launch. The kernels themselves are in mfft_tile.h.
#include <cstdlib>
#include <mutex>
#include <set>

#include "mfft_tile.h"

namespace mfft {

// The fixed stage shapes: sub-problem length, radix sequence, lane order per side, tile copies,
// and the tile (G sub-problems, NT threads). The shapes are those of the default plans of the
// benchmarked sizes; each G / NT is the best of the variants measured on MI355X (256 MB
// batches; per launch): 6 points G = 256 89 us / G = 341 95; 16 points G = 128 87 / 256 95;
// 2100 points G = 1 of 256 threads 89 / G = 2 of 512 97; 270 and 280 points 256 threads ahead
// of 512 (75600: 230 / 246 us); 375 and 441 points 512 threads ahead of 256, and G = 8 ahead of
// G = 4 (165375: 317 / 401 us). Reading the small sizes butterfly by butterfly instead of
// staging the tile costs 1.1x (2 points) to 7x (16 points).
const TileSchedule kTiles[] = {
    // single-stage contiguous batches of small transforms (the tile staged through LDS)
    MFFT_T(2, 1, RS_(2), 0, 0, 1, 1024, 256, 2),
    MFFT_T(4, 1, RS_(4), 0, 0, 1, 512, 256, 4),
    MFFT_T(8, 1, RS_(8), 0, 0, 1, 256, 256, 8),
    MFFT_T(16, 1, RS_(16), 0, 0, 1, 128, 256, 16),
    MFFT_T(6, 1, RS_(6), 0, 0, 1, 256, 256, 6),
    // the complex parts of real-data plans of powers of two up to 64 points and of 6 points
    // (the generated tables hold every other size): staged tiles whose copies take
    // the real-data maps (tile copies 2; r2c per 256 MB of complex-equivalent batch on MI355X:
    // 16 and 32 points 221 / 207 -> 44 / 52 us against the unfused composition, 64 points
    // 65 -> 52 against the register kernel's fused maps, which are faster from 128 points)
    MFFT_T(2, 1, RS_(2), 0, 0, 2, 1024, 256, 2),
    MFFT_T(4, 1, RS_(4), 0, 0, 2, 512, 256, 4),
    MFFT_T(8, 1, RS_(8), 0, 0, 2, 256, 256, 8),
    MFFT_T(16, 1, RS_(16), 0, 0, 2, 128, 256, 16),
    MFFT_T(6, 1, RS_(6), 0, 0, 2, 256, 256, 6),
    MFFT_T(32, 2, RS_(8, 4), 0, 0, 2, 64, 256, 8, 4),
    MFFT_T(64, 2, RS_(8, 8), 0, 0, 2, 32, 256, 8, 8),
    // single-stage contiguous batches read butterfly by butterfly; from 64 points on in place
    // of the register-resident kernel (MI355X, 256 MB batches: 64 ... 1024 points 78-83 us
    // against 88-98, 2048 points 82 against 90; radix 16 measured 1-2 us slower at 256, 512
// and 2048 points, 12 slower as 16 x 16 x 4 at 1024, 6 faster as 16 x 8 x 8)
    MFFT_T(32, 2, RS_(8, 4), 0, 0, 0, 64, 256, 8, 4),
    MFFT_T(64, 2, RS_(8, 8), 0, 0, 0, 32, 256, 8, 8),
    MFFT_T(128, 3, RS_(8, 8, 2), 0, 0, 0, 16, 256, 8, 8, 2),
    MFFT_T(256, 3, RS_(8, 8, 4), 0, 0, 0, 8, 256, 8, 8, 4),
    MFFT_T(512, 3, RS_(8, 8, 8), 0, 0, 0, 4, 256, 8, 8, 8),
    MFFT_T(1024, 4, RS_(8, 8, 8, 2), 0, 0, 0, 2, 256, 8, 8, 8, 2),
    MFFT_T(2048, 4, RS_(8, 8, 8, 4), 0, 0, 0, 1, 256, 8, 8, 8, 4),
    // the default plain schedule of 2048 points (80.5 us against 82.9 for 8 x 8 x 8 x 4; other
    // orders of those radices 84-90, 8 x 16 x 16 81-82)
    MFFT_T(2048, 3, RS_(16, 8, 16), 0, 0, 0, 1, 256, 16, 8, 16),
    // the default plain schedule of 1024 points (radix 16: one LDS pass fewer); a single
    // transform takes a tile of one
    MFFT_T(1024, 3, RS_(16, 8, 8), 0, 0, 0, 2, 256, 16, 8, 8),
    MFFT_T(1024, 3, RS_(16, 8, 8), 0, 0, 0, 1, 256, 16, 8, 8),
    MFFT_T(60, 2, RS_(12, 5), 0, 0, 0, 34, 256, 12, 5),
    MFFT_T(360, 3, RS_(12, 10, 3), 0, 0, 0, 5, 256, 12, 10, 3),
    MFFT_T(1000, 3, RS_(10, 10, 10), 0, 0, 0, 2, 256, 10, 10, 10),
    MFFT_T(2100, 3, RS_(15, 14, 10), 0, 0, 0, 1, 256, 15, 14, 10),
    // stages of two-stage plans: A reads rows of G adjacent columns and writes each column
    // contiguously; B reads and writes rows of G adjacent residue classes. The wider tiles come
    // first (MI355X, 256 MB batches: 10368 = 96 x 108 at G = 42 / 36 on 512 threads 211 ->
    // 183 us against G = 21 / 18 on 256; 165375's 441-point stage at G = 16 on 1024 threads
    // 315 -> 285 us against G = 8 on 512, 295 at G = 16 on 512; 75600's 280-point stage at
    // G = 12 on 512 threads 218-222 -> 212-213 us against G = 8 on 256, G = 8 or 16 and 384 or
    // 768 threads slower; wider tiles of 270 and 375 points measured equal or slower), the
    // narrower ones take stages of fewer sub-problems
    MFFT_T(96, 2, RS_(12, 8), 1, 0, 0, 42, 512, 12, 8),
    MFFT_T(108, 2, RS_(12, 9), 1, 1, 0, 36, 512, 12, 9),
    MFFT_T(441, 3, RS_(9, 7, 7), 1, 1, 0, 16, 1024, 9, 7, 7),
    MFFT_T(280, 3, RS_(10, 7, 4), 1, 1, 0, 12, 512, 10, 7, 4),
    MFFT_T(96, 2, RS_(12, 8), 1, 0, 0, 21, 256, 12, 8),
    MFFT_T(108, 2, RS_(12, 9), 1, 1, 0, 18, 256, 12, 9),
    MFFT_T(270, 3, RS_(10, 9, 3), 1, 0, 0, 8, 256, 10, 9, 3),
    MFFT_T(280, 3, RS_(10, 7, 4), 1, 1, 0, 8, 256, 10, 7, 4),
    MFFT_T(375, 4, RS_(5, 5, 5, 3), 1, 0, 0, 8, 512, 5, 5, 5, 3),
    MFFT_T(441, 3, RS_(9, 7, 7), 1, 1, 0, 8, 512, 9, 7, 7),
    // the same stages of the real-data plans of those sizes, with the maps in their first /
    // last passes' accesses (stage A reads the real input or Hermitian half, stage B writes the
    // half or the real parts)
    MFFT_T(96, 2, RS_(12, 8), 1, 0, 3, 42, 512, 12, 8),
    MFFT_T(108, 2, RS_(12, 9), 1, 1, 3, 36, 512, 12, 9),
    MFFT_T(270, 3, RS_(10, 9, 3), 1, 0, 3, 8, 256, 10, 9, 3),
    MFFT_T(280, 3, RS_(10, 7, 4), 1, 1, 3, 12, 512, 10, 7, 4),
    MFFT_T(375, 4, RS_(5, 5, 5, 3), 1, 0, 3, 8, 512, 5, 5, 5, 3),
    MFFT_T(441, 3, RS_(9, 7, 7), 1, 1, 3, 16, 1024, 9, 7, 7),
    // batches of adjacent transforms (the columns of a row-major array: lanes walk the
    // transforms on both sides), and the stages of two-stage column plans (2-D 64 x 64:
    // 173 -> 167 us, 4096 x 4096: 276 -> 267 per 2^24 elements; a tile for the 2048-point
    // columns, one workgroup per CU either way, measured the same as the stage kernel; the
    // 32-point stage of 2048-point columns as 32 x 64: 2048 x 2048 316-322 -> 285-287 us, with
    // tiles of 64 columns 325, of 16 or 32 on 128 threads 291-293, of 64 on 512 threads 323)
    MFFT_T(64, 2, RS_(8, 8), 1, 1, 0, 32, 256, 8, 8),
    MFFT_T(32, 2, RS_(8, 4), 1, 1, 0, 32, 256, 8, 4),
};
#undef RS_
#undef MFFT_T
constexpr int kNumTiles = (int)(sizeof(kTiles) / sizeof(kTiles[0]));

// Whether a launch of this stage remaps its workgroups by XCD, for rows that span at least 8
// tiles of width G: stages whose twiddle base depends on the sub-problem (kb_u != 0, the second stage of
// a two-stage plan, whose twiddles span the whole table; MI355X: 75600 points 233 -> 218 us,
// 165375 points 331 -> 313 per 256 MB batch), and stages reading rows of adjacent
// sub-problems whose stride is not a whole number of 128-byte lines, so that the tiles that
// share a line meet in one L2 (165375's first stage, rows 7056 B apart: 288 -> 277-279 us;
// 75600's, 4480 B apart, lost 11 us to it). MFFT_XCD=0 turns it off.
int xcd_remap(const StageArgs &a, int G) {
  static const int on = [] {
    const char *e = std::getenv("MFFT_XCD");
    return e ? std::atoi(e) : 1;
  }();
  // (a 128-byte line holds 8 complex values, or 16 doubles of real input)
  const bool split_lines = a.gfast_in && a.in_e % (a.io_in == 1 ? 16 : 8) != 0;
  return on && (a.kb_u != 0 || split_lines) && a.nsub >= 8 * (int64_t)G ? 1 : 0;
}

extern const TileSchedule *const kTileGenPart0, *const kTileGenPart1, *const kTileGenPart2,
    *const kTileGenPart3, *const kTileGenPart4, *const kTileGenPart5, *const kTileGenPart6,
    *const kTileGenPart7;
extern const int kTileGenPart0Count, kTileGenPart1Count, kTileGenPart2Count, kTileGenPart3Count,
    kTileGenPart4Count, kTileGenPart5Count, kTileGenPart6Count, kTileGenPart7Count;
const TileSchedule *const kTileGen[kTileGenParts] = {
    kTileGenPart0, kTileGenPart1, kTileGenPart2, kTileGenPart3,
    kTileGenPart4, kTileGenPart5, kTileGenPart6, kTileGenPart7};
const int kTileGenCount[kTileGenParts] = {
    kTileGenPart0Count, kTileGenPart1Count, kTileGenPart2Count, kTileGenPart3Count,
    kTileGenPart4Count, kTileGenPart5Count, kTileGenPart6Count, kTileGenPart7Count};

// Schedule i of the hand table followed by the generated tables (nullptr past the end).
const TileSchedule *tile_at(int i) {
  if (i < 0) return nullptr;
  if (i < kNumTiles) return &kTiles[i];
  i -= kNumTiles;
  for (int q = 0; q < kTileGenParts; ++q) {
    if (i < kTileGenCount[q]) return &kTileGen[q][i];
    i -= kTileGenCount[q];
  }
  return nullptr;
}

// The tile kernel of a stage (1 + index, 0 if none; MFFT_TILE=0 turns them off): same length,
// radices, lane orders and tile copies, and at least one full tile of sub-problems; the hand
// table first. Register stage kernels (reg 2) keep their own kernel.
int tile_lookup(const StageArgs &a) {
  static const int on = [] {
    const char *e = std::getenv("MFFT_TILE");
    return e ? std::atoi(e) : 1;
  }();
  if (!on || a.reg == 2) return 0;
  for (int i = 0;; ++i) {
    const TileSchedule *t = tile_at(i);
    if (!t) return 0;
    if (t->L != a.L || t->npass != a.npass || t->gfi != (a.gfast_in ? 1 : 0) ||
        t->gfo != (a.gfast_out ? 1 : 0) || t->io != a.stage_io || a.nsub < t->G)
      continue;
    bool same = true;
    for (int p = 0; p < a.npass; ++p) same = same && t->rs[p] == a.pass[p].R;
    if (same) return i + 1;
  }
}

// Whether tile kernel `tile` (1 + index) fits the stage: its shape, radices, lane orders and
// tile copies, and at least one full tile of sub-problems.
bool tile_fits(const StageArgs &a, int tile) {
  const TileSchedule *t = tile_at(tile - 1);
  if (!t || a.reg == 2) return false;
  if (t->L != a.L || t->npass != a.npass || t->gfi != (a.gfast_in ? 1 : 0) ||
      t->gfo != (a.gfast_out ? 1 : 0) || t->io != a.stage_io || a.nsub < t->G)
    return false;
  for (int p = 0; p < a.npass; ++p)
    if (t->rs[p] != a.pass[p].R) return false;
  return true;
}

// Every tile kernel that fits the stage (1 + index each, table order), at most `max`; returns
// how many fit. The autotuner races them.
int tile_candidates(const StageArgs &a, int *tiles, int max) {
  int n = 0;
  for (int i = 0;; ++i) {
    if (!tile_at(i)) return n;
    if (tile_fits(a, i + 1)) {
      if (n < max) tiles[n] = i + 1;
      ++n;
    }
  }
}

// The tile (sub-problems, threads) of tile kernel `tile` (1 + index), the identity wisdom
// records.
bool tile_shape(int tile, int *G, int *NT) {
  const TileSchedule *t = tile_at(tile - 1);
  if (!t) return false;
  *G = t->G;
  *NT = t->NT;
  return true;
}

hipError_t launch_tile(bool fwd, const StageArgs &a, int64_t ntrans, const double2 *in,
                       double2 *out, const double2 *tw, hipStream_t stream) {
  const TileSchedule *tp = tile_at(a.tile - 1);
  if (!tp) return hipErrorInvalidValue;
  const TileSchedule &t = *tp;
  const void *k = t.k[fwd ? 1 : 0];
  {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.insert(k).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax);
  }
  StageArgs b = a;
  b.xcd = xcd_remap(a, t.G);  // the tile's own width, which may differ from the plan's G
  if (a.reg == 1) {
    // in place of the register kernel, whose load / store policy is its own: ordinary loads
    // and non-temporal stores measured best here (ordinary stores 1.2x slower, non-temporal
    // loads 1.1x)
    b.cached_in = 1;
    b.cached_out = 0;
  }
  int64_t nx = (a.nsub + t.G - 1) / t.G;
  if (b.xcd) nx = (nx + 7) / 8 * 8;
  const size_t lds = (size_t)stage_lds_elems(t.G, t.L, t.npass, t.io) * sizeof(double2);
  void *args[] = {(void *)&in, (void *)&out, (void *)&tw, (void *)&b};
  return hipLaunchKernel(k, dim3((unsigned)nx, (unsigned)ntrans), dim3(t.NT), args, lds, stream);
}

}  // namespace mfft
