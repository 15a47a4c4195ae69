// Fixed-schedule LDS-tile stage kernels for gfx950.
//
// The stage kernel of mfft_kernels.hip reads its radix sequence, tile shape and pass
// strides at run time: one kernel holds every radix of its set, and the pass loop's
// run-time divisors, twiddle strides and 64-bit offsets cost it 88-196 VGPRs and hundreds of
// spilled SGPRs. Here the sub-problem length L, the sub-problems per workgroup G, the
// workgroup size NT, the lane order on each HBM side and the radix sequence are template
// parameters: the passes become straight-line code over constant strides, with only the
// radices the stage uses. The arithmetic is the stage kernel's (the reference's passWith,
// LeanFFTW/Gen/Stockham.lean:88-98, with the butterflies of csrc/bfly_gen.h and the same
// twiddle reads), so the bits are the same.
//
// A workgroup takes sub-problems g0 .. g0 + G - 1 of one row (transform) of the grid (its tile
// and row come from its block index, remapped by XCD for second stages). Pass 0 reads HBM (or,
// with IO, a coalesced copy stages the tile through LDS first), the intermediate passes run in
// place in one padded LDS buffer, the last pass writes HBM (or the tile is copied out).
// GFI / GFO: consecutive lanes walk the sub-problem index on the input / output side (the
// sub-problems are adjacent in memory), else the butterfly index. IO: 0 the tile is read and
// written by the first and last passes, 1 staged through LDS by coalesced tile copies, 2 staged
// with the real-data maps of r2c / c2r plans in the tile copies.
#include <cstdlib>
#include <mutex>
#include <set>

#include "mfft_pass.h"

namespace mfft {

template <bool FWD, int R, int L, int G, int NT, int NSL, bool SRC_HBM, bool DST_HBM, bool GF>
__device__ __forceinline__ void tile_pass(double2 *buf, const HbmSide &hin, const HbmSide &hout,
                                          int tid, int gcount, const double2 *__restrict__ tw,
                                          const PassArgs &pa, int64_t kb0, int64_t kb_u) {
  constexpr int LR = L / R, NBT = G * LR, NB = (NBT + NT - 1) / NT, LL = lds_len(L);
  static_assert(L % R == 0, "radix must divide the sub-problem length");
  const bool twd = pa.twoff >= 0;
  const double2 *twp = tw + (twd ? pa.twoff : 0);
  double2 x[NB][R];
  int gs[NB], js[NB];
  bool ok[NB];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int b = tid + u * NT;
    int g, j;
    if constexpr (GF) {
      j = b / G;
      g = b - j * G;
    } else {
      g = b / LR;
      j = b - g * LR;
    }
    gs[u] = g;
    js[u] = j;
    ok[u] = (NBT % NT == 0 || b < NBT) && g < gcount;
    if (ok[u]) {
      if constexpr (SRC_HBM) {
        const double2 *s = hin.base + (int64_t)g * hin.gstride + (int64_t)j * hin.stride;
        const int64_t st = (int64_t)LR * hin.stride;
        if (hin.cached) {
#pragma unroll
          for (int r = 0; r < R; ++r) x[u][r] = s[r * st];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) x[u][r] = ld_stream(s + r * st);
        }
      } else {
        const double2 *s = buf + g * LL;
#pragma unroll
        for (int r = 0; r < R; ++r) x[u][r] = s[lpad(j + r * LR)];
      }
    }
  }
  if constexpr (!SRC_HBM && !DST_HBM) __syncthreads();  // all reads before the in-place writes
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (!ok[u]) continue;
    const int g = gs[u], j = js[u];
    const int bb = j / NSL;
    const int k = j - bb * NSL;
    if (twd) {
      const double2 *tp = twp + (kb0 + g) * kb_u + (int64_t)k * pa.kmul;
#pragma unroll
      for (int r = 1; r < R; ++r) x[u][r] = twmul(tp[(r - 1) * pa.twns], x[u][r]);
    }
    Bfly<FWD, R, false>::run(x[u]);
    const int e0 = bb * NSL * R + k;
    if constexpr (DST_HBM) {
      double2 *d = hout.base + (int64_t)g * hout.gstride + (int64_t)e0 * hout.stride;
      const int64_t st = (int64_t)NSL * hout.stride;
      if (hout.cached) {
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * st] = x[u][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) st_stream(d + r * st, x[u][r]);
      }
    } else {
      double2 *d = buf + g * LL;
#pragma unroll
      for (int r = 0; r < R; ++r) d[lpad(e0 + r * NSL)] = x[u][r];
    }
  }
  if constexpr (!DST_HBM) __syncthreads();
}

// Pass P of the stage (NSL: product of the stage's earlier radices).
template <bool FWD, int L, int G, int NT, int GFI, int GFO, int IO, int P, int NSL, int R,
          int... Rest>
__device__ __forceinline__ void tile_passes(double2 *buf, const HbmSide &hin, const HbmSide &hout,
                                            int tid, int gcount, const StageArgs &a,
                                            const double2 *__restrict__ tw, int64_t kb0) {
  constexpr bool FIRST = P == 0 && !IO, LAST = sizeof...(Rest) == 0 && !IO;
  constexpr bool GF = FIRST ? GFI != 0 : (LAST ? GFO != 0 : false);
  tile_pass<FWD, R, L, G, NT, NSL, FIRST, LAST, GF>(buf, hin, hout, tid, gcount, tw, a.pass[P],
                                                    kb0, a.kb_u);
  if constexpr (sizeof...(Rest) > 0)
    tile_passes<FWD, L, G, NT, GFI, GFO, IO, P + 1, NSL * R, Rest...>(buf, hin, hout, tid, gcount,
                                                                       a, tw, kb0);
}

// Coalesced copies of a tile of contiguous sub-problems (sub-problem g at g * L) between HBM
// and LDS: consecutive lanes take consecutive elements. With the real-data maps of a staged
// tile (contiguous transforms of length L, rows of in_g / out_g elements of the arrays' own
// types, `ib` / `ob` the tile's first row): io_in 1 reads doubles as (x, 0), 2 the Hermitian
// extension of L/2 + 1 complex values; io_out 1 stores real parts, 2 the first L/2 + 1 values.
template <int L, int G, int NT, bool MAPS>
__device__ __forceinline__ void tile_in(double2 *buf, const HbmSide &h, int tid, int gcount,
                                        const StageArgs &a, const void *ib) {
  constexpr int EP = (G * L + NT - 1) / NT, LL = lds_len(L);
  const int total = gcount * L;
  double2 v[EP];
  if (!MAPS || a.io_in == 0) {
#pragma unroll
    for (int u = 0; u < EP; ++u) {
      const int f = tid + u * NT;
      if (f < total) v[u] = h.cached ? h.base[f] : ld_stream(h.base + f);
    }
  } else if (a.io_in == 1) {
    const double *src = static_cast<const double *>(ib);
#pragma unroll
    for (int u = 0; u < EP; ++u) {
      const int f = tid + u * NT;
      if (f < total) {
        const int g = f / L;
        v[u] = make_double2(src[g * a.in_g + (f - g * L)], 0.0);
      }
    }
  } else {
    const double2 *src = static_cast<const double2 *>(ib);
#pragma unroll
    for (int u = 0; u < EP; ++u) {
      const int f = tid + u * NT;
      if (f < total) {
        const int g = f / L, e = f - g * L;
        const bool mirror = e > L / 2;
        v[u] = src[g * a.in_g + (mirror ? L - e : e)];
        if (mirror) v[u].y = -v[u].y;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < EP; ++u) {
    const int f = tid + u * NT;
    if (f < total) {
      const int g = f / L;
      buf[g * LL + lpad(f - g * L)] = v[u];
    }
  }
  __syncthreads();
}

template <int L, int G, int NT, bool MAPS>
__device__ __forceinline__ void tile_out(const double2 *buf, const HbmSide &h, int tid,
                                         int gcount, const StageArgs &a, void *ob) {
  constexpr int EP = (G * L + NT - 1) / NT, LL = lds_len(L);
  const int total = gcount * L;
#pragma unroll
  for (int u = 0; u < EP; ++u) {
    const int f = tid + u * NT;
    if (f < total) {
      const int g = f / L, e = f - g * L;
      const double2 v = buf[g * LL + lpad(e)];
      if (!MAPS || a.io_out == 0) {
        if (h.cached)
          h.base[f] = v;
        else
          st_stream(h.base + f, v);
      } else if (a.io_out == 1) {
        __builtin_nontemporal_store(v.x, static_cast<double *>(ob) + g * a.out_g + e);
      } else if (e <= L / 2) {
        st_stream(static_cast<double2 *>(ob) + g * a.out_g + e, v);
      }
    }
  }
}

template <bool FWD, int L, int G, int NT, int GFI, int GFO, int IO, int... Rs>
__global__ __launch_bounds__(NT) void tile_kernel(const double2 *in, double2 *out,
                                                  const double2 *__restrict__ tw,
                                                  const StageArgs a) {
  extern __shared__ double2 lds[];  // G padded sub-problems
  int bx, by;
  if (!stage_block(a.xcd, a.nsub, G, bx, by)) return;
  const int g0 = bx * G;
  const int64_t h = by;
  const int gcount = (int)(a.nsub - g0 < G ? a.nsub - g0 : G);
  const HbmSide hin{const_cast<double2 *>(in) + h * a.in_h + g0 * a.in_g, nullptr, a.in_g, a.in_e,
                    nullptr, 0, 0, a.cached_in};
  const HbmSide hout{out + h * a.out_h + g0 * a.out_g, nullptr, a.out_g, a.out_e, nullptr, 0, 0,
                     a.cached_out};
  const int tid = threadIdx.x;
  constexpr bool MAPS = IO == 2;
  if constexpr (IO != 0) {
    // the real-data maps address the arrays in their own element types
    const size_t isz = MAPS && a.io_in == 1 ? sizeof(double) : sizeof(double2);
    const void *ib = reinterpret_cast<const char *>(in) + (h * a.in_h + (int64_t)g0 * a.in_g) * isz;
    tile_in<L, G, NT, MAPS>(lds, hin, tid, gcount, a, ib);
  }
  tile_passes<FWD, L, G, NT, GFI, GFO, IO, 0, 1, Rs...>(lds, hin, hout, tid, gcount, a,
                                                        tw + h * a.kb_h, g0);
  if constexpr (IO != 0) {
    const size_t osz = MAPS && a.io_out == 1 ? sizeof(double) : sizeof(double2);
    void *ob = reinterpret_cast<char *>(out) + (h * a.out_h + (int64_t)g0 * a.out_g) * osz;
    tile_out<L, G, NT, MAPS>(lds, hout, tid, gcount, a, ob);
  }
}

// The fixed stage shapes: sub-problem length, radix sequence, lane order per side, tile copies,
// and the tile (G sub-problems, NT threads). The shapes are those of the default plans of the
// benchmarked sizes; each G / NT is the best of the variants measured on MI355X (256 MB
// batches; per launch): 6 points G = 256 89 us / G = 341 95; 16 points G = 128 87 / 256 95;
// 2100 points G = 1 of 256 threads 89 / G = 2 of 512 97; 270 and 280 points 256 threads ahead
// of 512 (75600: 230 / 246 us); 375 and 441 points 512 threads ahead of 256, and G = 8 ahead of
// G = 4 (165375: 317 / 401 us). Reading the small sizes butterfly by butterfly instead of
// staging the tile costs 1.1x (2 points) to 7x (16 points).
struct TileSchedule {
  int L, npass;
  int rs[6];
  int gfi, gfo, io;
  int G, NT;
  const void *k[2];
};

#define MFFT_T(LL, NP, RS, FI, FO, IO, GG, NN, ...)                                      \
  {LL, NP, RS, FI, FO, IO, GG, NN,                                                       \
   {(const void *)tile_kernel<false, LL, GG, NN, FI, FO, IO, __VA_ARGS__>,               \
    (const void *)tile_kernel<true, LL, GG, NN, FI, FO, IO, __VA_ARGS__>}}
#define RS_(...) {__VA_ARGS__}
const TileSchedule kTiles[] = {
    // single-stage contiguous batches of small transforms (the tile staged through LDS)
    MFFT_T(2, 1, RS_(2), 0, 0, 1, 1024, 256, 2),
    MFFT_T(4, 1, RS_(4), 0, 0, 1, 512, 256, 4),
    MFFT_T(8, 1, RS_(8), 0, 0, 1, 256, 256, 8),
    MFFT_T(16, 1, RS_(16), 0, 0, 1, 128, 256, 16),
    MFFT_T(6, 1, RS_(6), 0, 0, 1, 256, 256, 6),
    // the complex parts of real-data plans up to 64 points: staged tiles whose copies take
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
    // the default plain schedule of 1024 points (radix 16: one LDS pass fewer); a single
    // transform takes a tile of one
    MFFT_T(1024, 3, RS_(16, 8, 8), 0, 0, 0, 2, 256, 16, 8, 8),
    MFFT_T(1024, 3, RS_(16, 8, 8), 0, 0, 0, 1, 256, 16, 8, 8),
    MFFT_T(60, 2, RS_(12, 5), 0, 0, 0, 34, 256, 12, 5),
    MFFT_T(360, 3, RS_(12, 10, 3), 0, 0, 0, 5, 256, 12, 10, 3),
    MFFT_T(1000, 3, RS_(10, 10, 10), 0, 0, 0, 2, 256, 10, 10, 10),
    MFFT_T(2100, 3, RS_(15, 14, 10), 0, 0, 0, 1, 256, 15, 14, 10),
    // stages of two-stage plans: A reads rows of G adjacent columns and writes each column
    // contiguously; B reads and writes rows of G adjacent residue classes
    MFFT_T(96, 2, RS_(12, 8), 1, 0, 0, 21, 256, 12, 8),
    MFFT_T(108, 2, RS_(12, 9), 1, 1, 0, 18, 256, 12, 9),
    MFFT_T(270, 3, RS_(10, 9, 3), 1, 0, 0, 8, 256, 10, 9, 3),
    MFFT_T(280, 3, RS_(10, 7, 4), 1, 1, 0, 8, 256, 10, 7, 4),
    MFFT_T(375, 4, RS_(5, 5, 5, 3), 1, 0, 0, 8, 512, 5, 5, 5, 3),
    MFFT_T(441, 3, RS_(9, 7, 7), 1, 1, 0, 8, 512, 9, 7, 7),
    // batches of adjacent transforms (the columns of a row-major array: lanes walk the
    // transforms on both sides), and the stages of two-stage column plans (2-D 64 x 64:
    // 173 -> 167 us, 4096 x 4096: 276 -> 267 per 2^24 elements; a tile for the 2048-point
    // columns, one workgroup per CU either way, measured the same as the stage kernel)
    MFFT_T(64, 2, RS_(8, 8), 1, 1, 0, 32, 256, 8, 8),
};
#undef RS_
#undef MFFT_T
constexpr int kNumTiles = (int)(sizeof(kTiles) / sizeof(kTiles[0]));

// Whether a launch of this stage remaps its workgroups by XCD: stages whose twiddle base
// depends on the sub-problem (kb_u != 0, the second stage of a two-stage plan, whose twiddles
// span the whole table) and whose rows span at least 8 tiles. MI355X: 75600 points 233 ->
// 218 us, 165375 points 331 -> 313 per 256 MB batch (MFFT_XCD=0 turns it off).
int xcd_remap(const StageArgs &a) {
  static const int on = [] {
    const char *e = std::getenv("MFFT_XCD");
    return e ? std::atoi(e) : 1;
  }();
  return on && a.kb_u != 0 && a.nsub >= 8 * (int64_t)a.G ? 1 : 0;
}

// The tile kernel of a stage (1 + index, 0 if none; MFFT_TILE=0 turns them off): same length,
// radices, lane orders and tile copies, and at least one full tile of sub-problems. Register
// stage kernels (reg 2) keep their own kernel.
int tile_lookup(const StageArgs &a) {
  static const int on = [] {
    const char *e = std::getenv("MFFT_TILE");
    return e ? std::atoi(e) : 1;
  }();
  if (!on || a.reg == 2) return 0;
  for (int i = 0; i < kNumTiles; ++i) {
    const TileSchedule &t = kTiles[i];
    if (t.L != a.L || t.npass != a.npass || t.gfi != (a.gfast_in ? 1 : 0) ||
        t.gfo != (a.gfast_out ? 1 : 0) || t.io != a.stage_io || a.nsub < t.G)
      continue;
    bool same = true;
    for (int p = 0; p < a.npass; ++p) same = same && t.rs[p] == a.pass[p].R;
    if (same) return i + 1;
  }
  return 0;
}

hipError_t launch_tile(bool fwd, const StageArgs &a, int64_t ntrans, const double2 *in,
                       double2 *out, const double2 *tw, hipStream_t stream) {
  if (a.tile <= 0 || a.tile > kNumTiles) return hipErrorInvalidValue;
  const TileSchedule &t = kTiles[a.tile - 1];
  const void *k = t.k[fwd ? 1 : 0];
  {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.insert(k).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax);
  }
  StageArgs b = a;
  b.xcd = xcd_remap(a);
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
  (void)hipLaunchKernel(k, dim3((unsigned)nx, (unsigned)ntrans), dim3(t.NT), args, lds, stream);
  return hipGetLastError();
}

}  // namespace mfft
