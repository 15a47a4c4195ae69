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
// with the real-data maps of r2c / c2r plans in the tile copies, 3 read and written by the
// first and last passes with the real-data maps in those accesses (the stages of two-stage
// real-data plans: stage A takes the input map, stage B the output map).
#pragma once

#include "mfft_pass.h"

namespace mfft {

// LDS position of element e of sub-problem g: sub-problems contiguous and padded (CST = 0), or,
// for the columns of a 2-D array held row-major in LDS, row e at e * CST and column g at lpad(g).
template <int LL, int CST>
__device__ __forceinline__ int lds_pos(int g, int e) {
  if constexpr (CST == 0)
    return g * LL + lpad(e);
  else
    return e * CST + lpad(g);
}

// The real-data maps of an unstaged tile (IO 3): the raw arrays, the transform's start in
// elements of each array's own type, and the tile's first sub-problem offset on each side, so
// that an access's index within its transform is g0 offset + its offset from the tile's base.
struct TileMaps {
  const StageArgs *a;
  const double2 *in;
  double2 *out;
  int64_t bin, bout, gin0, gout0;
};

template <bool FWD, int R, int L, int G, int NT, int NSL, bool SRC_HBM, bool DST_HBM, bool GF,
          int CST = 0, bool MAPI = false, bool MAPO = false>
__device__ __forceinline__ void tile_pass(double2 *buf, const HbmSide &hin, const HbmSide &hout,
                                          int tid, int gcount, const double2 *__restrict__ tw,
                                          const PassArgs &pa, int64_t kb0, int64_t kb_u,
                                          const TileMaps *maps = nullptr) {
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
      if constexpr (SRC_HBM && MAPI) {
        if (maps->a->io_in) {
          const int64_t o = maps->gin0 + (int64_t)g * hin.gstride + (int64_t)j * hin.stride;
          const int64_t st = (int64_t)LR * hin.stride;
#pragma unroll
          for (int r = 0; r < R; ++r) x[u][r] = load_io(*maps->a, maps->in, maps->bin, o + r * st);
          continue;
        }
      }
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
#pragma unroll
        for (int r = 0; r < R; ++r) x[u][r] = buf[lds_pos<LL, CST>(g, j + r * LR)];
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
    if constexpr (DST_HBM && MAPO) {
      if (maps->a->io_out) {
        const int64_t o = maps->gout0 + (int64_t)g * hout.gstride + (int64_t)e0 * hout.stride;
        const int64_t st = (int64_t)NSL * hout.stride;
#pragma unroll
        for (int r = 0; r < R; ++r) store_io(*maps->a, maps->out, maps->bout, o + r * st, x[u][r]);
        continue;
      }
    }
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
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lds_pos<LL, CST>(g, e0 + r * NSL)] = x[u][r];
    }
  }
  if constexpr (!DST_HBM) __syncthreads();
}

// Pass P of the stage (NSL: product of the stage's earlier radices).
template <bool FWD, int L, int G, int NT, int GFI, int GFO, int IO, int P, int NSL, int R,
          int... Rest>
__device__ __forceinline__ void tile_passes(double2 *buf, const HbmSide &hin, const HbmSide &hout,
                                            int tid, int gcount, const StageArgs &a,
                                            const double2 *__restrict__ tw, int64_t kb0,
                                            const TileMaps *maps) {
  constexpr bool STAGED = IO == 1 || IO == 2;
  constexpr bool FIRST = P == 0 && !STAGED, LAST = sizeof...(Rest) == 0 && !STAGED;
  constexpr bool GF = FIRST ? GFI != 0 : (LAST ? GFO != 0 : false);
  tile_pass<FWD, R, L, G, NT, NSL, FIRST, LAST, GF, 0, FIRST && IO == 3, LAST && IO == 3>(
      buf, hin, hout, tid, gcount, tw, a.pass[P], kb0, a.kb_u, maps);
  if constexpr (sizeof...(Rest) > 0)
    tile_passes<FWD, L, G, NT, GFI, GFO, IO, P + 1, NSL * R, Rest...>(buf, hin, hout, tid, gcount,
                                                                       a, tw, kb0, maps);
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
  // IO 3: the maps in the first / last passes' accesses, which address the arrays in their
  // own element types from the transform's start
  TileMaps tm{&a, in, out, h * a.in_h, h * a.out_h, (int64_t)g0 * a.in_g, (int64_t)g0 * a.out_g};
  if constexpr (IO == 1 || IO == 2) {
    // the real-data maps address the arrays in their own element types
    const size_t isz = MAPS && a.io_in == 1 ? sizeof(double) : sizeof(double2);
    const void *ib = reinterpret_cast<const char *>(in) + (h * a.in_h + (int64_t)g0 * a.in_g) * isz;
    tile_in<L, G, NT, MAPS>(lds, hin, tid, gcount, a, ib);
  }
  tile_passes<FWD, L, G, NT, GFI, GFO, IO, 0, 1, Rs...>(lds, hin, hout, tid, gcount, a,
                                                        tw + h * a.kb_h, g0, &tm);
  if constexpr (IO == 1 || IO == 2) {
    const size_t osz = MAPS && a.io_out == 1 ? sizeof(double) : sizeof(double2);
    void *ob = reinterpret_cast<char *>(out) + (h * a.out_h + (int64_t)g0 * a.out_g) * osz;
    tile_out<L, G, NT, MAPS>(lds, hout, tid, gcount, a, ob);
  }
}

// A fixed stage shape: sub-problem length, radix sequence, lane order per side, tile copies
// (IO), and the tile (G sub-problems, NT threads), with its kernels by direction.
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

// The generated tables (mfft_tile_gen*.hip, written by lean-fftw_amd/tilegen.py): every
// 13-smooth size up to 4096 points that is not a power of two, with the default plan's radices.
constexpr int kTileGenParts = 8;
extern const TileSchedule *const kTileGen[kTileGenParts];
extern const int kTileGenCount[kTileGenParts];

}  // namespace mfft
