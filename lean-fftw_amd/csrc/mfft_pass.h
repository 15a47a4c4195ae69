// One Stockham pass of a batch of sub-problems (device code shared by the stage kernels
// and the ragged persistent kernel).
//
// Pass layout inside a sub-problem of length L (LeanFFTW/Gen/Stockham.lean:12-20 and
// passWith, :88-98): butterfly j (< L/R) reads x[j + r L/R]; with k = j mod ns and
// b = j / ns it writes output r to x'[b ns R + k + r ns]. The inputs of a pass come from
// HBM (first pass of a stage) or from the LDS buffer, its outputs go to the LDS buffer or to
// HBM (last pass of a stage). Each thread reads all its butterflies' inputs before a barrier
// and writes their outputs after it, so the intermediate passes run in place in one buffer.
#pragma once
#include "bfly_gen.h"
#include "mfft_internal.h"

namespace mfft {

// LDS position of element e of a sub-problem: one padding element after every 16, so that
// butterfly outputs written at a stride of R elements (R <= 16) by consecutive lanes fall
// into different banks. A sub-problem of length L occupies lds_len(L) elements.
__device__ __forceinline__ int lpad(int e) { return e + (e >> 4); }
__host__ __device__ constexpr int lds_len(int L) { return L + (L >> 4) + 1; }

// LDS of a stage: none for one pass (HBM to HBM), one buffer of G padded sub-problems
// otherwise (the intermediate passes run in place), and always when the tile is staged.
__host__ __device__ constexpr int stage_lds_elems(int G, int L, int npass, int stage_io = 0) {
  return (npass <= 1 && !stage_io) ? 0 : G * lds_len(L);
}

// x * w for a twiddle input, exactly as the reference's general product mulGen(w, x)
// (LeanFFTW/Gen/Sym.lean:94-97): re = fma(w.re, x.re, -(w.im x.im)),
// im = fma(w.re, x.im, w.im x.re). A twiddled butterfly is this product on inputs 1..R-1
// followed by the untwiddled butterfly: no smart constructor or contraction of the
// generator sees through the product, so both DAGs round identically.
__device__ __forceinline__ double2 twmul(double2 w, double2 x) {
  return make_double2(fma(w.x, x.x, -(w.y * x.y)), fma(w.x, x.y, w.y * x.x));
}

// Streaming HBM access: transform inputs are read once and outputs written once, so they
// are loaded and stored non-temporally (the copy rate of MI355X rises from about 5.1 to
// 6.4 TB/s with this hint, and the end-of-kernel L2 write-back has less to do).
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2 *p) {
  const d2v t = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ void st_stream(double2 *p, double2 v) {
  const d2v t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<d2v *>(p));
}

// Real-data element maps fused into a transform's first loads and last stores (contiguous
// layouts; `off` is the element's index within its transform, `base` the transform's start in
// elements of the array's own type). Input: io_in 1 reads doubles as (x, 0); 2 reads the
// Hermitian extension of n/2 + 1 complex values (conj of element n - off above n/2; cached
// loads, since each half element is read twice). Output: io_out 1 stores real parts as
// doubles; 2 stores the first n/2 + 1 elements only; non-temporal stores unless cached_out.
__device__ __forceinline__ double2 load_io(const StageArgs &a, const double2 *in, int64_t base,
                                           int64_t off) {
  if (a.io_in == 1)
    return make_double2(__builtin_nontemporal_load(reinterpret_cast<const double *>(in) + base + off),
                        0.0);
  const bool mirror = off > a.io_n / 2;
  double2 z = in[base + (mirror ? a.io_n - off : off)];
  if (mirror) z.y = -z.y;
  return z;
}

__device__ __forceinline__ void store_io(const StageArgs &a, double2 *out, int64_t base,
                                         int64_t off, double2 y) {
  if (a.io_out == 1) {
    double *d = reinterpret_cast<double *>(out) + base + off;
    if (a.cached_out)
      *d = y.x;
    else
      __builtin_nontemporal_store(y.x, d);
  } else if (off <= a.io_n / 2) {
    if (a.cached_out)
      out[base + off] = y;
    else
      st_stream(out + base + off, y);
  }
}

// q = n / d for n * d < 2^32, with m = ceil(2^32 / d) (m = 0 encodes d = 1)
__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t m) {
  return m ? __umulhi(n, m) : n;
}

// Workgroup coordinates of a stage launch (sub-problem tile bx, row by). With a.xcd the grid's
// x extent is a multiple of 8 and linear workgroup id i (which runs on XCD i mod 8) takes tile
// (i mod 8) * c + (i / 8) mod c of row (i / 8) / c, c = gridDim.x / 8: a bijection of the grid
// in which every XCD keeps to its own eighth of the tiles. Returns false for the padding
// workgroups past the last tile (the whole workgroup returns before any barrier).
__device__ __forceinline__ bool stage_block(int xcd, int64_t nsub, int G, int &bx, int &by) {
  bx = blockIdx.x;
  by = blockIdx.y;
  if (xcd) {
    const uint32_t c = gridDim.x >> 3;
    const uint32_t lin = blockIdx.x + blockIdx.y * gridDim.x;
    const uint32_t idx = lin >> 3;
    bx = (int)((lin & 7u) * c + idx % c);
    by = (int)(idx / c);
  }
  return (int64_t)bx * G < nsub;
}

// Where a pass reads or writes in HBM: element e of sub-problem g of the tile is
// base[off(g) + e * stride], off(g) = list ? list[g] : g * gstride, optionally scaled by a
// real weight wt[g * wg + e * wstride] (weights are shared across the batch); on the input
// side of the kernels compiled with PS (a scaled input weight) the weight is wk * wt[...].
// Only those kernels read wk: the others keep the registers it would cost (the generic stage
// kernel of the mid radix set, 118 -> 134 VGPRs and 109 -> 128 plus spills, at the time).
// With SQ the last pass also returns the sum of |x|^2 of the elements the thread stored.
struct HbmSide {
  double2 *base;
  const int64_t *list;
  int64_t gstride, stride;
  const double *wt;
  int64_t wg, wstride;
  int cached = 0;  // ordinary loads / stores on this side instead of non-temporal ones
  double wk = 1.0;
};

__device__ __forceinline__ int64_t hbm_off(const HbmSide &s, int g) {
  return s.list ? s.list[g] : (int64_t)g * s.gstride;
}

// One pass; src_hbm / dst_hbm: the inputs come from HBM / the outputs go to HBM (else the LDS
// buffer). pass_run fixes them at compile time; the ragged kernel passes them at run time, so
// that it holds one copy of each radix's pass instead of four.
template <bool FWD, int R, int EPT, bool SQ, bool PS = false>
__device__ __forceinline__ double pass_body(double2 *buf, const HbmSide &hin, const HbmSide &hout,
                                          int tid, int nthreads, int G, int gcount, int L,
                                          const PassArgs &pa, bool gfast,
                                          const double2 *__restrict__ tw, int64_t kb0,
                                          int64_t kb_u, const bool SRC_HBM, const bool DST_HBM) {
  // every butterfly of the pass, over all threads: a thread holds NB of them in registers
  // (reads, then barrier, then writes), so the pass runs in place in one LDS buffer. A
  // workgroup holds at most EPT * nthreads elements.
  constexpr int NB = (EPT + R - 1) / R;
  const int LR = L / R;
  const int nb_total = G * LR;
  const bool twd = pa.twoff >= 0;
  const double2 *twp = tw + (twd ? pa.twoff : 0);
  const int ll = lds_len(L);
  double2 x[NB][R];
  int gs[NB], js[NB];
  bool ok[NB];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int b = tid + u * nthreads;
    int g, j;
    if (gfast) {
      j = (int)fdiv((uint32_t)b, pa.m_G);
      g = b - j * G;
    } else {
      g = (int)fdiv((uint32_t)b, pa.m_LR);
      j = b - g * LR;
    }
    gs[u] = g;
    js[u] = j;
    ok[u] = b < nb_total && g < gcount;
    if (ok[u]) {
      if (SRC_HBM) {
        const double2 *s = hin.base + hbm_off(hin, g) + (int64_t)j * hin.stride;
        const int64_t st = (int64_t)LR * hin.stride;
        if (hin.cached) {
#pragma unroll
          for (int r = 0; r < R; ++r) x[u][r] = s[r * st];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) x[u][r] = ld_stream(s + r * st);
        }
        if (hin.wt) {
          const double *w = hin.wt + (int64_t)g * hin.wg + (int64_t)j * hin.wstride;
          const int64_t ws = (int64_t)LR * hin.wstride;
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const double v = PS ? hin.wk * w[r * ws] : w[r * ws];
            x[u][r].x = x[u][r].x * v;
            x[u][r].y = x[u][r].y * v;
          }
        }
      } else {
        const double2 *s = buf + g * ll;
#pragma unroll
        for (int r = 0; r < R; ++r) x[u][r] = s[lpad(j + r * LR)];
      }
    }
  }
  if (!SRC_HBM && !DST_HBM) __syncthreads();  // all reads before the in-place writes
  double sumsq = 0.0;  // sum of |output|^2 over this thread's HBM stores, when requested
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (!ok[u]) continue;
    const int g = gs[u], j = js[u];
    const int bb = (int)fdiv((uint32_t)j, pa.m_nsl);
    const int k = j - bb * pa.nsl;
    if (twd) {
      const double2 *tp = twp + (kb0 + g) * kb_u + (int64_t)k * pa.kmul;
#pragma unroll
      for (int r = 1; r < R; ++r) x[u][r] = twmul(tp[(r - 1) * pa.twns], x[u][r]);
    }
    Bfly<FWD, R, false>::run(x[u]);
    const int e0 = bb * pa.nsl * R + k;
    if (DST_HBM) {
      double2 *d = hout.base + hbm_off(hout, g) + (int64_t)e0 * hout.stride;
      const int64_t st = (int64_t)pa.nsl * hout.stride;
      if (hout.wt) {
        const double *w = hout.wt + (int64_t)g * hout.wg + (int64_t)e0 * hout.wstride;
        const int64_t ws = (int64_t)pa.nsl * hout.wstride;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double v = w[r * ws];
          x[u][r].x = x[u][r].x * v;
          x[u][r].y = x[u][r].y * v;
        }
      }
      if (hout.cached) {
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * st] = x[u][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) st_stream(d + r * st, x[u][r]);
      }
      if constexpr (SQ) {
#pragma unroll
        for (int r = 0; r < R; ++r) sumsq = sumsq + (x[u][r].x * x[u][r].x + x[u][r].y * x[u][r].y);
      }
    } else {
      double2 *d = buf + g * ll;
#pragma unroll
      for (int r = 0; r < R; ++r) d[lpad(e0 + r * pa.nsl)] = x[u][r];
    }
  }
  if (!DST_HBM) __syncthreads();
  return sumsq;
}

template <bool FWD, int R, bool SRC_HBM, bool DST_HBM, int EPT, bool SQ, bool PS = false>
__device__ __forceinline__ double pass_run(double2 *buf, const HbmSide &hin, const HbmSide &hout,
                                         int tid, int nthreads, int G, int gcount, int L,
                                         const PassArgs &pa, bool gfast,
                                         const double2 *__restrict__ tw, int64_t kb0,
                                         int64_t kb_u) {
  return pass_body<FWD, R, EPT, SQ, PS>(buf, hin, hout, tid, nthreads, G, gcount, L, pa, gfast, tw,
                                    kb0, kb_u, SRC_HBM, DST_HBM);
}

// Coalesced copy of a tile of contiguous sub-problems between HBM and LDS: consecutive lanes
// take consecutive elements f = g * L + i (sub-problem g at g * gstride, gstride == L).
template <int EPT, bool PS = false>
__device__ __forceinline__ void tile_to_lds(double2 *buf, const HbmSide &h, int tid, int nthreads,
                                            int gcount, int L, uint32_t m_L) {
  const int total = gcount * L, ll = lds_len(L);
  double2 v[EPT];
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int f = tid + u * nthreads;
    if (f < total) {
      v[u] = ld_stream(h.base + f);
      if (h.wt) {
        const int g = (int)fdiv((uint32_t)f, m_L);
        const double w = PS ? h.wk * h.wt[(int64_t)g * h.wg + (int64_t)(f - g * L) * h.wstride]
                            : h.wt[(int64_t)g * h.wg + (int64_t)(f - g * L) * h.wstride];
        v[u].x = v[u].x * w;
        v[u].y = v[u].y * w;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int f = tid + u * nthreads;
    if (f < total) {
      const int g = (int)fdiv((uint32_t)f, m_L);
      buf[g * ll + lpad(f - g * L)] = v[u];
    }
  }
  __syncthreads();
}

template <int EPT, bool SQ>
__device__ __forceinline__ double tile_from_lds(const double2 *buf, const HbmSide &h, int tid,
                                                int nthreads, int gcount, int L, uint32_t m_L) {
  const int total = gcount * L, ll = lds_len(L);
  double sumsq = 0.0;
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int f = tid + u * nthreads;
    if (f < total) {
      const int g = (int)fdiv((uint32_t)f, m_L);
      double2 v = buf[g * ll + lpad(f - g * L)];
      if (h.wt) {
        const double w = h.wt[(int64_t)g * h.wg + (int64_t)(f - g * L) * h.wstride];
        v.x = v.x * w;
        v.y = v.y * w;
      }
      if (h.cached)
        h.base[f] = v;
      else
        st_stream(h.base + f, v);
      if constexpr (SQ) sumsq = sumsq + (v.x * v.x + v.y * v.y);
    }
  }
  return sumsq;
}

// All passes of a stage: HBM -> (LDS ->)* HBM, or with `stage_io` the tile is first copied
// to LDS and finally back to HBM by coalesced tile copies (for contiguous small
// sub-problems, whose butterfly reads would otherwise leave lanes far apart). Twiddle index
// of pass p for local index k of sub-problem g: (kb0 + g) * kb_u + k * kmul.
template <bool FWD, int RSET, int EPT, bool SQ = false, bool PS = false>
__device__ __forceinline__ double run_stage(double2 *buf, const HbmSide &hin, const HbmSide &hout,
                                          int tid, int nthreads, int G, int gcount, int L,
                                          int npass, const PassArgs *pas, int gfast_in,
                                          int gfast_out, const double2 *__restrict__ tw,
                                          int64_t kb0, int64_t kb_u, int stage_io = 0,
                                          uint32_t m_L = 0) {
  double sumsq = 0.0;
  if (stage_io) tile_to_lds<EPT, PS>(buf, hin, tid, nthreads, gcount, L, m_L);
  for (int p = 0; p < npass; ++p) {
    const PassArgs pa = pas[p];
    const bool first = p == 0 && !stage_io, last = p == npass - 1 && !stage_io;
    const bool gfast = first ? gfast_in : (last ? gfast_out : false);
#define MFFT_CASE(RR)                                                                            \
  case RR:                                                                                       \
    if (!in_set(RSET, RR)) break;                                                                \
    if (first && last)                                                                           \
      sumsq = pass_run<FWD, RR, true, true, EPT, SQ, PS>(buf, hin, hout, tid, nthreads, G, gcount, L, \
        pa, gfast, tw,  \
                                     kb0, kb_u);                                             \
    else if (first)                                                                              \
      pass_run<FWD, RR, true, false, EPT, SQ, PS>(buf, hin, hout, tid, nthreads, G, gcount, L, \
        pa, gfast, tw,  \
                                     kb0, kb_u);                                     \
    else if (last)                                                                               \
      sumsq = pass_run<FWD, RR, false, true, EPT, SQ, PS>(buf, hin, hout, tid, nthreads, G, gcount, L, \
        pa, gfast, tw,  \
                                     kb0, kb_u);                                     \
    else                                                                                         \
      pass_run<FWD, RR, false, false, EPT, SQ, PS>(buf, hin, hout, tid, nthreads, G, gcount, L, \
        pa, gfast, tw,  \
                                     kb0, kb_u);                                    \
    break;
    switch (pa.R) {
      MFFT_CASE(2)
      MFFT_CASE(3)
      MFFT_CASE(4)
      MFFT_CASE(5)
      MFFT_CASE(6)
      MFFT_CASE(7)
      MFFT_CASE(8)
      MFFT_CASE(9)
      MFFT_CASE(10)
      MFFT_CASE(11)
      MFFT_CASE(12)
      MFFT_CASE(13)
      MFFT_CASE(14)
      MFFT_CASE(15)
      MFFT_CASE(16)
      default:
        break;
    }
#undef MFFT_CASE
  }
  if (stage_io) sumsq = tile_from_lds<EPT, SQ>(buf, hout, tid, nthreads, gcount, L, m_L);
  return sumsq;
}

// All passes of a stage as run_stage, with one pass body per radix whose HBM / LDS sides are
// chosen at run time (a quarter of the code, and fewer registers, than a copy per side
// combination).
template <bool FWD, int RSET, int EPT, bool SQ = false>
__device__ __forceinline__ double run_stage_compact(double2 *buf, const HbmSide &hin,
                                                   const HbmSide &hout, int tid, int nthreads,
                                                   int G, int gcount, int L, int npass,
                                                   const PassArgs *pas, int gfast_in,
                                                   int gfast_out, const double2 *__restrict__ tw,
                                                   int64_t kb0, int64_t kb_u, int stage_io = 0,
                                                   uint32_t m_L = 0) {
  double sumsq = 0.0;
  if (stage_io) tile_to_lds<EPT>(buf, hin, tid, nthreads, gcount, L, m_L);
  for (int p = 0; p < npass; ++p) {
    const PassArgs pa = pas[p];
    const bool first = p == 0 && !stage_io, last = p == npass - 1 && !stage_io;
    const bool gfast = first ? gfast_in : (last ? gfast_out : false);
#define MFFT_CASE(RR)                                                                          \
  case RR:                                                                                     \
    if (in_set(RSET, RR))                                                                      \
      sumsq = pass_body<FWD, RR, EPT, SQ>(buf, hin, hout, tid, nthreads, G, gcount, L, pa,     \
                                          gfast, tw, kb0, kb_u, first, last);                  \
    break;
    switch (pa.R) {
      MFFT_CASE(2)
      MFFT_CASE(3)
      MFFT_CASE(4)
      MFFT_CASE(5)
      MFFT_CASE(6)
      MFFT_CASE(7)
      MFFT_CASE(8)
      MFFT_CASE(9)
      MFFT_CASE(10)
      MFFT_CASE(11)
      MFFT_CASE(12)
      MFFT_CASE(13)
      MFFT_CASE(14)
      MFFT_CASE(15)
      MFFT_CASE(16)
      default:
        break;
    }
#undef MFFT_CASE
  }
  if (stage_io) sumsq = tile_from_lds<EPT, SQ>(buf, hout, tid, nthreads, gcount, L, m_L);
  return sumsq;
}

}  // namespace mfft
