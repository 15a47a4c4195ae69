// One Stockham pass of a batch of sub-problems (device code shared by the stage kernels
// and the ragged persistent kernel).
//
// Pass layout inside a sub-problem of length L (LeanFFTW/Gen/Stockham.lean:12-20 and
// passWith, :88-98): butterfly j (< L/R) reads x[j + r L/R]; with k = j mod ns and
// b = j / ns it writes output r to x'[b ns R + k + r ns]. The inputs of a pass come from
// HBM (first pass of a stage) or from an LDS buffer, its outputs go to the other LDS
// buffer or to HBM (last pass of a stage), so a thread holds one group of butterflies at a
// time and the intermediate passes alternate between two LDS buffers.
#pragma once
#include "bfly_gen.h"
#include "mfft_internal.h"

namespace mfft {

// q = n / d for n * d < 2^32, with m = ceil(2^32 / d) (m = 0 encodes d = 1)
__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t m) {
  return m ? __umulhi(n, m) : n;
}

// Where a pass reads or writes in HBM: element e of sub-problem g of the tile is
// base[off(g) + e * stride], off(g) = list ? list[g] : g * gstride, optionally scaled by a
// real weight wt[g * wg + e * wstride] (weights are shared across the batch).
struct HbmSide {
  double2 *base;
  const int64_t *list;
  int64_t gstride, stride;
  const double *wt;
  int64_t wg, wstride;
};

__device__ __forceinline__ int64_t hbm_off(const HbmSide &s, int g) {
  return s.list ? s.list[g] : (int64_t)g * s.gstride;
}

template <bool FWD, int R, bool SRC_HBM, bool DST_HBM>
__device__ __forceinline__ void pass_run(const double2 *src_lds, double2 *dst_lds,
                                         const HbmSide &hin, const HbmSide &hout, int tid,
                                         int nthreads, int G, int gcount, int L,
                                         const PassArgs &pa, bool gfast,
                                         const double2 *__restrict__ tw, int64_t kb0,
                                         int64_t kb_u) {
  const int LR = L / R;
  const int nb = G * LR;
  // several butterflies per iteration when reading HBM, so that >= 8 loads are in flight
  constexpr int UF = SRC_HBM ? (R >= 8 ? 1 : (8 + R - 1) / R) : 1;
  const bool twd = pa.twoff >= 0;
  const double2 *twp = tw + (twd ? pa.twoff : 0);
#pragma unroll 1
  for (int b0 = tid; b0 < nb; b0 += nthreads * UF) {
    double2 x[UF][R];
    int gs[UF], js[UF];
    bool ok[UF];
#pragma unroll
    for (int u = 0; u < UF; ++u) {
      const int b = b0 + u * nthreads;
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
      ok[u] = b < nb && g < gcount;
      if (ok[u]) {
        if constexpr (SRC_HBM) {
          const double2 *s = hin.base + hbm_off(hin, g) + (int64_t)j * hin.stride;
          const int64_t st = (int64_t)LR * hin.stride;
#pragma unroll
          for (int r = 0; r < R; ++r) x[u][r] = s[r * st];
          if (hin.wt) {
            const double *w = hin.wt + (int64_t)g * hin.wg + (int64_t)j * hin.wstride;
            const int64_t ws = (int64_t)LR * hin.wstride;
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const double v = w[r * ws];
              x[u][r].x = x[u][r].x * v;
              x[u][r].y = x[u][r].y * v;
            }
          }
        } else {
          const double2 *s = src_lds + g * L + j;
#pragma unroll
          for (int r = 0; r < R; ++r) x[u][r] = s[r * LR];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < UF; ++u) {
      if (!ok[u]) continue;
      const int g = gs[u], j = js[u];
      const int bb = (int)fdiv((uint32_t)j, pa.m_nsl);
      const int k = j - bb * pa.nsl;
      if (twd) {
        const int64_t kg = (kb0 + g) * kb_u + (int64_t)k * pa.kmul;
        double2 w[R - 1];
#pragma unroll
        for (int r = 1; r < R; ++r) w[r - 1] = twp[(r - 1) * pa.twns + kg];
        Bfly<FWD, R, true>::run(x[u], w);
      } else {
        Bfly<FWD, R, false>::run(x[u]);
      }
      const int e0 = bb * pa.nsl * R + k;
      if constexpr (DST_HBM) {
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
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * st] = x[u][r];
      } else {
        double2 *d = dst_lds + g * L + e0;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * pa.nsl] = x[u][r];
      }
    }
  }
  if constexpr (!DST_HBM) __syncthreads();
}

// All passes of a stage: HBM -> (LDS ->)* HBM. Twiddle index of pass p for local index k
// of sub-problem g: (kb0 + g) * kb_u + k * kmul.
template <bool FWD>
__device__ __forceinline__ void run_stage(double2 *buf0, double2 *buf1, const HbmSide &hin,
                                          const HbmSide &hout, int tid, int nthreads, int G,
                                          int gcount, int L, int npass, const PassArgs *pas,
                                          int gfast_in, int gfast_out,
                                          const double2 *__restrict__ tw, int64_t kb0,
                                          int64_t kb_u) {
  for (int p = 0; p < npass; ++p) {
    const PassArgs pa = pas[p];
    const bool first = p == 0, last = p == npass - 1;
    const double2 *src = (p & 1) ? buf1 : buf0;
    double2 *dst = (p & 1) ? buf0 : buf1;
    const bool gfast = first ? gfast_in : (last ? gfast_out : false);
#define MFFT_CASE(RR)                                                                            \
  case RR:                                                                                       \
    if (first && last)                                                                           \
      pass_run<FWD, RR, true, true>(src, dst, hin, hout, tid, nthreads, G, gcount, L, pa, gfast, \
                                    tw, kb0, kb_u);                                             \
    else if (first)                                                                              \
      pass_run<FWD, RR, true, false>(src, dst, hin, hout, tid, nthreads, G, gcount, L, pa,      \
                                     gfast, tw, kb0, kb_u);                                     \
    else if (last)                                                                               \
      pass_run<FWD, RR, false, true>(src, dst, hin, hout, tid, nthreads, G, gcount, L, pa,      \
                                     gfast, tw, kb0, kb_u);                                     \
    else                                                                                         \
      pass_run<FWD, RR, false, false>(src, dst, hin, hout, tid, nthreads, G, gcount, L, pa,     \
                                      gfast, tw, kb0, kb_u);                                    \
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
}

}  // namespace mfft
