// One Stockham pass of a sub-problem batch held in LDS (device code shared by the
// stage kernels and the ragged persistent kernel).
//
// Pass layout inside a sub-problem (LeanFFTW/Gen/Stockham.lean:12-20 and passWith,
// :88-98): butterfly j (< L/R) reads x[j + r L/R]; with k = j mod ns and b = j / ns it
// writes output r to x'[b ns R + k + r ns]. Passes alternate between two LDS buffers, so
// a thread holds one butterfly at a time.
#pragma once
#include "bfly_gen.h"
#include "mfft_internal.h"

namespace mfft {

template <bool FWD, int R>
__device__ __forceinline__ void pass_lds(const double2 *src, double2 *dst, int tid, int gcount,
                                         int L, int nsl, const double2 *__restrict__ twp,
                                         bool twiddled, int64_t twns, int64_t kmul, int64_t kb_u,
                                         int64_t s0, int64_t U) {
  const int LR = L / R;
  const int nbf = gcount * LR;
#pragma unroll 1
  for (int b = tid; b < nbf; b += WG_THREADS) {
    const int g = b / LR, j = b - g * LR;
    const int k = j % nsl, bb = j / nsl;
    double2 x[R];
    const double2 *s = src + g * L + j;
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = s[r * LR];
    if (twiddled) {
      const int64_t kg = (kb_u ? ((s0 + g) % U) : 0) + (int64_t)k * kmul;
      double2 w[R - 1];
#pragma unroll
      for (int r = 1; r < R; ++r) w[r - 1] = twp[(r - 1) * twns + kg];
      Bfly<FWD, R, true>::run(x, w);
    } else {
      Bfly<FWD, R, false>::run(x);
    }
    double2 *d = dst + g * L + bb * nsl * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * nsl] = x[r];
  }
  __syncthreads();
}

}  // namespace mfft
