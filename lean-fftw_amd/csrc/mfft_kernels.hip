// Stockham stage kernels for gfx950: FP64 complex, LDS-resident passes.
//
// A workgroup loads G sub-problems of length L into LDS (coalesced: consecutive lanes walk
// whichever of the element or sub-problem index is adjacent in memory), runs the stage's
// Stockham passes in place in LDS -- each pass is the reference's passWith
// (LeanFFTW/Gen/Stockham.lean:88-98) restricted to the sub-problem, with the butterflies
// of csrc/bfly_gen.h -- and stores the result. All arithmetic is FP64 vector FMA; nothing
// here is a dense contraction, so MFMA is not used.
//
#include "mfft_pass.h"

namespace mfft {

template <bool FWD>
__global__ __launch_bounds__(WG_THREADS) void stage_kernel(const double2 *in, double2 *out,
                                                           const double2 *__restrict__ tw,
                                                           const StageArgs a) {
  extern __shared__ double2 lds[];  // two buffers of G*L elements
  const int tid = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * a.G;
  const int gcount = (int)(a.nsub - s0 < a.G ? a.nsub - s0 : a.G);
  const int L = a.L;
  const int total = gcount * L;

  for (int f = tid; f < total; f += WG_THREADS) {
    int g, i;
    if (a.gfast_in) {
      i = f / gcount;
      g = f - i * gcount;
    } else {
      g = f / L;
      i = f - g * L;
    }
    const int64_t s = s0 + g;
    const int64_t h = s / a.U, u = s - h * a.U;
    double2 v = in[h * a.idist + u * a.in_u + (int64_t)i * a.in_i];
    if (a.pre) {
      const double wt = a.pre[u * a.pre_u + (int64_t)i * a.pre_i];
      v.x = v.x * wt;
      v.y = v.y * wt;
    }
    lds[g * L + i] = v;
  }
  __syncthreads();

  double2 *const buf0 = lds, *const buf1 = lds + a.G * L;
  for (int p = 0; p < a.npass; ++p) {
    const bool twd = a.twoff[p] >= 0;
    const double2 *twp = tw + (twd ? a.twoff[p] : 0);
    const int nsl = a.nsl[p];
    const int64_t twns = a.twns[p];
#define MFFT_PASS(RR)                                                                       \
  case RR:                                                                                  \
    pass_lds<FWD, RR>((p & 1) ? buf1 : buf0, (p & 1) ? buf0 : buf1, tid, gcount, L, nsl, twp, twd, twns,  \
                      a.kmul, a.kb_u, s0, a.U);                                             \
    break;
    switch (a.R[p]) {
      MFFT_PASS(2)
      MFFT_PASS(3)
      MFFT_PASS(4)
      MFFT_PASS(5)
      MFFT_PASS(6)
      MFFT_PASS(7)
      MFFT_PASS(8)
      MFFT_PASS(9)
      MFFT_PASS(10)
      MFFT_PASS(11)
      MFFT_PASS(12)
      MFFT_PASS(13)
      MFFT_PASS(14)
      MFFT_PASS(15)
      MFFT_PASS(16)
      default:
        break;
    }
#undef MFFT_PASS
  }

  for (int f = tid; f < total; f += WG_THREADS) {
    int g, i;
    if (a.gfast_out) {
      i = f / gcount;
      g = f - i * gcount;
    } else {
      g = f / L;
      i = f - g * L;
    }
    const int64_t s = s0 + g;
    const int64_t h = s / a.U, u = s - h * a.U;
    double2 v = ((a.npass & 1) ? buf1 : buf0)[g * L + i];
    if (a.post) {
      const double wt = a.post[u * a.post_u + (int64_t)i * a.post_i];
      v.x = v.x * wt;
      v.y = v.y * wt;
    }
    out[h * a.odist + u * a.out_u + (int64_t)i * a.out_i] = v;
  }
}

hipError_t launch_stage(bool fwd, const StageArgs &a, const double2 *in, double2 *out,
                        const double2 *tw, hipStream_t stream) {
  if (a.nsub <= 0) return hipSuccess;
  const int64_t nwg = (a.nsub + a.G - 1) / a.G;
  const size_t lds = 2 * (size_t)a.G * a.L * sizeof(double2);
  if (fwd)
    hipLaunchKernelGGL(stage_kernel<true>, dim3((unsigned)nwg), dim3(WG_THREADS), lds, stream,
                       in, out, tw, a);
  else
    hipLaunchKernelGGL(stage_kernel<false>, dim3((unsigned)nwg), dim3(WG_THREADS), lds, stream,
                       in, out, tw, a);
  return hipGetLastError();
}

}  // namespace mfft
