// Stockham stage kernels for gfx950: FP64 complex, LDS-resident passes.
//
// A workgroup takes G sub-problems of length L. The stage's first pass reads its
// butterfly inputs straight from HBM (consecutive lanes walk whichever of the element or
// sub-problem index is adjacent in memory, so the reads coalesce), the intermediate passes
// run between two LDS buffers, and the last pass writes straight to HBM. Each pass is the
// reference's passWith (LeanFFTW/Gen/Stockham.lean:88-98) restricted to the sub-problem,
// with the butterflies of csrc/bfly_gen.h. All arithmetic is FP64 vector FMA; nothing here
// is a dense contraction, so MFMA is not used.
#include "mfft_pass.h"

namespace mfft {

// PS: input weights scaled by a device scalar (the adjoint of the differentiable path), a
// variant of its own so that the other instances keep their registers.
template <bool FWD, int RSET, int EPT, int MAXT, bool SQ, bool PS>
__global__ __launch_bounds__(MAXT) void stage_kernel(const double2 *in, double2 *out,
                                                     const double2 *__restrict__ tw,
                                                     const StageArgs a) {
  extern __shared__ double2 lds[];  // G padded sub-problems
  const int g0 = blockIdx.x * a.G;
  const int64_t h = blockIdx.y, z = blockIdx.z;
  const int gcount = (int)(a.nsub - g0 < a.G ? a.nsub - g0 : a.G);
  HbmSide hin{const_cast<double2 *>(in) + z * a.in_z + h * a.in_h + g0 * a.in_g, nullptr, a.in_g,
              a.in_e, a.pre ? a.pre + g0 * a.pre_g : nullptr, a.pre_g, a.pre_e, a.cached_in};
  HbmSide hout{out + z * a.out_z + h * a.out_h + g0 * a.out_g, nullptr, a.out_g, a.out_e,
               a.post ? a.post + g0 * a.post_g : nullptr, a.post_g, a.post_e, a.cached_out};
  if constexpr (PS) {
    if (a.pre_s) hin.wk = a.pre_k * a.pre_s[0];
  }
  // (the run-time-sided pass body of run_stage_compact measured the same here: MI355X, mixed
  // sizes on this kernel 2.91 / 2.92 TB/s)
  const double part = run_stage<FWD, RSET, EPT, SQ, PS>(lds, hin, hout, threadIdx.x, blockDim.x, a.G, gcount, a.L,
                 a.npass, a.pass, a.gfast_in, a.gfast_out, tw + h * a.kb_h, g0, a.kb_u, a.stage_io, a.m_L);
  if constexpr (SQ) {
    // deterministic per-workgroup partial of sum |out|^2: a fixed reduction tree
    __shared__ double wsum[MAXT / 64];
    double v = part;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = t + wsum[w];
      a.sumsq[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
  }
}

// the kernel instance for (direction, radix set, elements per thread, sum of squares, scaled
// input weights; sum-of-squares calls take no scaled weights)
const void *kernel_for(int fwd, int set, int ept, int sq, int big, int ps) {
#define MFFT_K(F, S, E, T)                                                                   \
  (sq ? (const void *)stage_kernel<F, S, E, T, true, false>                                 \
      : ps ? (const void *)stage_kernel<F, S, E, T, false, true>                            \
           : (const void *)stage_kernel<F, S, E, T, false, false>)
#define MFFT_DIR(S, E, T) (fwd ? MFFT_K(true, S, E, T) : MFFT_K(false, S, E, T))
  if (ept == 16) {
    if (set == RSET_POW2) return MFFT_DIR(RSET_POW2, 16, 512);
    if (set == RSET_SMALL) return MFFT_DIR(RSET_SMALL, 16, 512);
    return MFFT_DIR(RSET_ALL, 16, 512);
  }
  if (set == RSET_SMALL) return big ? MFFT_DIR(RSET_SMALL, 8, 1024) : MFFT_DIR(RSET_SMALL, 8, 256);
  if (set == RSET_MID) return big ? MFFT_DIR(RSET_MID, 8, 1024) : MFFT_DIR(RSET_MID, 8, 256);
  return MFFT_DIR(RSET_ALL, 8, 256);
#undef MFFT_DIR
#undef MFFT_K
}

hipError_t launch_stage(bool fwd, const StageArgs &a, int64_t ntrans, const double2 *in,
                        double2 *out, const double2 *tw, hipStream_t stream, int64_t nz) {
  if (a.nsub <= 0 || ntrans <= 0 || nz <= 0) return hipSuccess;
  // a z extent only on the stage kernel (no tile, register, weights or partials)
  if (nz > 1 && (a.tile > 0 || a.reg != 0 || a.sumsq || nz > 65535 || ntrans > 65535))
    return hipErrorInvalidValue;
  // the real-data element maps exist in the register kernels (input maps at 16 elements per
  // thread: real input into a schedule opening with radix 16), in the staged tile kernels and
  // in the tile kernels of two-stage real-data plans
  const bool tile_io =
      a.tile > 0 && (a.stage_io == 2 || a.stage_io == 3) && !a.pre && !a.post && !a.sumsq;
  if ((a.io_in || a.io_out) && !tile_io &&
      (a.reg == 0 || (a.reg == 2 && a.sumsq) ||
       (a.io_in && a.ept != 8 && !(a.io_in == 1 && a.reg == 1 && a.pass[0].R == 16))))
    return hipErrorInvalidValue;
  if (a.tile > 0 && !a.pre && !a.post && !a.sumsq && (tile_io || (!a.io_in && !a.io_out)))
    return launch_tile(fwd, a, ntrans, in, out, tw, stream);
  if (a.reg == 1) return ntrans == 1 ? launch_reg(fwd, a, in, out, tw, stream) : hipErrorInvalidValue;
  if (a.reg == 2 && !a.sumsq) return launch_reg_stage(fwd, a, ntrans, in, out, tw, stream);
  const int64_t nx = (a.nsub + a.G - 1) / a.G;
  const size_t lds = (size_t)stage_lds_elems(a.G, a.L, a.npass, a.stage_io) * sizeof(double2);
  const dim3 grid((unsigned)nx, (unsigned)ntrans, (unsigned)nz);
  // variants: <direction, radix set, elements per thread, max threads, sum of squares>
  bool small = true, mid = true, pow2 = true;
  for (int p = 0; p < a.npass; ++p) {
    const int r = a.pass[p].R;
    small = small && small_radix(r);
    mid = mid && mid_radix(r);
    pow2 = pow2 && (r == 2 || r == 4 || r == 8 || r == 16);
  }
  const int set = (a.ept == 16 && pow2) ? RSET_POW2
                  : small               ? RSET_SMALL
                  : (mid && a.ept == 8) ? RSET_MID
                                        : RSET_ALL;
  const void *k = kernel_for(fwd, set, a.ept, a.sumsq != nullptr, a.threads > 256,
                             a.pre && a.pre_s && !a.sumsq);
  static const bool attr = [] {
    for (int f = 0; f < 2; ++f)
      for (int st = 0; st < 4; ++st)
        for (int e = 8; e <= 16; e += 8)
          for (int q = 0; q < 2; ++q)
            for (int b = 0; b < 2; ++b)
              for (int ps = 0; ps < 2; ++ps)
                (void)hipFuncSetAttribute(kernel_for(f, st, e, q, b, ps),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax);
    return true;
  }();
  (void)attr;
  void *args[] = {(void *)&in, (void *)&out, (void *)&tw, (void *)&a};
  return hipLaunchKernel(k, grid, dim3(a.threads), args, lds, stream);
}

}  // namespace mfft
