// Element maps of the composite transforms (mfft_plan_r2c / mfft_plan_c2r): the reference's
// generic engine computes them through the complex DFT (LeanFFTW/Core/Problem.lean:128-145):
//   r2c: F = DFT(x_j + 0 i), output F_0 .. F_{n/2}                  (LineOp.r2c)
//   c2r: G = backward DFT(hermExt(h)), output Re G_j                 (LineOp.c2r, hermExt :121-123)
// These kernels are the exact, arithmetic-free maps around the complex transform: embedding a
// real batch, keeping the first n/2 + 1 outputs, the Hermitian extension (conjugate mirror)
// and the real part. Streaming accesses, one element per thread and iteration.
#include "mfft_pass.h"

namespace mfft {

namespace {

constexpr int kMapThreads = 256;

unsigned map_blocks(int64_t total) {
  const int64_t b = (total + kMapThreads - 1) / kMapThreads;
  return (unsigned)(b < 65536 * 16 ? (b < 1 ? 1 : b) : 65536 * 16);
}

// work[h n + j] = (x[h n + j], 0)
__global__ __launch_bounds__(kMapThreads) void embed_real_kernel(const double *__restrict__ x,
                                                                 double2 *__restrict__ work,
                                                                 int64_t total) {
  for (int64_t i = (int64_t)blockIdx.x * kMapThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kMapThreads) {
    const double v = __builtin_nontemporal_load(x + i);
    st_stream(work + i, make_double2(v, 0.0));
  }
}

// out[h (n/2+1) + k] = work[h n + k], k <= n/2
__global__ __launch_bounds__(kMapThreads) void extract_half_kernel(const double2 *__restrict__ work,
                                                                   double2 *__restrict__ out,
                                                                   int64_t n, int64_t total) {
  const int64_t nh = n / 2 + 1;
  for (int64_t i = (int64_t)blockIdx.x * kMapThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kMapThreads) {
    const int64_t h = i / nh, k = i - h * nh;
    st_stream(out + i, ld_stream(work + h * n + k));
  }
}

// work[h n + k] = in[h (n/2+1) + k] for k <= n/2, else conj(in[h (n/2+1) + n - k])
__global__ __launch_bounds__(kMapThreads) void herm_ext_kernel(const double2 *__restrict__ in,
                                                               double2 *__restrict__ work,
                                                               int64_t n, int64_t total) {
  const int64_t nh = n / 2 + 1;
  for (int64_t i = (int64_t)blockIdx.x * kMapThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kMapThreads) {
    const int64_t h = i / n, k = i - h * n;
    double2 z;
    if (k <= n / 2) {
      z = in[h * nh + k];
    } else {
      z = in[h * nh + (n - k)];
      z.y = -z.y;
    }
    st_stream(work + i, z);
  }
}

// out[h n + j] = Re work[h n + j]
__global__ __launch_bounds__(kMapThreads) void real_part_kernel(const double2 *__restrict__ work,
                                                                double *__restrict__ out,
                                                                int64_t total) {
  for (int64_t i = (int64_t)blockIdx.x * kMapThreads + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kMapThreads) {
    __builtin_nontemporal_store(ld_stream(work + i).x, out + i);
  }
}

}  // namespace

hipError_t launch_embed_real(const double *x, double2 *work, int64_t n, int64_t howmany,
                             hipStream_t st) {
  const int64_t total = n * howmany;
  if (total <= 0) return hipSuccess;
  void *args[] = {(void *)&x, (void *)&work, (void *)&total};
  return hipLaunchKernel((const void *)embed_real_kernel, dim3(map_blocks(total)), dim3(kMapThreads),
                         args, 0, st);
}

hipError_t launch_extract_half(const double2 *work, double2 *out, int64_t n, int64_t howmany,
                               hipStream_t st) {
  const int64_t total = (n / 2 + 1) * howmany;
  if (total <= 0) return hipSuccess;
  void *args[] = {(void *)&work, (void *)&out, (void *)&n, (void *)&total};
  return hipLaunchKernel((const void *)extract_half_kernel, dim3(map_blocks(total)),
                         dim3(kMapThreads), args, 0, st);
}

hipError_t launch_herm_ext(const double2 *in, double2 *work, int64_t n, int64_t howmany,
                           hipStream_t st) {
  const int64_t total = n * howmany;
  if (total <= 0) return hipSuccess;
  void *args[] = {(void *)&in, (void *)&work, (void *)&n, (void *)&total};
  return hipLaunchKernel((const void *)herm_ext_kernel, dim3(map_blocks(total)), dim3(kMapThreads),
                         args, 0, st);
}

hipError_t launch_real_part(const double2 *work, double *out, int64_t n, int64_t howmany,
                            hipStream_t st) {
  const int64_t total = n * howmany;
  if (total <= 0) return hipSuccess;
  void *args[] = {(void *)&work, (void *)&out, (void *)&total};
  return hipLaunchKernel((const void *)real_part_kernel, dim3(map_blocks(total)), dim3(kMapThreads),
                         args, 0, st);
}

}  // namespace mfft
