// Single-pass 2-D transforms of small arrays (n0 * n1 <= 4096 points): one workgroup holds a
// whole row-major n0 x n1 array in LDS, runs the axis-0 passes on its columns and then the
// axis-1 passes on its rows, and writes it back, so the array crosses HBM once instead of twice
// (the 2-D plan otherwise runs its axis-0 part out of place and its axis-1 part in place on the
// output). The passes are the two parts' own: the column plan's radices and twiddle table for
// axis 0, the row plan's for axis 1, with the pass code of the tile kernels (mfft_tile.h), so
// the bits are the two-launch composition's (the reference's generic engine applies the line
// transform along axis 0, then axis 1: LeanFFTW/Core/Problem.lean:165-182).
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <mutex>
#include <set>

#include "mfft_tile.h"

namespace mfft {

template <int... Rs>
struct RList {};

// Passes P... of one axis: sub-problems of length L (G of them), LDS addressing CST (0 for
// rows, the padded row length for columns), lanes walking the sub-problem index for columns.
template <bool FWD, int L, int G, int NT, int CST, int P, int NSL, int R, int... Rest>
__device__ __forceinline__ void axis_passes(double2 *buf, int tid, const StageArgs &a,
                                            const double2 *__restrict__ tw) {
  const HbmSide none{};
  tile_pass<FWD, R, L, G, NT, NSL, false, false, CST != 0, CST>(buf, none, none, tid, G, tw,
                                                                a.pass[P], 0, a.kb_u);
  if constexpr (sizeof...(Rest) > 0)
    axis_passes<FWD, L, G, NT, CST, P + 1, NSL * R, Rest...>(buf, tid, a, tw);
}

template <bool FWD, int N0, int N1, int NT, int... R0, int... R1>
__device__ __forceinline__ void run_2d(double2 *buf, int tid, const StageArgs &col,
                                       const StageArgs &row, const double2 *__restrict__ tw0,
                                       const double2 *__restrict__ tw1, RList<R0...>,
                                       RList<R1...>) {
  constexpr int LL1 = lds_len(N1);
  axis_passes<FWD, N0, N1, NT, LL1, 0, 1, R0...>(buf, tid, col, tw0);  // columns
  axis_passes<FWD, N1, N0, NT, 0, 0, 1, R1...>(buf, tid, row, tw1);    // rows
}

template <bool FWD, int N0, int N1, int NT, class C, class W>
__global__ __launch_bounds__(NT) void tile2d_kernel(const double2 *in, double2 *out,
                                                    const double2 *__restrict__ tw0,
                                                    const double2 *__restrict__ tw1,
                                                    const StageArgs col, const StageArgs row) {
  extern __shared__ double2 lds[];  // N0 rows of lds_len(N1) elements
  constexpr int LL1 = lds_len(N1), TOT = N0 * N1, EP = (TOT + NT - 1) / NT;
  const int tid = threadIdx.x;
  const double2 *src = in + (int64_t)blockIdx.x * TOT;
  double2 *dst = out + (int64_t)blockIdx.x * TOT;
  double2 v[EP];
#pragma unroll
  for (int u = 0; u < EP; ++u) {
    const int f = tid + u * NT;
    if (TOT % NT == 0 || f < TOT) v[u] = src[f];
  }
#pragma unroll
  for (int u = 0; u < EP; ++u) {
    const int f = tid + u * NT;
    if (TOT % NT == 0 || f < TOT) lds[(f / N1) * LL1 + lpad(f % N1)] = v[u];
  }
  __syncthreads();
  run_2d<FWD, N0, N1, NT>(lds, tid, col, row, tw0, tw1, C{}, W{});
#pragma unroll
  for (int u = 0; u < EP; ++u) {
    const int f = tid + u * NT;
    if (TOT % NT == 0 || f < TOT) st_stream(dst + f, lds[(f / N1) * LL1 + lpad(f % N1)]);
  }
}

struct Tile2D {
  int n0, n1, np0, np1;
  int r0[4], r1[4];
  int NT;
  const void *k[2];
};

#define MFFT_T2(A, B, NP0, R0S, NP1, R1S, NN, C, W)                                           \
  {A, B, NP0, NP1, R0S, R1S, NN,                                                             \
   {(const void *)tile2d_kernel<false, A, B, NN, C, W>,                                      \
    (const void *)tile2d_kernel<true, A, B, NN, C, W>}}
#define RS_(...) {__VA_ARGS__}
// the default radices of the parts: 16 = 16, 32 = 8 x 4, 64 = 8 x 8, 128 = 8 x 8 x 2 (columns
// and rows alike); 512 threads for 4096 points (64 x 64: 93 us per 2^24 elements against 105 at
// 256), 256 below
using R16 = RList<16>;
using R84 = RList<8, 4>;
using R88 = RList<8, 8>;
using R882 = RList<8, 8, 2>;
const Tile2D k2D[] = {
    MFFT_T2(64, 64, 2, RS_(8, 8), 2, RS_(8, 8), 512, R88, R88),
    MFFT_T2(32, 128, 2, RS_(8, 4), 3, RS_(8, 8, 2), 512, R84, R882),
    MFFT_T2(128, 32, 3, RS_(8, 8, 2), 2, RS_(8, 4), 512, R882, R84),
    MFFT_T2(32, 32, 2, RS_(8, 4), 2, RS_(8, 4), 256, R84, R84),
    MFFT_T2(32, 64, 2, RS_(8, 4), 2, RS_(8, 8), 256, R84, R88),
    MFFT_T2(64, 32, 2, RS_(8, 8), 2, RS_(8, 4), 256, R88, R84),
    MFFT_T2(16, 16, 1, RS_(16), 1, RS_(16), 256, R16, R16),
    MFFT_T2(16, 32, 1, RS_(16), 2, RS_(8, 4), 256, R16, R84),
    MFFT_T2(32, 16, 2, RS_(8, 4), 1, RS_(16), 256, R84, R16),
    MFFT_T2(16, 64, 1, RS_(16), 2, RS_(8, 8), 256, R16, R88),
    MFFT_T2(64, 16, 2, RS_(8, 8), 1, RS_(16), 256, R88, R16),
    MFFT_T2(16, 128, 1, RS_(16), 3, RS_(8, 8, 2), 256, R16, R882),
    MFFT_T2(128, 16, 3, RS_(8, 8, 2), 1, RS_(16), 256, R882, R16),
};
#undef RS_
#undef MFFT_T2
constexpr int kNum2D = (int)(sizeof(k2D) / sizeof(k2D[0]));

int tile2d_lookup(int64_t n0, int64_t n1, const StageArgs &col, const StageArgs &row) {
  static const int on = [] {
    const char *e = std::getenv("MFFT_TILE");
    return e ? std::atoi(e) : 1;
  }();
  if (!on) return 0;
  for (int i = 0; i < kNum2D; ++i) {
    const Tile2D &t = k2D[i];
    if (t.n0 != n0 || t.n1 != n1 || t.np0 != col.npass || t.np1 != row.npass) continue;
    bool same = true;
    for (int p = 0; p < t.np0; ++p) same = same && t.r0[p] == col.pass[p].R;
    for (int p = 0; p < t.np1; ++p) same = same && t.r1[p] == row.pass[p].R;
    if (same) return i + 1;
  }
  return 0;
}

hipError_t launch_tile2d(bool fwd, int idx, const StageArgs &col, const StageArgs &row,
                         const double2 *in, double2 *out, const double2 *tw0,
                         const double2 *tw1, int64_t howmany, hipStream_t stream) {
  if (howmany == 0) return hipSuccess;
  if (idx <= 0 || idx > kNum2D || howmany < 0) return hipErrorInvalidValue;
  const Tile2D &t = k2D[idx - 1];
  const void *k = t.k[fwd ? 1 : 0];
  {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.insert(k).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax);
  }
  if (howmany > INT_MAX) return hipErrorInvalidValue;
  const size_t lds = (size_t)t.n0 * lds_len(t.n1) * sizeof(double2);
  void *args[] = {(void *)&in, (void *)&out, (void *)&tw0, (void *)&tw1, (void *)&col,
                  (void *)&row};
  return hipLaunchKernel(k, dim3((unsigned)howmany), dim3(t.NT), args, lds, stream);
}

}  // namespace mfft
