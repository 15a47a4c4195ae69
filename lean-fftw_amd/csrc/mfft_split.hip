// Split register kernel for gfx950: one power-of-two transform of N = P * Q points per
// workgroup, held in registers (E per thread), computed as P-point column transforms and
// Q-point row transforms inside the workgroup.
//
// The passes are the plan's Stockham passes (LeanFFTW/Gen/Stockham.lean:88-98) split after
// the passes whose radices multiply to P, as a two-stage plan splits them: the first passes
// act on the columns x[q + Q p] (p < P) of the transform, and leave column q's result at
// q P + o; the rest act on the rows k + P c (c < Q) with twiddle base k and multiplier P, and
// write row k's output o to k + P o. Every butterfly takes the inputs and twiddles that
// refTransform gives it, so the result is bit-identical to refTransform.
//
// A column (P / E threads) and a row (Q / E threads) each lie inside one wave, so the
// exchanges between the passes of a column, or of a row, go through that wave's own part of
// LDS with no workgroup barrier; only the hand-over from columns to rows is a workgroup-wide
// exchange. The register kernel of mfft_reg.hip exchanges through the whole workgroup after
// every pass; with one workgroup per CU (16384 points) its waves then wait at every barrier
// for the slowest, and LDS traffic and arithmetic never overlap.
//
// Lanes: in the column phase lane l of wave w holds column q = w C + l mod C (C = 64 / (P/E)
// columns per wave) as thread t = l / C of that column, elements q + Q (t + (P/E) m); in the
// row phase it holds row k = w R + l mod R as thread l / R. Loads and stores therefore move
// runs of C (or R) consecutive elements, whole 128-byte lines.
#include <mutex>
#include <set>

#include "mfft_pass.h"

namespace mfft {

namespace {

template <int... Rs>
struct Radices {};

template <int... Rs>
constexpr int product(Radices<Rs...>) {
  return (1 * ... * Rs);
}


// LDS index of double e of a column or row region: one padding word after every 16 and one
// more after every 64. A 64-bit LDS access is served 16 lanes at a time over 16 bank pairs;
// with this layout (and a region length of n + n/16 + n/64) the exchanges of 64- and
// 128-point sub-problems with 4 or 8 threads each are free of bank conflicts (a bank model,
// confirmed by SQ_LDS_BANK_CONFLICT = 0; one padding word per 32 left 2-way conflicts in the
// 128-point exchanges).
__device__ __forceinline__ int spad(int e) { return e + (e >> 4) + (e >> 6); }
__host__ __device__ constexpr int spad_len(int n) { return n + (n >> 4) + (n >> 6); }

// Orders the wave's LDS accesses (a wave's LDS operations complete in order; this keeps the
// compiler from moving them across the exchange).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One pass of radix R on the E registers of thread t of a T-thread sub-problem: butterfly u
// (u < E / R) is j = t + T u, its inputs are registers u + r E/R. Twiddle of input r:
// tw[(r - 1) TWNS + kb + k KMUL], k = j mod NS (NS: the product of the sub-problem's
// earlier radices); no twiddles in the plan's first pass.
template <bool FWD, int R, int T, int E, int NS, bool TWD, int TWNS, int KMUL>
__device__ __forceinline__ void sbfly(double2 (&v)[E], int t, const double2 *__restrict__ tw,
                                      int kb) {
  constexpr int NB = E / R;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    double2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[u + r * NB];
    if constexpr (TWD) {
      const int k = (t + u * T) & (NS - 1);
      const double2 *tp = tw + (kb + k * KMUL);
#pragma unroll
      for (int r = 1; r < R; ++r) x[r] = twmul(tp[(r - 1) * TWNS], x[r]);
    }
    Bfly<FWD, R, false>::run(x);
#pragma unroll
    for (int r = 0; r < R; ++r) v[u + r * NB] = x[r];
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Exchange after a pass of radix R inside one sub-problem's LDS region `sub` (T threads of
// one wave, T dividing 16): output r of butterfly u goes to its Stockham position
// (j / NS) NS R + j mod NS + r NS, then thread t reads elements t + T m, at spad(t + T m) =
// t + T m + (T m) / 16 + (T m) / 64. Real parts, then imaginary parts.
template <int R, int T, int E, int NS>
__device__ __forceinline__ void wave_exchange(double2 (&v)[E], double *sub, int t) {
  static_assert(16 % T == 0, "fixed read offsets need T | 16");
  constexpr int NB = E / R;
  int ts = t;
  asm volatile("" : "+v"(ts));  // the write positions are computed here, not hoisted
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int j = ts + u * T;
    const int k = j & (NS - 1);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) sub[spad(base + r * NS)] = v[u + r * NB].x;
  }
  wave_sync();
  const double *rd = sub + t;
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].x = rd[T * m + ((T * m) >> 4) + ((T * m) >> 6)];
  wave_sync();
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int j = ts + u * T;
    const int k = j & (NS - 1);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) sub[spad(base + r * NS)] = v[u + r * NB].y;
  }
  wave_sync();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].y = rd[T * m + ((T * m) >> 4) + ((T * m) >> 6)];
  wave_sync();
}

// The passes of the column phase (plan passes 0 ... ): global ns = local ns.
template <bool FWD, int T, int E, int PI, int NS, int R, int... Rest>
__device__ __forceinline__ void col_passes(double2 (&v)[E], int t, double *sub, const StageArgs &a,
                                           const double2 *__restrict__ tw) {
  sbfly<FWD, R, T, E, NS, (PI > 0), NS, 1>(v, t, tw + (PI > 0 ? a.pass[PI].twoff : 0), 0);
  if constexpr (sizeof...(Rest) > 0) {
    wave_exchange<R, T, E, NS>(v, sub, t);
    col_passes<FWD, T, E, PI + 1, NS * R, Rest...>(v, t, sub, a, tw);
  }
}

// The passes of the row phase (plan passes PI ... ): row k, twiddle base k, multiplier P,
// table row length P NS.
template <bool FWD, int T, int E, int P, int PI, int NS, int R, int... Rest>
__device__ __forceinline__ void row_passes(double2 (&v)[E], int t, int k, double *sub,
                                           const StageArgs &a, const double2 *__restrict__ tw) {
  sbfly<FWD, R, T, E, NS, true, P * NS, P>(v, t, tw + a.pass[PI].twoff, k);
  if constexpr (sizeof...(Rest) > 0) {
    wave_exchange<R, T, E, NS>(v, sub, t);
    row_passes<FWD, T, E, P, PI + 1, NS * R, Rest...>(v, t, k, sub, a, tw);
  }
}

// Swizzle of the hand-over: row o's elements are permuted within 32-element blocks by
// xsw(o) = 9 o mod 32, conflict-free for both access patterns of the schedules below (without
// it the 16384-point kernel runs 20% slower, with xsw(o) = o mod 32 5%).
__device__ __forceinline__ int xsw(int o) { return (o * 9) & 31; }

template <int... As, int... Bs>
constexpr int split_lds_doubles(Radices<As...> ra, Radices<Bs...> rb) {
  constexpr int P = (1 * ... * As), Q = (1 * ... * Bs);
  // column regions (Q of them), row regions (P), and the hand-over (P rows of Q)
  const int c = Q * spad_len(P), r = P * spad_len(Q), x = P * Q;
  return c > r ? (c > x ? c : x) : (r > x ? r : x);
}

// Loads of the column phase: element q + Q (tc + TC m) into register m, times the input
// weight if there is one.
template <int Q, int TC, int E, bool PRE>
__device__ __forceinline__ void split_load(double2 (&v)[E], const double2 *src, const StageArgs &a,
                                           int q, int tc) {
  const double2 *s = src + (q + Q * tc);
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = ld_stream(s + Q * TC * m);
  if constexpr (PRE) {
    const double *w = a.pre + (int64_t)(q + Q * tc) * a.pre_e;
    const double wk = a.pre_s ? a.pre_k * a.pre_s[0] : 1.0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double c = wk * w[(int64_t)(Q * TC * m) * a.pre_e];
      v[m].x = v[m].x * c;
      v[m].y = v[m].y * c;
    }
  }
}

// Stores of the row phase: register m is row k's output o = tr + TR m, element k + P o;
// returns the thread's sum of |out|^2 (SQ).
template <int P, int TR, int E, bool POST, bool CACHED, bool SQ>
__device__ __forceinline__ double split_store(const double2 (&v)[E], double2 *dst,
                                              const StageArgs &a, int k, int tr) {
  int trs = tr;
  asm volatile("" : "+v"(trs));  // keep the store addresses out of the pass code
  const int e0 = k + P * trs;
  double2 *d = dst + e0;
  const double *w = POST ? a.post + (int64_t)e0 * a.post_e : nullptr;
  double sumsq = 0.0;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    double2 y = v[m];
    if constexpr (POST) {
      const double c = w[(int64_t)(P * TR * m) * a.post_e];
      y.x = y.x * c;
      y.y = y.y * c;
    }
    if constexpr (CACHED)
      d[P * TR * m] = y;
    else
      st_stream(d + P * TR * m, y);
    if constexpr (SQ) sumsq = sumsq + (y.x * y.x + y.y * y.y);
  }
  return sumsq;
}

template <bool FWD, bool SQ, bool PRE, bool POST, bool CACHED, int E, int... As, int... Bs>
__device__ __forceinline__ void split_body(Radices<As...>, Radices<Bs...>, const double2 *in,
                                           double2 *out, const double2 *__restrict__ tw,
                                           const StageArgs &a, double *lds) {
  constexpr int P = (1 * ... * As), Q = (1 * ... * Bs), N = P * Q;
  constexpr int TC = P / E, TR = Q / E;     // threads per column, per row
  constexpr int CW = 64 / TC, RW = 64 / TR;  // columns, rows per wave
  constexpr int SA = sizeof...(As);
  static_assert(Q >= 32, "the transpose swizzle permutes 32-element blocks of a row");
  static_assert(TC <= 32 && TR <= 32 && 32 % TC == 0 && 32 % TR == 0, "columns and rows fit a wave");
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int64_t h = blockIdx.x;  // one transform per workgroup

  // ---- column phase
  const int cw = lane % CW, tc = lane / CW, q = w * CW + cw;
  double2 v[E];
  split_load<Q, TC, E, PRE>(v, in + h * a.in_g, a, q, tc);
  col_passes<FWD, TC, E, 0, 1, As...>(v, tc, lds + q * spad_len(P), a, tw);
  // ---- hand-over: column q's output o (register m, o = tc + TC m) to row o, element q, at
  // o Q + (q ^ xsw(o)): the swizzle spreads both the writes (runs of consecutive q for a few
  // o) and the reads (runs of consecutive elements of several rows) over the LDS banks
  __syncthreads();  // every wave is done with its column regions
  const int rw = lane % RW, tr = lane / RW, k = w * RW + rw;
  const int sk = xsw(k);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int o = tc + TC * m;
    lds[o * Q + (q ^ xsw(o))] = v[m].x;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].x = lds[k * Q + ((tr + TR * m) ^ sk)];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int o = tc + TC * m;
    lds[o * Q + (q ^ xsw(o))] = v[m].y;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].y = lds[k * Q + ((tr + TR * m) ^ sk)];
  __syncthreads();  // the row regions overlap the transpose

  // ---- row phase
  row_passes<FWD, TR, E, P, SA, 1, Bs...>(v, tr, k, lds + k * spad_len(Q), a, tw);
  const double sumsq = split_store<P, TR, E, POST, CACHED, SQ>(v, out + h * a.out_g, a, k, tr);
  if constexpr (SQ) {
    __shared__ double wsum[N / E / 64];
    double s = sumsq;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = s + __shfl_xor(s, o, 64);
    if (lane == 0) wsum[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < N / E / 64; ++i) t = t + wsum[i];
      a.sumsq[blockIdx.x] = t;
    }
  }
}

template <bool FWD, bool SQ, int V, int E, class RA, class RB>
__global__ __launch_bounds__(product(RA{}) * product(RB{}) / E, 4) void split_kernel(
    const double2 *in, double2 *out, const double2 *__restrict__ tw, const StageArgs a) {
  extern __shared__ double lds_split[];
  // V: bit 0 input weights, bit 1 output weights, bit 2 ordinary (cached) stores
  split_body<FWD, SQ, (V & 1) != 0, (V & 2) != 0, (V & 4) != 0, E>(RA{}, RB{}, in, out, tw, a,
                                                                  lds_split);
}

// ---------------------------------------------------------------- 2-D arrays
// A row-major P x Q array (element (i0, i1) at i0 Q + i1) in one workgroup: the column phase
// above is the axis-0 transform (column q = elements q + Q i0, the column plan's passes and
// table), the hand-over leaves row k0 = column outputs k0, and the row phase runs the row
// plan's passes and table on each row (no factor between the axes) and writes X[k0, k1] at
// k0 Q + k1. The two-launch 2-D plan applies the same line transforms in the same order (axis
// 0, then axis 1), so the bits are its bits. Row-phase lanes walk along the row
// (lane l: thread l mod TR of row w RW + l / TR), so the stores move whole 128-byte lines.
template <bool FWD, int E, int... As, int... Bs>
__device__ __forceinline__ void split2d_body(Radices<As...>, Radices<Bs...>, const double2 *in,
                                             double2 *out, const double2 *__restrict__ tw0,
                                             const double2 *__restrict__ tw1,
                                             const StageArgs &col, const StageArgs &row,
                                             double *lds) {
  constexpr int P = (1 * ... * As), Q = (1 * ... * Bs), N = P * Q;
  constexpr int TC = P / E, TR = Q / E;
  constexpr int CW = 64 / TC, RW = 64 / TR;
  static_assert(Q >= 32, "the transpose swizzle permutes 32-element blocks of a row");
  static_assert(TC <= 32 && TR <= 32 && 32 % TC == 0 && 32 % TR == 0, "columns and rows fit a wave");
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int64_t h = blockIdx.x;  // one array per workgroup
  const int cw = lane % CW, tc = lane / CW, q = w * CW + cw;
  double2 v[E];
  split_load<Q, TC, E, false>(v, in + h * N, col, q, tc);
  col_passes<FWD, TC, E, 0, 1, As...>(v, tc, lds + q * spad_len(P), col, tw0);
  __syncthreads();
  const int tr = lane % TR, k = w * RW + lane / TR;
  const int sk = xsw(k);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int o = tc + TC * m;
    lds[o * Q + (q ^ xsw(o))] = v[m].x;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].x = lds[k * Q + ((tr + TR * m) ^ sk)];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int o = tc + TC * m;
    lds[o * Q + (q ^ xsw(o))] = v[m].y;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].y = lds[k * Q + ((tr + TR * m) ^ sk)];
  __syncthreads();
  col_passes<FWD, TR, E, 0, 1, Bs...>(v, tr, lds + k * spad_len(Q), row, tw1);
  int trs = tr;
  asm volatile("" : "+v"(trs));  // keep the store addresses out of the pass code
  double2 *d = out + h * N + (int64_t)k * Q + trs;
#pragma unroll
  for (int m = 0; m < E; ++m) st_stream(d + TR * m, v[m]);
}

template <bool FWD, int E, class RA, class RB>
__global__ __launch_bounds__(product(RA{}) * product(RB{}) / E, 4) void split2d_kernel(
    const double2 *in, double2 *out, const double2 *__restrict__ tw0,
    const double2 *__restrict__ tw1, const StageArgs col, const StageArgs row) {
  extern __shared__ double lds_split[];
  split2d_body<FWD, E>(RA{}, RB{}, in, out, tw0, tw1, col, row, lds_split);
}

struct Split2dSchedule {
  int n0, n1, e, na, nb;
  int ra[4], rb[4];
  size_t lds;  // bytes
  const void *k[2];
};

using R84 = Radices<8, 4>;
using R88 = Radices<8, 8>;
using R882 = Radices<8, 8, 2>;
using R884 = Radices<8, 8, 4>;
// arrays of 8192 and 16384 points whose columns and rows each fit one wave (16 elements per
// thread), with the default radices of their parts
template <class RA, class RB>
constexpr Split2dSchedule s2d(int n0, int n1, int na, const int (&ra)[3], int nb, const int (&rb)[3]) {
  return {n0, n1, 16, na, nb, {ra[0], ra[1], ra[2], 0}, {rb[0], rb[1], rb[2], 0},
          (size_t)split_lds_doubles(RA{}, RB{}) * sizeof(double),
          {(const void *)split2d_kernel<false, 16, RA, RB>, (const void *)split2d_kernel<true, 16, RA, RB>}};
}
const Split2dSchedule kSplit2d[] = {
    s2d<R882, R882>(128, 128, 3, {8, 8, 2}, 3, {8, 8, 2}),
    s2d<R88, R884>(64, 256, 2, {8, 8, 0}, 3, {8, 8, 4}),
    s2d<R884, R88>(256, 64, 3, {8, 8, 4}, 2, {8, 8, 0}),
    s2d<R88, R882>(64, 128, 2, {8, 8, 0}, 3, {8, 8, 2}),
    s2d<R882, R88>(128, 64, 3, {8, 8, 2}, 2, {8, 8, 0}),
    s2d<R84, R884>(32, 256, 2, {8, 4, 0}, 3, {8, 8, 4}),
    s2d<R884, R84>(256, 32, 3, {8, 8, 4}, 2, {8, 4, 0}),
};

struct SplitSchedule {
  int n, e, na, nb;
  int rs[8];
  size_t lds;          // bytes
  int cached;          // 4: ordinary stores of the output, 0: non-temporal
  const void *k[2][2][4];  // [direction][sum of squares][input weights + 2 output weights]
};

#define MFFT_SPW(F, S, C, EE, RA, RB)                                                         \
  {(const void *)split_kernel<F, S, C | 0, EE, RA, RB>,                                      \
   (const void *)split_kernel<F, S, C | 1, EE, RA, RB>,                                      \
   (const void *)split_kernel<F, S, C | 2, EE, RA, RB>,                                      \
   (const void *)split_kernel<F, S, C | 3, EE, RA, RB>}
#define MFFT_SP(C, EE, RA, RB)                                                                \
  C, {{MFFT_SPW(false, false, C, EE, RA, RB), MFFT_SPW(false, true, C, EE, RA, RB)},          \
      {MFFT_SPW(true, false, C, EE, RA, RB), MFFT_SPW(true, true, C, EE, RA, RB)}}
#define MFFT_LDS(RA, RB) (size_t)split_lds_doubles(RA{}, RB{}) * sizeof(double)
using R168 = Radices<16, 8>;
using R164 = Radices<16, 4>;
const SplitSchedule kSplit[] = {
    {8192, 16, 2, 2, {16, 4, 16, 8}, MFFT_LDS(R164, R168), MFFT_SP(0, 16, R164, R168)},
    {16384, 16, 2, 2, {16, 8, 16, 8}, MFFT_LDS(R168, R168), MFFT_SP(0, 16, R168, R168)},
};
#undef MFFT_SP
#undef MFFT_SPW
#undef MFFT_LDS

}  // namespace

const void *split_schedule(int64_t n, const int *radices, int nradix) {
  for (const auto &s : kSplit) {
    if (s.n != n || s.na + s.nb != nradix) continue;
    bool same = true;
    for (int i = 0; i < nradix; ++i) same = same && s.rs[i] == radices[i];
    if (same) return &s;
  }
  return nullptr;
}

// The single-pass 2-D kernel for an n0 x n1 plan whose axis parts are single stages of these
// radices (nullptr if there is none).
const void *split2d_lookup(int64_t n0, int64_t n1, const StageArgs &col, const StageArgs &row) {
  for (const auto &s : kSplit2d) {
    if (s.n0 != n0 || s.n1 != n1 || s.na != col.npass || s.nb != row.npass) continue;
    bool same = true;
    for (int i = 0; i < s.na; ++i) same = same && s.ra[i] == col.pass[i].R;
    for (int i = 0; i < s.nb; ++i) same = same && s.rb[i] == row.pass[i].R;
    if (same) return &s;
  }
  return nullptr;
}

hipError_t launch_split2d(bool fwd, const void *sched, const StageArgs &col, const StageArgs &row,
                          const double2 *in, double2 *out, const double2 *tw0,
                          const double2 *tw1, int64_t arrays, hipStream_t stream) {
  const auto *s = static_cast<const Split2dSchedule *>(sched);
  if (!s || arrays < 1) return arrays < 1 ? hipSuccess : hipErrorInvalidValue;
  const void *k = s->k[fwd ? 1 : 0];
  {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.insert(k).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->lds);
  }
  void *args[] = {(void *)&in, (void *)&out, (void *)&tw0, (void *)&tw1, (void *)&col,
                  (void *)&row};
  return hipLaunchKernel(k, dim3((unsigned)arrays), dim3(s->n0 * s->n1 / s->e), args, s->lds,
                         stream);
}

// The split kernel for a single-stage register plan of these radices (complex data, one
// transform per workgroup); returns false when there is none.
bool launch_split(bool fwd, const StageArgs &a, const double2 *in, double2 *out,
                  const double2 *tw, hipStream_t stream, hipError_t *err) {
  if (a.io_in || a.io_out || a.in_e != 1 || a.out_e != 1 || a.G != 1) return false;
  int rs[MAX_PASSES];
  for (int p = 0; p < a.npass; ++p) rs[p] = a.pass[p].R;
  const auto *s = static_cast<const SplitSchedule *>(split_schedule(a.L, rs, a.npass));
  if (!s) return false;
  const void *k = s->k[fwd ? 1 : 0][a.sumsq ? 1 : 0][(a.pre ? 1 : 0) | (a.post ? 2 : 0)];
  {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.insert(k).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->lds);
  }
  void *args[] = {(void *)&in, (void *)&out, (void *)&tw, (void *)&a};
  *err = a.nsub <= 0 ? hipSuccess
                     : hipLaunchKernel(k, dim3((unsigned)a.nsub), dim3(s->n / s->e), args, s->lds,
                                       stream);
  return true;
}

}  // namespace mfft
