// Register-resident Stockham kernel for gfx950: power-of-two transforms of N points with
// T = N / E threads per transform, each holding E elements (8, or 16 at N = 16384) in
// registers for the whole transform; a workgroup runs 256 / T transforms when T < 256, else
// one. Used for N = 64 ... 16384: up to 8192 points it beats the LDS-tile kernel, and at
// 16384 a transform (256 KiB, more than the 160 KiB of LDS) still fits the registers of one
// CU, so it runs in one HBM round trip.
//
// Every pass is the reference's passWith (LeanFFTW/Gen/Stockham.lean:88-98) with the
// butterflies of csrc/bfly_gen.h and the plan's pass tables, so the result is bit-identical
// to refTransform for the same radix sequence. Thread t always holds the elements
// t + T m (m < E): for radix R its butterflies are j = t + T u (u < E / R), whose inputs
// x[j + r N/R] are exactly those elements (m = u + r E/R). So the first pass reads HBM and
// the last pass writes HBM with consecutive lanes on consecutive addresses, and between
// passes the outputs go to LDS at their Stockham positions and come back as the next
// pass's columns. Up to 8192 points the exchange moves whole complex values (N padded
// double2, 136 KiB at 8192); at 16384 it moves the real parts, then the imaginary parts,
// through one buffer of N doubles (132 KiB), half of what the transform occupies.
#include <cstdlib>
#include <mutex>
#include <set>

#include "mfft_pass.h"

namespace mfft {

// LDS index of double e: one padding word after every 32, so that outputs written at a
// stride of R <= 16 doubles by consecutive lanes spread over the banks.
__device__ __forceinline__ int rpad(int e) { return e + (e >> 5); }
__host__ __device__ constexpr int reg_lds_doubles(int n) { return n + (n >> 5) + 1; }

// One pass of radix R on the E registers v, in place: butterfly u (u < E / R) takes its
// inputs from registers u + r E/R (columns t + T (u + r E/R)) and leaves output r there.
template <bool FWD, int R, int T, int E>
__device__ __forceinline__ void reg_bfly(double2 (&v)[E], int t, int ns,
                                         const double2 *__restrict__ twp, bool twd) {
  constexpr int NB = E / R;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    double2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[u + r * NB];
    if (twd) {
      const int k = (t + u * T) & (ns - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) x[r] = twmul(twp[(r - 1) * ns + k], x[r]);
    }
    Bfly<FWD, R, false>::run(x);
#pragma unroll
    for (int r = 0; r < R; ++r) v[u + r * NB] = x[r];
    // one butterfly's twiddles in flight at a time: hoisting the next butterfly's loads
    // above this one costs registers that 16 waves per CU do not have
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Exchange through LDS after a pass of radix R: output r of butterfly u (register
// u + r E/R) goes to its Stockham position (j / ns) ns R + j mod ns + r ns, j = t + T u;
// then every thread reads its columns t + T m.
template <int R, int T, int E>
__device__ __forceinline__ void reg_exchange(double2 (&v)[E], double *lds, int t, int ns) {
  constexpr int NB = E / R;
  int pos[E];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int j = t + u * T;
    const int k = j & (ns - 1);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) pos[u + r * NB] = rpad(base + r * ns);
  }
#pragma unroll
  for (int i = 0; i < E; ++i) lds[pos[i]] = v[i].x;
  __syncthreads();
  // T is a multiple of 32, so rpad(t + T m) = rpad(t) + m (T + T / 32): fixed offsets
  const double *col = lds + rpad(t);
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].x = col[m * (T + T / 32)];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) lds[pos[i]] = v[i].y;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].y = col[m * (T + T / 32)];
  __syncthreads();
}

// The same exchange with whole complex values through N padded double2 (two barriers).
template <int R, int T, int E>
__device__ __forceinline__ void reg_exchange_full(double2 (&v)[E], double2 *lds, int t, int ns) {
  constexpr int NB = E / R;
  int pos[E];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int j = t + u * T;
    const int k = j & (ns - 1);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) pos[u + r * NB] = lpad(base + r * ns);
  }
#pragma unroll
  for (int i = 0; i < E; ++i) lds[pos[i]] = v[i];
  __syncthreads();
  if constexpr (T % 16 == 0) {  // lpad(t + T m) = lpad(t) + m (T + T / 16)
    const double2 *col = lds + lpad(t);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = col[m * (T + T / 16)];
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = lds[lpad(t + T * m)];
  }
  __syncthreads();
}

template <int R, int T, int E>
__device__ __forceinline__ void reg_exchange_any(double2 (&v)[E], void *lds, int t, int ns) {
  if constexpr (E == 8)
    reg_exchange_full<R, T, E>(v, static_cast<double2 *>(lds), t, ns);
  else
    reg_exchange<R, T, E>(v, static_cast<double *>(lds), t, ns);
}

// Workgroup size: T threads per transform, and G = 256 / T transforms per workgroup when a
// transform has fewer than 256 threads (N < 2048).
template <int T>
constexpr int reg_block() {
  return T < 256 ? 256 : T;
}

// Load a transform's columns t + T m (non-temporal), times the input weights (scaled by the
// device scalar pre_s only when PS: the fixed kernels without it keep 60-64 VGPRs at 8
// elements per thread, which a scaled weight took to 68-72, and no spills at 16).
template <int T, int E, bool REAL16 = false, bool PS = true>
__device__ __forceinline__ void reg_load(double2 (&v)[E], const StageArgs &a, const double2 *in,
                                         int64_t base, int t, bool valid) {
  // 16 elements per thread: real input only, and only in the kernels that take it without
  // spilling (REAL16); the others run no map (the host does not fuse for them)
  if ((E == 16 && !REAL16) || a.io_in == 0) {
    const double2 *src = in + base;
    if (a.cached_in) {
#pragma unroll
      for (int m = 0; m < E; ++m)
        v[m] = valid ? src[(int64_t)(t + T * m) * a.in_e] : make_double2(0.0, 0.0);
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m)
        v[m] = valid ? ld_stream(src + (int64_t)(t + T * m) * a.in_e) : make_double2(0.0, 0.0);
    }
  } else if constexpr (E == 16) {
    const double *src = reinterpret_cast<const double *>(in) + base;
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = make_double2(
          valid ? __builtin_nontemporal_load(src + (int64_t)(t + T * m) * a.in_e) : 0.0, 0.0);
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = valid ? load_io(a, in, base, (int64_t)(t + T * m) * a.in_e) : make_double2(0.0, 0.0);
  }
  if (a.pre) {
    if constexpr (PS) {
      const double wk = a.pre_s ? a.pre_k * a.pre_s[0] : 1.0;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double w = wk * a.pre[(int64_t)(t + T * m) * a.pre_e];
        v[m].x = v[m].x * w;
        v[m].y = v[m].y * w;
      }
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double w = a.pre[(int64_t)(t + T * m) * a.pre_e];
        v[m].x = v[m].x * w;
        v[m].y = v[m].y * w;
      }
    }
  }
}

// Store after the last pass (ns = N / R): output r of butterfly j = t + T u goes to
// j + r N/R, the column of the register that holds it, so register m goes to element t + T m.
// Returns this thread's sum of |out|^2. (The opaque copy of t keeps the compiler from
// computing these addresses before the passes and holding them in registers through them.)
template <int T, int E>
__device__ __forceinline__ double reg_store(const double2 (&v)[E], const StageArgs &a,
                                            double2 *out, int64_t base, int t, bool valid) {
  int ts = t;
  asm volatile("" : "+v"(ts));
  if (a.io_out != 0) {  // real-data maps: composite plans take no weights or partials
    if (valid) {
#pragma unroll
      for (int m = 0; m < E; ++m) store_io(a, out, base, (int64_t)(ts + T * m) * a.out_e, v[m]);
    }
    return 0.0;
  }
  double sumsq = 0.0;
  double2 *dst = out + base;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int e = ts + T * m;
    double2 y = v[m];
    if (a.post) {
      const double w = a.post[(int64_t)e * a.post_e];
      y.x = y.x * w;
      y.y = y.y * w;
    }
    if (valid) {
      if (a.cached_out)
        dst[(int64_t)e * a.out_e] = y;
      else
        st_stream(dst + (int64_t)e * a.out_e, y);
    }
    sumsq = sumsq + (y.x * y.x + y.y * y.y);
  }
  return sumsq;
}

// One sum-of-squares partial per workgroup, a fixed reduction tree.
template <int BT>
__device__ __forceinline__ void block_partial(double sumsq, double *dst) {
  __shared__ double wsum[BT / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sumsq = sumsq + __shfl_xor(sumsq, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sumsq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < BT / 64; ++w) s = s + wsum[w];
    *dst = s;
  }
}

// Any radix sequence in {2, 4, 8} (16 elements per thread: radix 4 only), dispatched pass by
// pass at run time. Registers are held to 128 per thread so that 16 waves fit on a CU.
template <bool FWD, int T, int E, bool SQ>
__global__ __launch_bounds__(reg_block<T>(), 4) void reg_kernel(const double2 *in, double2 *out,
                                                               const double2 *__restrict__ tw,
                                                               const StageArgs a) {
  constexpr int G = reg_block<T>() / T;
  static_assert(G == 1 || E == 8, "several transforms per workgroup use the whole-complex exchange");
  extern __shared__ double2 lds_c[];
  const int t = G == 1 ? (int)threadIdx.x : (int)threadIdx.x % T;
  const int g = G == 1 ? 0 : (int)threadIdx.x / T;
  // transform g of the workgroup: its own LDS region; a transform past the batch runs the
  // passes on zeros and stores nothing
  void *lds_d = lds_c + g * lds_len(T * E);
  const int64_t h = (int64_t)blockIdx.x * G + g;
  const bool valid = h < a.nsub;
  double2 v[E];
  reg_load<T, E>(v, a, in, (valid ? h : 0) * a.in_g, t, valid);
  for (int p = 0; p < a.npass; ++p) {
    const PassArgs pa = a.pass[p];
    const bool twd = pa.twoff >= 0;
    const double2 *twp = tw + (twd ? pa.twoff : 0);
    const bool last = p == a.npass - 1;
#define MFFT_REG(RR)                                              \
  case RR:                                                        \
    reg_bfly<FWD, RR, T, E>(v, t, pa.nsl, twp, twd);              \
    if (!last) reg_exchange_any<RR, T, E>(v, lds_d, t, pa.nsl);   \
    break;
    // 16 elements per thread (N = 16384) run radix 4 only: radix 2's eight butterflies per
    // thread, or radix 8's larger ones, exceed the 128 registers of 16 waves per CU
    switch (pa.R) {
      case 2:
        if constexpr (E == 8) {
          reg_bfly<FWD, 2, T, E>(v, t, pa.nsl, twp, twd);
          if (!last) reg_exchange_any<2, T, E>(v, lds_d, t, pa.nsl);
        }
        break;
      MFFT_REG(4)
      case 8:
        if constexpr (E == 8) {
          reg_bfly<FWD, 8, T, E>(v, t, pa.nsl, twp, twd);
          if (!last) reg_exchange_any<8, T, E>(v, lds_d, t, pa.nsl);
        }
        break;
      default:
        break;
    }
#undef MFFT_REG
  }
  const double sumsq = reg_store<T, E>(v, a, out, (valid ? h : 0) * a.out_g, t, valid);
  if constexpr (SQ) block_partial<reg_block<T>()>(sumsq, a.sumsq + blockIdx.x);
}

// ---------------------------------------------------------------- fixed schedules
// The default schedules of 64 ... 16384 points as compile-time radix sequences: ns, the
// twiddle flag and the exchange offsets become constants and the passes straight-line code,
// which takes the kernel from about 104 to about 61 registers at 8 elements per thread (and
// lets 16 elements per thread run radix 8 and 16 without spilling). With the exchange of
// real then imaginary parts (N doubles of LDS) two workgroups of 1024 threads share a CU at
// 8192 points, four of 512 at 4096.
template <bool FWD, int T, int E, bool HALF, int P, int NS, int R, int... Rest>
__device__ __forceinline__ void fixed_passes(double2 (&v)[E], int t, void *lds,
                                             const StageArgs &a, const double2 *__restrict__ tw) {
  reg_bfly<FWD, R, T, E>(v, t, NS, tw + (P > 0 ? a.pass[P].twoff : 0), P > 0);
  if constexpr (sizeof...(Rest) > 0) {
    if constexpr (HALF)
      reg_exchange<R, T, E>(v, static_cast<double *>(lds), t, NS);
    else
      reg_exchange_full<R, T, E>(v, static_cast<double2 *>(lds), t, NS);
    fixed_passes<FWD, T, E, HALF, P + 1, NS * R, Rest...>(v, t, lds, a, tw);
  }
}

template <int R, int... Rest>
constexpr int first_radix() {
  return R;
}

// Below 2048 points a workgroup runs G = 256 / T transforms, each with its own LDS region,
// and exchanges whole complex values (LDS is not the limit there); from 2048 on one transform
// and the real/imaginary exchange.
template <bool FWD, int T, int E, bool SQ, bool PS, int... Rs>
__global__ __launch_bounds__(reg_block<T>(), 4) void fixed_kernel(const double2 *in, double2 *out,
                                                                 const double2 *__restrict__ tw,
                                                                 const StageArgs a) {
  constexpr int G = reg_block<T>() / T;
  constexpr bool HALF = G == 1;
  extern __shared__ double2 lds_c[];
  const int t = G == 1 ? (int)threadIdx.x : (int)threadIdx.x % T;
  const int g = G == 1 ? 0 : (int)threadIdx.x / T;
  void *lds_d = lds_c + g * lds_len(T * E);
  const int64_t h = (int64_t)blockIdx.x * G + g;
  const bool valid = G == 1 || h < a.nsub;  // one transform per workgroup: the grid is the batch
  double2 v[E];
  reg_load<T, E, first_radix<Rs...>() == 16, PS>(v, a, in, (valid ? h : 0) * a.in_g, t, valid);
  fixed_passes<FWD, T, E, HALF, 0, 1, Rs...>(v, t, lds_d, a, tw);
  const double sumsq = reg_store<T, E>(v, a, out, (valid ? h : 0) * a.out_g, t, valid);
  if constexpr (SQ) block_partial<reg_block<T>()>(sumsq, a.sumsq + blockIdx.x);
}

// The fixed schedules: N, elements per thread, radix sequence, and the kernels by
// (direction, sum of squares, input weights scaled by a device scalar; sum-of-squares calls
// take no scaled weights).
struct FixedSchedule {
  int n, e, npass;
  int rs[8];
  const void *k[2][2][2];
};

#define MFFT_FX(TT, EE, ...)                                                                 \
  {{{(const void *)fixed_kernel<false, TT, EE, false, false, __VA_ARGS__>,                    \
     (const void *)fixed_kernel<false, TT, EE, false, true, __VA_ARGS__>},                    \
    {(const void *)fixed_kernel<false, TT, EE, true, false, __VA_ARGS__>, nullptr}},          \
   {{(const void *)fixed_kernel<true, TT, EE, false, false, __VA_ARGS__>,                     \
     (const void *)fixed_kernel<true, TT, EE, false, true, __VA_ARGS__>},                     \
    {(const void *)fixed_kernel<true, TT, EE, true, false, __VA_ARGS__>, nullptr}}}
const FixedSchedule kFixed[] = {
    {64, 8, 2, {8, 8}, MFFT_FX(8, 8, 8, 8)},
    {128, 8, 3, {8, 8, 2}, MFFT_FX(16, 8, 8, 8, 2)},
    {256, 8, 3, {8, 8, 4}, MFFT_FX(32, 8, 8, 8, 4)},
    {512, 8, 3, {8, 8, 8}, MFFT_FX(64, 8, 8, 8, 8)},
    {1024, 8, 4, {8, 8, 8, 2}, MFFT_FX(128, 8, 8, 8, 8, 2)},
    {2048, 8, 4, {8, 8, 8, 4}, MFFT_FX(256, 8, 8, 8, 8, 4)},
    {4096, 8, 4, {8, 8, 8, 8}, MFFT_FX(512, 8, 8, 8, 8, 8)},
    {8192, 8, 5, {8, 8, 8, 8, 2}, MFFT_FX(1024, 8, 8, 8, 8, 8, 2)},
    {8192, 16, 6, {8, 4, 4, 4, 4, 4}, MFFT_FX(512, 16, 8, 4, 4, 4, 4, 4)},
    {16384, 16, 4, {16, 16, 16, 4}, MFFT_FX(1024, 16, 16, 16, 16, 4)},
    {16384, 16, 5, {8, 8, 8, 8, 4}, MFFT_FX(1024, 16, 8, 8, 8, 8, 4)},
    // radix 16 at 16 elements per thread: 8192 = 16 x 16 x 8 x 4 is the plain default
    // (98-101 us per 256 MB batch against 106-110 for 8^4 x 2; 16^3 x 2 115, 16 x 16 x 4 x 8 108)
    {8192, 16, 4, {16, 16, 8, 4}, MFFT_FX(512, 16, 16, 16, 8, 4)},
};
#undef MFFT_FX

const FixedSchedule *fixed_schedule(int64_t n, const int *radices, int nradix) {
  static const int64_t min_n = [] {  // MFFT_FIXED_MIN: smallest size with a fixed schedule
    const char *e = std::getenv("MFFT_FIXED_MIN");
    return e ? (int64_t)std::atoll(e) : (int64_t)0;
  }();
  if (n < min_n) return nullptr;
  for (const auto &f : kFixed) {
    if (f.n != n || f.npass != nradix) continue;
    bool same = true;
    for (int i = 0; i < nradix; ++i) same = same && f.rs[i] == radices[i];
    if (same) return &f;
  }
  return nullptr;
}

// ---------------------------------------------------------------- register-resident stages
// A kernel stage of a two-stage plan, N = P * Q, with the same scheme: the workgroup takes G
// sub-problems of length L (Q decimated columns of length P, or P residue classes of length
// Q), T = L / E threads each, and holds them in registers through the stage's passes. Loads
// walk the sub-problem index with consecutive lanes (rows of G elements: 512 bytes at
// G = 32), and so do the stores of stage B; stage A's stores walk the element index instead
// (its intermediate is contiguous per column), so the exchange before the last pass
// redistributes the columns to that order. Each sub-problem has its own LDS region of
// reg_lds_doubles(L) doubles (real then imaginary parts), so G = 32 sub-problems of 256
// points take 68 KiB and two workgroups share a CU. Twiddles as in the stage kernel: pass p
// of sub-problem g0 + g reads tw[twoff + (r - 1) twns + (g0 + g) kb_u + k kmul].

// Twiddled butterflies of a stage pass (k = j mod NS, NS the product of the stage's earlier
// radices).
template <bool FWD, int R, int T, int E, int NS>
__device__ __forceinline__ void stage_bfly(double2 (&v)[E], int t, const double2 *__restrict__ twp,
                                           bool twd, int64_t twns, int64_t kb, int64_t kmul) {
  constexpr int NB = E / R;
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    double2 x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[u + r * NB];
    if (twd) {
      const int k = (t + u * T) & (NS - 1);
      const double2 *tp = twp + kb + (int64_t)k * kmul;
#pragma unroll
      for (int r = 1; r < R; ++r) x[r] = twmul(tp[(r - 1) * twns], x[r]);
    }
    Bfly<FWD, R, false>::run(x);
#pragma unroll
    for (int r = 0; r < R; ++r) v[u + r * NB] = x[r];
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Exchange of a stage pass of radix R: the writer (region gw, thread tw) puts output r of
// butterfly u at its Stockham position; the reader (region gr, thread tr) takes columns
// tr + T m. Real parts, then imaginary parts.
template <int R, int T, int E, int NS, int SUB>
__device__ __forceinline__ void stage_exchange(double2 (&v)[E], double *lds, int gw, int tw,
                                               int gr, int tr) {
  constexpr int NB = E / R;
  double *wr = lds + gw * SUB;
  int pos[E];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int j = tw + u * T;
    const int k = j & (NS - 1);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) pos[u + r * NB] = rpad(base + r * NS);
  }
  const double *rd = lds + gr * SUB;
#pragma unroll
  for (int i = 0; i < E; ++i) wr[pos[i]] = v[i].x;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].x = rd[rpad(tr + T * m)];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) wr[pos[i]] = v[i].y;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m].y = rd[rpad(tr + T * m)];
  __syncthreads();
}

// The passes of a stage from pass P on: the threads keep the input order (gi, ti) until the
// exchange before the last pass, which hands them the output order (go, to). kbi / kbo are
// the twiddle bases of the sub-problems in the two orders (those of a valid sub-problem for
// the lanes past a partial tile, whose results are not stored, so that every twiddle read
// stays inside the table).
template <bool FWD, int T, int E, int SUB, int P, int NS, int R, int... Rest>
__device__ __forceinline__ void stage_passes(double2 (&v)[E], double *lds, int gi, int ti, int go,
                                             int to, const StageArgs &a,
                                             const double2 *__restrict__ tw, int64_t kbi,
                                             int64_t kbo) {
  const PassArgs &pa = a.pass[P];
  const bool twd = P > 0 || pa.twoff >= 0;
  const int t = sizeof...(Rest) > 0 ? ti : to;
  const int64_t kb = sizeof...(Rest) > 0 ? kbi : kbo;
  stage_bfly<FWD, R, T, E, NS>(v, t, tw + (twd ? pa.twoff : 0), twd, pa.twns, kb, pa.kmul);
  if constexpr (sizeof...(Rest) > 0) {
    constexpr bool before_last = sizeof...(Rest) == 1;
    stage_exchange<R, T, E, NS, SUB>(v, lds, gi, ti, before_last ? go : gi, before_last ? to : ti);
    stage_passes<FWD, T, E, SUB, P + 1, NS * R, Rest...>(v, lds, gi, ti, go, to, a, tw, kbi, kbo);
  }
}

template <bool FWD, int L, int E, int G, int... Rs>
__global__ __launch_bounds__(L / E * G, 4) void stage_reg_kernel(const double2 *in, double2 *out,
                                                                const double2 *__restrict__ tw,
                                                                const StageArgs a) {
  constexpr int T = L / E, SUB = reg_lds_doubles(L);
  static_assert(sizeof...(Rs) >= 2, "a stage of one pass has no exchange to reorder its threads");
  extern __shared__ double lds_s[];
  int bx, by;
  if (!stage_block(a.xcd, a.nsub, G, bx, by)) return;
  const int i = threadIdx.x;
  const int64_t g0 = (int64_t)bx * G;
  const int64_t h = by;
  const int gcount = (int)(a.nsub - g0 < G ? a.nsub - g0 : G);
  const int gi = a.gfast_in ? i % G : i / T, ti = a.gfast_in ? i / G : i % T;
  const int go = a.gfast_out ? i % G : i / T, to = a.gfast_out ? i / G : i % T;
  double2 v[E];
  {
    const bool ok = gi < gcount;
    if (a.io_in == 0) {
      const double2 *src = in + h * a.in_h + (g0 + (ok ? gi : 0)) * a.in_g;
#pragma unroll
      for (int m = 0; m < E; ++m)
        v[m] = ok ? (a.cached_in ? src[(int64_t)(ti + T * m) * a.in_e]
                                 : ld_stream(src + (int64_t)(ti + T * m) * a.in_e))
                  : make_double2(0.0, 0.0);
    } else {
      const int64_t o = (g0 + (ok ? gi : 0)) * a.in_g;
#pragma unroll
      for (int m = 0; m < E; ++m)
        v[m] = ok ? load_io(a, in, h * a.in_h, o + (int64_t)(ti + T * m) * a.in_e)
                  : make_double2(0.0, 0.0);
    }
    if (a.pre) {
      const double *w = a.pre + (g0 + (ok ? gi : 0)) * a.pre_g;
      const double wk = a.pre_s ? a.pre_k * a.pre_s[0] : 1.0;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double c = wk * w[(int64_t)(ti + T * m) * a.pre_e];
        v[m].x = v[m].x * c;
        v[m].y = v[m].y * c;
      }
    }
  }
  const int64_t kbi = (g0 + (gi < gcount ? gi : 0)) * a.kb_u;
  const int64_t kbo = (g0 + (go < gcount ? go : 0)) * a.kb_u;
  stage_passes<FWD, T, E, SUB, 0, 1, Rs...>(v, lds_s, gi, ti, go, to, a, tw + h * a.kb_h, kbi, kbo);
  int tos = to;
  asm volatile("" : "+v"(tos));  // keep the store addresses out of the pass code
  if (go < gcount) {
    double2 *dst = out + h * a.out_h + (g0 + go) * a.out_g;
    const double *w = a.post ? a.post + (g0 + go) * a.post_g : nullptr;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int e = tos + T * m;
      double2 y = v[m];
      if (w) {
        const double c = w[(int64_t)e * a.post_e];
        y.x = y.x * c;
        y.y = y.y * c;
      }
      if (a.io_out == 0) {
        if (a.cached_out)
          dst[(int64_t)e * a.out_e] = y;
        else
          st_stream(dst + (int64_t)e * a.out_e, y);
      } else
        store_io(a, out, h * a.out_h, (g0 + go) * a.out_g + (int64_t)e * a.out_e, y);
    }
  }
}

// The stage schedules: sub-problem length, sub-problems per workgroup, radix sequence, and
// whether it is the one the planner prefers for that length. Measured on MI355X (256 MB
// batches; per stage): 128 points 112 / 102 / 104 us at G = 8 / 16 / 32, 256 points
// 108 / 103 / 118 us at 8 / 16 / 32, 512 points 108 / 120 us at 8 / 16 (the LDS-tile stage
// kernel: 107, 113, 116 us), 1024 points at G = 4 with the XCD remap (three 512-thread
// workgroups per CU against one of 1024 threads at G = 8; 2-D 1024 x 1024, columns then rows:
// 229-235 -> 200 us per 2^24 elements). Wider rows lose more to occupancy than they gain.
struct StageSchedule {
  int L, G, npass;
  int rs[4];
  bool pref;
  const void *k[2];
};

#define MFFT_SK(LL, GG, ...)                                               \
  {(const void *)stage_reg_kernel<false, LL, 8, GG, __VA_ARGS__>,         \
   (const void *)stage_reg_kernel<true, LL, 8, GG, __VA_ARGS__>}
const StageSchedule kStages[] = {
    {128, 8, 3, {8, 8, 2}, false, MFFT_SK(128, 8, 8, 8, 2)},
    {128, 16, 3, {8, 8, 2}, true, MFFT_SK(128, 16, 8, 8, 2)},
    {128, 32, 3, {8, 8, 2}, false, MFFT_SK(128, 32, 8, 8, 2)},
    {256, 8, 3, {8, 8, 4}, false, MFFT_SK(256, 8, 8, 8, 4)},
    {256, 16, 3, {8, 8, 4}, true, MFFT_SK(256, 16, 8, 8, 4)},
    {256, 32, 3, {8, 8, 4}, false, MFFT_SK(256, 32, 8, 8, 4)},
    {512, 8, 3, {8, 8, 8}, true, MFFT_SK(512, 8, 8, 8, 8)},
    {512, 16, 3, {8, 8, 8}, false, MFFT_SK(512, 16, 8, 8, 8)},
    {1024, 4, 4, {8, 8, 8, 2}, true, MFFT_SK(1024, 4, 8, 8, 8, 2)},
    {1024, 8, 4, {8, 8, 8, 2}, false, MFFT_SK(1024, 8, 8, 8, 8, 2)},
    {2048, 4, 4, {8, 8, 8, 4}, true, MFFT_SK(2048, 4, 8, 8, 8, 4)},
};
#undef MFFT_SK

int reg_stage_g(int L, const int *radices, int nradix, int gmax) {
  int best = 0;
  for (const auto &st : kStages) {
    if (st.L != L || st.npass != nradix || st.G > gmax) continue;
    bool same = true;
    for (int i = 0; i < nradix; ++i) same = same && st.rs[i] == radices[i];
    if (!same) continue;
    if (st.pref) return st.G;
    best = st.G > best ? st.G : best;
  }
  return best;
}

hipError_t launch_reg_stage(bool fwd, const StageArgs &a, int64_t ntrans, const double2 *in,
                            double2 *out, const double2 *tw, hipStream_t stream) {
  const void *k = nullptr;
  for (const auto &st : kStages) {
    if (st.L != a.L || st.G != a.G || st.npass != a.npass) continue;
    bool same = true;
    for (int i = 0; i < a.npass; ++i) same = same && st.rs[i] == a.pass[i].R;
    if (same) k = st.k[fwd ? 1 : 0];
  }
  if (!k) return hipErrorInvalidValue;
  {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.insert(k).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax);
  }
  // XCD remap (neighbouring tiles of a row on one XCD) where a tile's rows leave 128-byte
  // lines shared with the neighbouring tile, which then meet in one L2: the real-data maps
  // (rows of 8 doubles of real input, or n/2 + 1 values per transform of half output; MI355X,
  // r2c per 2^24-point batch: 2^17 / 2^18 points 176 / 207 -> 167 / 170 us, c2r within 1-2%;
  // c2r's real-part rows of 8 doubles measured within run-to-run noise with it, so they keep
  // the plain order)
  // and complex rows of fewer than 8 values (the 1024-point stage at G = 4: 2-D 1024 x 1024
  // 229-235 -> 200 us). For complex rows of whole lines it measured neutral to 4% slower
  // (2^15..2^18 points).
  StageArgs b = a;
  b.xcd = 0;
  int64_t nx = (a.nsub + a.G - 1) / a.G;
  const bool part_rows = (a.gfast_in || a.gfast_out) && a.G < 8;
  if ((a.io_in != 0 || a.io_out == 2 || part_rows) && a.nsub >= 8 * (int64_t)a.G) {
    b.xcd = 1;
    nx = (nx + 7) / 8 * 8;
  }
  const size_t lds = (size_t)a.G * reg_lds_doubles(a.L) * sizeof(double);
  const int bt = a.L / 8 * a.G;
  void *args[] = {(void *)&in, (void *)&out, (void *)&tw, (void *)&b};
  return hipLaunchKernel(k, dim3((unsigned)nx, (unsigned)ntrans), dim3(bt), args, lds, stream);
}

// Elements per thread: 8 (T = N / 8 <= 1024 threads) up to 8192 points, 16 at 16384.
int reg_elems(int64_t n) { return n > 8192 ? 16 : 8; }

bool reg_supported(int64_t n, const int *radices, int nradix) {
  if (n < 32 || n > 16384 || (n & (n - 1)) || nradix < 1 || nradix > MAX_PASSES) return false;
  if (fixed_schedule(n, radices, nradix) || split_schedule(n, radices, nradix)) return true;
  for (int i = 0; i < nradix; ++i) {
    const int r = radices[i];
    if (reg_elems(n) == 16 ? r != 4 : (r != 2 && r != 4 && r != 8)) return false;
  }
  return true;
}

hipError_t launch_reg(bool fwd, const StageArgs &a, const double2 *in, double2 *out,
                      const double2 *tw, hipStream_t stream) {
  if (a.nsub <= 0) return hipSuccess;
  {
    hipError_t e;
    if (launch_split(fwd, a, in, out, tw, stream, &e)) return e;
  }
  int rs[MAX_PASSES];
  for (int p = 0; p < a.npass; ++p) rs[p] = a.pass[p].R;
  if (const FixedSchedule *f = fixed_schedule(a.L, rs, a.npass)) {
    const bool ps = a.pre && a.pre_s;
    if (ps && a.sumsq) return hipErrorInvalidValue;  // no such call reaches here
    const void *k = f->k[fwd ? 1 : 0][a.sumsq ? 1 : 0][ps ? 1 : 0];
    const int T = a.L / f->e;
    const int BT = T < 256 ? 256 : T, G = BT / T;  // as reg_block<T>()
    const size_t lds = G == 1 ? (size_t)reg_lds_doubles(a.L) * sizeof(double)
                              : (size_t)G * lds_len(a.L) * sizeof(double2);
    {
      static std::mutex mu;
      static std::set<const void *> done;
      std::lock_guard<std::mutex> lk(mu);
      if (done.insert(k).second)
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax);
    }
    void *args[] = {(void *)&in, (void *)&out, (void *)&tw, (void *)&a};
    const int64_t nblk = (a.nsub + G - 1) / G;
    return hipLaunchKernel(k, dim3((unsigned)nblk), dim3(BT), args, lds, stream);
  }
  const int E = reg_elems(a.L), T = a.L / E;
  // the run-time dispatched kernel holds radices 2, 4, 8 at 8 elements per thread and radix 4
  // at 16 (a schedule only the split kernel runs, asked for with element maps, is refused)
  for (int p = 0; p < a.npass; ++p) {
    const int r = a.pass[p].R;
    if (E == 16 ? r != 4 : (r != 2 && r != 4 && r != 8)) return hipErrorInvalidValue;
  }
  const int BT = T < 256 ? 256 : T, G = BT / T;  // as reg_block<T>()
  const size_t lds = E == 8 ? (size_t)G * lds_len(a.L) * sizeof(double2)
                            : (size_t)reg_lds_doubles(a.L) * sizeof(double);
  const bool sq = a.sumsq != nullptr;
  const void *k = nullptr;
#define MFFT_RK(TT, EE)                                                        \
  (fwd ? (sq ? (const void *)reg_kernel<true, TT, EE, true>                    \
             : (const void *)reg_kernel<true, TT, EE, false>)                  \
       : (sq ? (const void *)reg_kernel<false, TT, EE, true>                   \
             : (const void *)reg_kernel<false, TT, EE, false>))
  if (E == 8 && T == 4) k = MFFT_RK(4, 8);
  else if (E == 8 && T == 8) k = MFFT_RK(8, 8);
  else if (E == 8 && T == 16) k = MFFT_RK(16, 8);
  else if (E == 8 && T == 32) k = MFFT_RK(32, 8);
  else if (E == 8 && T == 64) k = MFFT_RK(64, 8);
  else if (E == 8 && T == 128) k = MFFT_RK(128, 8);
  else if (E == 8 && T == 256) k = MFFT_RK(256, 8);
  else if (E == 8 && T == 512) k = MFFT_RK(512, 8);
  else if (E == 8 && T == 1024) k = MFFT_RK(1024, 8);
  else if (E == 16 && T == 1024) k = MFFT_RK(1024, 16);
  else return hipErrorInvalidValue;
#undef MFFT_RK
  {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.insert(k).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytesMax);
  }
  void *args[] = {(void *)&in, (void *)&out, (void *)&tw, (void *)&a};
  const int64_t nblk = (a.nsub + G - 1) / G;
  return hipLaunchKernel(k, dim3((unsigned)nblk), dim3(BT), args, lds, stream);
}

}  // namespace mfft
