// Host side of mfft: plans, twiddle tables, kernel-stage schedules and the C ABI.
//
// A plan is the reference's fast plan (a radix sequence rs of a Stockham schedule,
// LeanFFTW/Gen/Stockham.lean:101-115) plus the way it is cut into GPU kernel stages:
//
//   * N <= LDS_ELEMS: one stage; each workgroup runs every pass of G whole transforms
//     in LDS.
//   * larger N = P * Q: the first passes (product P) act on the Q decimated sequences
//     x[c + m Q] (m < P) -- the Stockham property that after passes with product P the
//     array holds, at c*P + k, the P-point DFT of that sequence -- and the remaining
//     passes act on the P residue classes {k1 + P i}, each a Q-point Stockham with the
//     plan's own twiddles (index k1 + P * k_local). Both stages compute exactly the
//     reference's passes, so the result is bit-identical to refTransform with rs.
//
// Twiddles are the reference's pass tables (LeanFFTW/Core/FastPlan.lean:27-37) computed
// with the exact-integer routine of mfft_trig.cpp, stored on the device per pass as
// (R-1) rows of ns interleaved (re, im) pairs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mfft.h"
#include "bfly_gen.h"
#include "mfft_internal.h"

using mfft::StageArgs;

struct mfft_plan {
  int64_t n = 0;
  int sign = -1;
  int64_t howmany = 1, istride = 1, idist = 0, ostride = 1, odist = 0;
  std::vector<int> radices;
  std::vector<int> stage_passes;
  std::vector<StageArgs> stages;
  double2 *d_tw = nullptr;
  size_t tw_elems = 0;
  std::vector<double2> h_tw;  // host copy of the pass tables (ragged batches reuse it)
  double2 *d_scratch = nullptr;  // multi-stage intermediate, chunk_h transforms
  int64_t chunk_h = 0;
  int device = 0;
  int64_t spec_in_off = 0, spec_out_off = 0;  // doubles, for mfft_plan_spec plans
  bool allow_big = true;  // may use the large-tile single-stage variant
  bool fuse = false;       // the complex part of a real-data plan (schedules the maps fuse into)
  // composite plans, built from complex plans (parts) the way the reference's generic engine
  // builds them (LeanFFTW/Core/Problem.lean:215-236): 0 complex DFT; 1 real to half complex
  // (r2c, part 0 the forward DFT on the embedded batch in d_work); 2 half complex to real (c2r,
  // part 0 the backward DFT on the Hermitian extension in d_work); 3 two-dimensional complex
  // DFT of n0 x n1 row-major arrays (part 0 along axis 0, then part 1 along axis 1)
  int kind = 0;
  mfft_plan *part[2] = {nullptr, nullptr};
  double2 *d_work = nullptr;
  bool fused_io = false;  // r2c / c2r: the real-data maps fuse into the complex part's kernels
  int64_t n0 = 0, n1 = 0;
  int fused2d = 0;  // 2-D: 1 + index of the single-pass kernel for small arrays, 0 for none
  const void *split2d = nullptr;  // 2-D: the single-pass register kernel (128 x 128), if any
  // an outer batch of single-stage plans: `outer` repetitions of the whole batch, the k-th
  // at in + k * outer_in, out + k * outer_out (the arrays of a 2-D plan's axis-0 part)
  int64_t outer = 1, outer_in = 0, outer_out = 0;
  // two-stage plan over a batch of adjacent transforms (columns): workgroups take G
  // adjacent transforms of one sub-problem index, which grid.y walks (see build_stages)
  bool columns = false;
};

namespace {

// planner flag of the complex parts of real-data plans: register-kernel schedules into whose
// loads and stores the real-data maps fuse (not part of the public flags)
constexpr unsigned kFuseRadices = 1u << 30;
// The largest power of two whose real-data plan stages its tile with the maps in the tile
// copies; the register kernels take the maps into their own loads and stores above it.
constexpr int64_t kFuseTileMax = 64;

std::atomic<size_t> g_threads{1};
std::mutex g_wisdom_mu;
std::string g_wisdom_path;

// Bytes of intermediate per chunk of a two-stage plan (MFFT_CHUNK_MB, default 1 GiB): the
// bound on the plan's scratch allocation. Smaller chunks do not keep the intermediate in the
// 256 MB Infinity Cache between the stages on MI355X (16 to 128 MB chunks of a 256 MB batch
// measured slower than one chunk), so a batch runs as one chunk when it fits.
int64_t chunk_bytes() {
  static const int64_t v = [] {
    const char *e = std::getenv("MFFT_CHUNK_MB");
    const int64_t mb = e ? std::atoll(e) : 1024;
    return std::max<int64_t>(1, mb) << 20;
  }();
  return v;
}

// Largest transform length whose single-stage tile is staged through LDS (coalesced copies).
int64_t stage_io_max() {
  static const int64_t v = [] {
    const char *e = std::getenv("MFFT_STAGE_IO_MAX");
    return e ? (int64_t)std::atoll(e) : (int64_t)16;
  }();
  return v;
}

bool smooth13(int64_t n) {
  for (int p : {2, 3, 5, 7, 11, 13})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Largest radix the planner uses by itself (MFFT_MAX_RADIX, default 16).
int max_radix() {
  static const int v = [] {
    const char *e = std::getenv("MFFT_MAX_RADIX");
    return e ? std::max(2, std::min(16, std::atoi(e))) : 16;
  }();
  return v;
}

bool radix_ok(int r) { return r >= 2 && r <= 16 && smooth13(r); }

// Cost model of a pass: arithmetic of the generated twiddled butterfly per point plus a
// fixed charge for the LDS round trip of the pass.
double pass_cost(int r) { return (double)mfft::kBflyOps[r] / r + 6.0; }

// Best radix sequence of length L over the radices accepted by `ok` (<= max_radix()): the
// fewest passes (every pass is one LDS round trip), then the least arithmetic. Returns
// false if L has no such factorisation. `cost` receives passes * 1000 + arithmetic.
bool best_radices(int64_t L, bool (*ok)(int), std::vector<int> &out, double &cost,
                  int cap = 0) {
  if (cap <= 0) cap = max_radix();
  std::vector<std::pair<double, std::vector<int>>> best;
  std::vector<int64_t> divs;
  for (int64_t d = 1; d <= L; ++d)
    if (L % d == 0) divs.push_back(d);
  best.assign(divs.size(), {1e300, {}});
  best[0] = {0.0, {}};
  for (size_t i = 1; i < divs.size(); ++i) {
    const int64_t d = divs[i];
    for (int r = cap; r >= 2; --r) {
      if (!radix_ok(r) || !ok(r) || d % r) continue;
      const size_t j = std::lower_bound(divs.begin(), divs.end(), d / r) - divs.begin();
      if (best[j].first >= 1e299) continue;
      const double c = best[j].first + 1000.0 + pass_cost(r);
      if (c < best[i].first) {
        best[i].first = c;
        best[i].second = best[j].second;
        best[i].second.push_back(r);
      }
    }
  }
  if (best.back().first >= 1e299) return false;
  out = best.back().second;
  std::sort(out.rbegin(), out.rend());
  cost = best.back().first;
  return true;
}

bool any_radix(int) { return true; }
bool small_radix_ok(int r) { return mfft::small_radix(r); }

bool mid_radix_ok(int r) { return mfft::mid_radix(r); }

// Radix sequence for an LDS stage of length L. Stages run the kernel variant of the smallest
// radix set that holds their radices (small: 2, 3, 4, 5, 6, 8; mid: 2 ... 13; all), whose
// occupancy is higher the smaller the set (about 88, 110 and 134 VGPRs). The small set
// measured faster than saving a pass with radix 16 (512 points: 5.2 TB/s as 8x8x8 against
// 3.9 as 16x16x2), so a sequence is charged 300 per pass when it needs the mid set and 600
// per pass when it needs the full set, against 1000 per pass. `mid_only` (ragged batches,
// whose kernel holds the mid set) leaves out radices 14, 15 and 16.
std::vector<int> choose_radices(int64_t L, bool mid_only = false) {
  std::vector<int> all, mid, small;
  double c_all = 1e300, c_mid = 1e300, c_small = 1e300;
  auto charge = [](const std::vector<int> &rs, double &c) {
    bool s = true, m = true;
    for (int r : rs) {
      s = s && mfft::small_radix(r);
      m = m && mfft::mid_radix(r);
    }
    c += (s ? 0.0 : m ? 300.0 : 600.0) * rs.size();
  };
  if (!mid_only && best_radices(L, any_radix, all, c_all)) charge(all, c_all);
  if (best_radices(L, mid_radix_ok, mid, c_mid)) charge(mid, c_mid);
  if (best_radices(L, small_radix_ok, small, c_small)) charge(small, c_small);
  if (c_small <= c_mid && c_small <= c_all) return small;
  if (c_mid <= c_all) return mid;
  return all;
}

// Radix sequence of a register-resident size: radices up to 8 (the fixed schedules of
// mfft_reg.hip for 2048 ... 8192), and 16 x 16 x 16 x 4 at 16384 (a fixed schedule with 16
// elements per thread: 125 us per 256 MB batch against 180 for radix 4 only). Plain complex
// plans (`fuse` false) take radix 16 where it measured faster on MI355X (256 MB batches):
// 1024 = 16 x 8 x 8 in a tile kernel (78 us against 85 for 8 x 8 x 8 x 2), 2048 = 16 x 8 x 16
// in a tile kernel (80.5 against 82.9 for 8 x 8 x 8 x 4), and the split kernel's schedules
// (mfft_split.hip: wave-local column and row passes, one workgroup-wide hand-over) at 8192 =
// 16 x 4 | 16 x 8 (101-102 us against 109-112 for 16 x 16 x 8 x 4 in the fixed kernel) and
// 16384 = 16 x 8 | 16 x 8 (111-113 against 129-135 for 16 x 16 x 16 x 4). (4096 = 16^3 at 256
// threads per transform measured 85-92 against 92, but it halves the threads of a batch and
// its weighted kernel took 34 us against 19 on a 512-transform batch; 4096 = 8 x 8 | 8 x 8 in
// the split kernel measured 92-94 against 91-93, four workgroups per CU already overlapping
// each other's barriers.) The parts of real-data plans (`fuse`) keep the fixed schedules,
// into whose loads and stores the real-data maps fuse.
std::vector<int> reg_radices(int64_t n, bool fuse = false) {
  if (n == 16384) return fuse ? std::vector<int>{16, 16, 16, 4} : std::vector<int>{16, 8, 16, 8};
  if (!fuse && n == 1024) return {16, 8, 8};
  if (!fuse && n == 2048) return {16, 8, 16};
  if (!fuse && n == 8192) return {16, 4, 16, 8};
  std::vector<int> rs;
  double c;
  best_radices(n, any_radix, rs, c, 8);
  return rs;
}

// Split N into P * Q for a two-stage plan: both within the LDS, as balanced as possible.
bool choose_split(int64_t N, int64_t &P) {
  P = 0;
  double best = 1e300;
  for (int64_t p = 2; p <= mfft::LDS_ELEMS && p < N; ++p) {
    if (N % p) continue;
    const int64_t q = N / p;
    if (q > mfft::LDS_ELEMS) continue;
    const double score = std::fabs(std::log((double)p) - std::log((double)q));
    if (score < best) {
      best = score;
      P = p;
    }
  }
  return P != 0;
}

// Split N into P1 * P2 * P3 for a three-stage plan (N whose two-way splits leave a factor
// above the LDS): each within the LDS, as balanced as possible.
bool choose_split3(int64_t N, int64_t &P1, int64_t &P2) {
  P1 = P2 = 0;
  double best = 1e300;
  const double third = std::log((double)N) / 3.0;
  for (int64_t a = 2; a <= mfft::LDS_ELEMS && a < N; ++a) {
    if (N % a) continue;
    for (int64_t b = 2; b <= mfft::LDS_ELEMS && a * b < N; ++b) {
      if ((N / a) % b) continue;
      const int64_t c = N / a / b;
      if (c > mfft::LDS_ELEMS || c < 2) continue;
      const double score = std::fabs(std::log((double)a) - third) +
                           std::fabs(std::log((double)b) - third) +
                           std::fabs(std::log((double)c) - third);
      if (score < best) {
        best = score;
        P1 = a;
        P2 = b;
      }
    }
  }
  return P1 != 0;
}

// Power-of-two sizes that run as one register-resident stage (mfft_reg.hip): from
// MFFT_REG_MIN (default 64) to 16384 points; MFFT_REG=0 turns the kernel off.
int64_t reg_min() {
  static const int64_t v = [] {
    const char *e = std::getenv("MFFT_REG");
    if (e && std::atoi(e) == 0) return (int64_t)1 << 40;
    const char *m = std::getenv("MFFT_REG_MIN");
    return m ? (int64_t)std::atoll(m) : (int64_t)64;
  }();
  return v;
}

bool reg_size(int64_t n) { return (n & (n - 1)) == 0 && n >= reg_min() && n <= 16384; }

// Whether a size runs as a single kernel stage (else it is split into two). The register-
// resident kernel walks the element index with consecutive lanes, so it takes unit-stride
// batches only (`contig`); strided batches run the stage kernel, which walks the batch.
bool single_stage(int64_t n, bool allow_big, bool contig = true) {
  const bool pow2 = (n & (n - 1)) == 0;
  return n <= mfft::LDS_ELEMS ||
         (allow_big &&
          (n <= (pow2 ? mfft::BIG_ELEMS : mfft::BIG_ELEMS / 2) || (contig && reg_size(n))));
}

hipError_t build_twiddles(mfft_plan *p) {
  const int64_t N = p->n;
  const bool fwd = p->sign < 0;
  const auto &rs = p->radices;
  std::vector<double> tab;
  if (rs.size() > 1) mfft::cos_sin_2pi_table((uint64_t)N, tab);
  std::vector<double2> host;
  int64_t ns = 1;
  std::vector<int64_t> off(rs.size(), -1);
  for (size_t t = 0; t < rs.size(); ++t) {
    const int R = rs[t];
    if (t >= 1) {
      off[t] = (int64_t)host.size();
      const int64_t step = N / (ns * R);
      for (int r = 1; r < R; ++r)
        for (int64_t k = 0; k < ns; ++k) {
          const int64_t idx = (int64_t)r * k * step;
          const double c = tab[2 * idx], s = tab[2 * idx + 1];
          host.push_back(make_double2(c, fwd ? (s == 0.0 ? 0.0 : -s) : s));
        }
    }
    ns *= R;
  }
  p->tw_elems = host.size();
  p->h_tw = host;
  if (!host.empty()) {
    hipError_t e = hipMalloc(&p->d_tw, host.size() * sizeof(double2));
    if (e != hipSuccess) return e;
    e = hipMemcpy(p->d_tw, host.data(), host.size() * sizeof(double2), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  // stamp the offsets into the stages
  size_t t = 0;
  for (auto &st : p->stages)
    for (int q = 0; q < st.npass; ++q, ++t) st.pass[q].twoff = off[t];
  return hipSuccess;
}

bool all_small(const std::vector<int> &rs) {
  for (int r : rs)
    if (!mfft::small_radix(r)) return false;
  return true;
}

bool all_mid(const std::vector<int> &rs) {
  for (int r : rs)
    if (!mfft::mid_radix(r)) return false;
  return true;
}

bool all_pow2(const std::vector<int> &rs) {
  for (int r : rs)
    if (r != 2 && r != 4 && r != 8 && r != 16) return false;
  return true;
}

// Workgroup shape of a stage whose G, L and radices `rs` are set: 256 threads with 8
// elements each; or, for a large tile (G * L > LDS_ELEMS), 8 elements per thread and up to
// 1024 threads when every radix is at most 10, else 16 elements per thread and up to 512.
// Thread counts are multiples of 64.
void set_tile(StageArgs &a, bool big, const std::vector<int> &rs) {
  if (!big) {
    a.ept = 8;
    a.threads = mfft::WG_THREADS;
    return;
  }
  a.ept = all_mid(rs) ? 8 : 16;
  const int64_t elems = (int64_t)a.G * a.L;
  const int64_t per = a.ept * 64;
  a.threads = (int)std::min<int64_t>(a.ept == 8 ? 1024 : 512, ((elems + per - 1) / per) * 64);
}

// A stage over sub-problems adjacent in memory (sub-problem g at g, element e at e*stride):
// G = 8 sub-problems (128-byte rows) when L allows, in a large tile if L > LDS_ELEMS / 8
// (rows of fewer than 8 elements cost more HBM time than a large tile's lower occupancy:
// 165375 = 375 x 441 runs 13% faster with 8-wide rows than with 4-wide ones).
void strided_tile(StageArgs &a, int64_t L, int64_t U, bool allow_big, const std::vector<int> &rs) {
  int64_t G = std::max<int64_t>(1, mfft::LDS_ELEMS / L);
  bool big = false;
  if (G < 8 && allow_big) {
    G = std::max<int64_t>(1, std::min<int64_t>(8, mfft::BIG_ELEMS / L));
    big = G * L > mfft::LDS_ELEMS;
  }
  a.G = (int)std::min<int64_t>(G, U);
  set_tile(a, big && a.G * L > mfft::LDS_ELEMS, rs);
}

// A strided stage whose length and radices have a register-resident stage schedule
// (mfft_reg.hip) runs it, with the schedule's preferred number of sub-problems per workgroup
// (MFFT_RSTAGE_G caps it; 0 turns these kernels off); the tile shape also suits the stage
// kernel, which runs the stage when sum-of-squares partials are requested.
void reg_stage(StageArgs &a, const std::vector<int> &rs, int64_t U) {
  static const int gmax = [] {
    const char *e = std::getenv("MFFT_RSTAGE_G");
    return e ? std::atoi(e) : 1 << 20;
  }();
  if (gmax <= 0 || !all_small(rs)) return;
  const int G = mfft::reg_stage_g(a.L, rs.data(), (int)rs.size(), gmax);
  if (G <= 0 || G > U) return;
  a.reg = 2;
  a.G = G;
  a.ept = 8;
  a.threads = a.L / 8 * G;
}

void fill_pass(mfft::PassArgs &pa, int R, int64_t nsl, int64_t twns, int64_t kmul, int L,
               int G) {
  pa.R = R;
  pa.nsl = (int)nsl;
  pa.m_LR = mfft::magic32(L / R);
  pa.m_nsl = mfft::magic32(nsl);
  pa.m_G = mfft::magic32(G);
  pa.twoff = -1;
  pa.twns = twns;
  pa.kmul = kmul;
}

// Transforms adjacent in memory on both sides (|dist| < |stride|), e.g. the columns of a
// row-major array.
bool columns_layout(const mfft_plan *p) {
  return std::llabs(p->idist) < std::llabs(p->istride) &&
         std::llabs(p->odist) < std::llabs(p->ostride);
}

// Two stages over a batch of columns. The passes are those of the row layout's split
// (stage A: sub-problems q < Q of length P, elements q + Q e; stage B: residue classes
// k1 < P of length Q, twiddle base k1), but a workgroup takes G adjacent transforms of one
// sub-problem, so every HBM access is a row of G consecutive complex values: g walks the
// transforms, grid.y walks q (stage A) or k1 (stage B). The intermediate of a chunk of H
// transforms is laid out as [q P + e][transform]; the H-dependent strides are set per chunk
// in run().
bool build_columns(mfft_plan *p, int split, const std::vector<int64_t> &ns) {
  const auto &rs = p->radices;
  const int s = (int)rs.size();
  const int64_t N = p->n, P = ns[split], Q = N / P, U = std::max<int64_t>(1, p->howmany);
  const std::vector<int> ra(rs.begin(), rs.begin() + split), rb(rs.begin() + split, rs.end());
  StageArgs A{}, B{};
  A.L = (int)P;
  A.npass = split;
  strided_tile(A, P, U, p->allow_big, ra);
  if (p->allow_big) reg_stage(A, ra, U);
  for (int t = 0; t < split; ++t) fill_pass(A.pass[t], rs[t], ns[t], ns[t], 1, A.L, A.G);
  A.in_g = p->idist;
  A.in_e = p->istride * Q;
  A.in_h = p->istride;
  A.out_g = 1;
  A.gfast_in = A.gfast_out = 1;
  A.pre_e = 1;
  A.post_e = 1;

  B.L = (int)Q;
  B.npass = s - split;
  strided_tile(B, Q, U, p->allow_big, rb);
  if (p->allow_big) reg_stage(B, rb, U);
  for (int t = split; t < s; ++t)
    fill_pass(B.pass[t - split], rs[t], ns[t] / P, ns[t], P, B.L, B.G);
  B.in_g = 1;
  B.out_g = p->odist;
  B.out_e = p->ostride * P;
  B.out_h = p->ostride;
  B.kb_h = 1;
  B.gfast_in = B.gfast_out = 1;
  A.cached_out = 1;  // the intermediate, as in the row layout
  B.pre_e = 1;
  B.post_e = 1;
  p->columns = true;
  p->stages = {A, B};
  p->stage_passes = {split, s - split};
  const int64_t cmax =
      std::max<int64_t>(1, std::min<int64_t>(U, chunk_bytes() / (16 * N)));
  const int64_t nchunk = (U + cmax - 1) / cmax;
  p->chunk_h = (U + nchunk - 1) / nchunk;
  return true;
}

// Three stages, N = P1 P2 P3 with passes [0, a), [a, b), [b, s) (ns_a = P1, ns_b = P1 P2): the
// two-stage decomposition applied again to its second stage. A: columns c < N/P1 of length
// P1, x[c + (N/P1) m] -> T[c P1 + k]. B: sub-problems (k1 < P1 along g, c2 < P3 along
// grid.y, transform along grid.z) of length P2, T[k1 + P1 c2 + (N/P2) m] -> U[k1 + P1 P2 c2 +
// P1 k], twiddle base k1 and multiplier P1. C: residue classes K < P1 P2 of length P3,
// U[K + P1 P2 i] -> y[K + P1 P2 e], twiddle base K and multiplier P1 P2. Every pass is the
// reference's, so the bits are refTransform's. B runs the stage kernel (the only one with a
// z extent); A and C may take tile kernels.
bool build_three(mfft_plan *p, int a, int b) {
  const int64_t N = p->n;
  const auto &rs = p->radices;
  const int s = (int)rs.size();
  if (a <= 0 || b <= a || b >= s || columns_layout(p)) return false;
  std::vector<int64_t> ns(s + 1, 1);
  for (int t = 0; t < s; ++t) ns[t + 1] = ns[t] * rs[t];
  const int64_t P1 = ns[a], P2 = ns[b] / ns[a], P3 = N / ns[b], NB = ns[b];
  if (P1 > mfft::LDS_ELEMS || P2 > mfft::LDS_ELEMS || P3 > mfft::LDS_ELEMS || a > mfft::MAX_PASSES ||
      b - a > mfft::MAX_PASSES || s - b > mfft::MAX_PASSES)
    return false;
  const std::vector<int> ra(rs.begin(), rs.begin() + a), rb(rs.begin() + a, rs.begin() + b),
      rc(rs.begin() + b, rs.end());
  StageArgs A{}, B{}, C{};
  A.L = (int)P1;
  A.npass = a;
  strided_tile(A, P1, N / P1, p->allow_big, ra);
  for (int t = 0; t < a; ++t) fill_pass(A.pass[t], rs[t], ns[t], ns[t], 1, A.L, A.G);
  A.nsub = N / P1;
  A.in_g = p->istride;
  A.in_e = p->istride * (N / P1);
  A.in_h = p->idist;
  A.out_g = P1;
  A.out_e = 1;
  A.out_h = N;
  A.gfast_in = 1;
  A.pre_g = 1;
  A.pre_e = N / P1;
  A.cached_out = 1;

  B.L = (int)P2;
  B.npass = b - a;
  strided_tile(B, P2, P1, p->allow_big, rb);
  for (int t = a; t < b; ++t) fill_pass(B.pass[t - a], rs[t], ns[t] / P1, ns[t], P1, B.L, B.G);
  B.nsub = P1;
  B.in_g = 1;
  B.in_e = N / P2;
  B.in_h = P1;      // grid.y: c2 < P3
  B.in_z = N;       // grid.z: the transform
  B.out_g = 1;
  B.out_e = P1;
  B.out_h = P1 * P2;
  B.out_z = N;
  B.kb_u = 1;
  B.gfast_in = 1;
  B.gfast_out = 1;
  B.cached_in = 1;
  B.cached_out = 1;

  C.L = (int)P3;
  C.npass = s - b;
  strided_tile(C, P3, NB, p->allow_big, rc);
  for (int t = b; t < s; ++t) fill_pass(C.pass[t - b], rs[t], ns[t] / NB, ns[t], NB, C.L, C.G);
  C.nsub = NB;
  C.in_g = 1;
  C.in_e = NB;
  C.in_h = N;
  C.out_g = p->ostride;
  C.out_e = p->ostride * NB;
  C.out_h = p->odist;
  C.kb_u = 1;
  C.gfast_in = 1;
  C.gfast_out = 1;
  C.post_g = 1;
  C.post_e = NB;
  C.cached_in = 1;
  p->stages = {A, B, C};
  p->stage_passes = {a, b - a, s - b};
  // equal chunks of at most chunk_bytes() over both intermediates
  int64_t cmax = std::max<int64_t>(1, std::min<int64_t>(p->howmany, chunk_bytes() / (32 * N)));
  cmax = std::min<int64_t>(cmax, 65535);  // grid.y of A and C, grid.z of B
  const int64_t nchunk = (std::max<int64_t>(1, p->howmany) + cmax - 1) / cmax;
  p->chunk_h = (std::max<int64_t>(1, p->howmany) + nchunk - 1) / nchunk;
  return true;
}

// Build the stage descriptors of a plan whose radices and split(s) are set (split2 > 0: three
// stages).
bool build_stages(mfft_plan *p, int split, int split2 = 0) {
  if (split2 > 0) return build_three(p, split, split2);
  const int64_t N = p->n;
  const auto &rs = p->radices;
  const int s = (int)rs.size();
  p->stages.clear();
  p->stage_passes.clear();
  std::vector<int64_t> ns(s + 1, 1);
  for (int t = 0; t < s; ++t) ns[t + 1] = ns[t] * rs[t];
  if (split <= 0 || split >= s) {
    // the complex part of a real-data plan runs a staged tile kernel, whose tile copies take
    // the real-data maps: powers of two up to 64 points (the register kernel's fused maps from
    // 128 on), every other single-stage size up to a tile of LDS
    const bool pow2 = (N & (N - 1)) == 0;
    const bool reg = p->allow_big && reg_size(N) && p->istride == 1 && p->ostride == 1 &&
                     mfft::reg_supported(N, rs.data(), s) && !(p->fuse && N <= kFuseTileMax);
    const bool big = !reg && N > mfft::LDS_ELEMS && p->allow_big &&
                     N <= ((N & (N - 1)) == 0 ? mfft::BIG_ELEMS : mfft::BIG_ELEMS / 2);
    if ((N > mfft::LDS_ELEMS && !big && !reg) || s > mfft::MAX_PASSES) return false;
    StageArgs a{};
    a.L = (int)N;
    a.npass = s;
    // a batch whose transforms are adjacent in memory (the columns of a row-major array) is
    // tiled like a two-stage plan's column stage: up to 8 transforms (128-byte rows) per
    // workgroup, in the register-resident stage kernel where a schedule exists
    const bool batch_fast = std::llabs(p->idist) < std::llabs(p->istride) ||
                            std::llabs(p->odist) < std::llabs(p->ostride);
    if (batch_fast && !reg) {
      strided_tile(a, N, std::max<int64_t>(1, p->howmany), p->allow_big, rs);
      if (p->allow_big && std::llabs(p->idist) < std::llabs(p->istride) &&
          std::llabs(p->odist) < std::llabs(p->ostride))
        reg_stage(a, rs, std::max<int64_t>(1, p->howmany));
    } else {
      a.G = big ? 1
                : (int)std::max<int64_t>(1, std::min<int64_t>(mfft::LDS_ELEMS / N, p->howmany));
      set_tile(a, big, rs);
    }
    for (int t = 0; t < s; ++t) fill_pass(a.pass[t], rs[t], ns[t], ns[t], 1, a.L, a.G);
    a.nsub = p->howmany;
    a.in_g = p->idist;
    a.in_e = p->istride;
    a.in_h = 0;
    a.out_g = p->odist;
    a.out_e = p->ostride;
    a.out_h = 0;
    a.kb_u = 0;
    a.gfast_in = std::llabs(p->idist) < std::llabs(p->istride) ? 1 : 0;
    a.gfast_out = std::llabs(p->odist) < std::llabs(p->ostride) ? 1 : 0;
    a.pre_g = 0;
    a.pre_e = 1;
    a.post_g = 0;
    a.post_e = 1;
    // small contiguous transforms: stage the tile through LDS with coalesced copies; the
    // complex parts of real-data plans whose sizes are not powers of two, up to the largest
    // tile kernel (4096 points), with the maps in the copies (a stage that finds no tile
    // kernel drops them again, untiled())
    a.m_L = mfft::magic32(N);
    const bool contiguous = p->istride == 1 && p->ostride == 1 && p->idist == N && p->odist == N;
    const int64_t io_max = p->fuse && pow2 ? std::max<int64_t>(kFuseTileMax, stage_io_max())
                                           : stage_io_max();
    if (contiguous && p->fuse && !pow2 && N <= 2 * (int64_t)mfft::LDS_ELEMS)
      a.stage_io = 2;
    else
      a.stage_io = contiguous && N <= io_max && a.ept == 8 ? (p->fuse ? 2 : 1) : 0;
    // the staged tile is read with ordinary loads (MI355X, tile kernels, 256 MB batches of
    // 2..16 points: 87-90 us with non-temporal loads, 80-82 with ordinary ones)
    if (a.stage_io) a.cached_in = 1;
    if (reg) {  // T = N / E threads per transform, 256 / T transforms per workgroup if T < 256
      const int E = mfft::reg_elems(N), T = (int)(N / E);
      a.reg = 1;
      a.ept = E;
      a.threads = T < 256 ? 256 : T;
      a.G = a.threads / T;
      a.stage_io = 0;
      // at 8 elements per thread, ordinary loads and non-temporal stores (MI355X, 4096 points
      // per 256 MB batch: 89.8 us with non-temporal loads and ordinary stores, 102.6 with both
      // ordinary, 82.2 with this); at 16 (8192, 16384 points) both non-temporal (ordinary
      // loads within +-2%)
      a.cached_in = E == 8 ? 1 : 0;
      a.cached_out = 0;
    } else if (a.reg == 0 && !a.gfast_in && !a.stage_io) {
      // the LDS-tile kernel's first pass reads each butterfly's inputs in runs of a few
      // elements, so several load instructions share a cache line: ordinary loads keep it
      // in L2 for them (MI355X: 60 points 117 -> 91 us, 360 and 1000 points 110 -> 101-106,
      // 32 points 94 -> 86 per 256 MB batch)
      a.cached_in = 1;
    }
    p->stages.push_back(a);
    p->stage_passes.push_back(s);
    p->chunk_h = p->howmany;
    return true;
  }
  const int64_t P = ns[split], Q = N / P;
  if (P > mfft::LDS_ELEMS || Q > mfft::LDS_ELEMS || split > mfft::MAX_PASSES ||
      s - split > mfft::MAX_PASSES)
    return false;
  if (columns_layout(p)) return build_columns(p, split, ns);
  StageArgs A{}, B{};
  A.L = (int)P;
  A.npass = split;
  strided_tile(A, P, Q, p->allow_big, std::vector<int>(rs.begin(), rs.begin() + split));
  if (p->allow_big) reg_stage(A, std::vector<int>(rs.begin(), rs.begin() + split), Q);
  for (int t = 0; t < split; ++t) fill_pass(A.pass[t], rs[t], ns[t], ns[t], 1, A.L, A.G);
  A.nsub = Q;
  A.in_g = p->istride;  // column c starts at element c
  A.in_e = p->istride * Q;
  A.in_h = p->idist;
  A.out_g = P;  // the contiguous intermediate: column c lands at c*P
  A.out_e = 1;
  A.out_h = N;
  A.kb_u = 0;
  A.gfast_in = 1;
  A.gfast_out = 0;
  A.pre_g = 1;
  A.pre_e = Q;

  B.L = (int)Q;
  B.npass = s - split;
  strided_tile(B, Q, P, p->allow_big, std::vector<int>(rs.begin() + split, rs.end()));
  if (p->allow_big) reg_stage(B, std::vector<int>(rs.begin() + split, rs.end()), P);
  for (int t = split; t < s; ++t)
    fill_pass(B.pass[t - split], rs[t], ns[t] / P, ns[t], P, B.L, B.G);
  B.nsub = P;  // residue classes k1 + P i
  B.in_g = 1;
  B.in_e = P;
  B.in_h = N;
  B.out_g = p->ostride;
  B.out_e = p->ostride * P;
  B.out_h = p->odist;
  B.kb_u = 1;
  B.gfast_in = 1;
  B.gfast_out = 1;
  B.post_g = 1;
  B.post_e = P;
  // the intermediate is written with ordinary stores: 10-15% faster two-stage plans on MI355X
  // than non-temporal ones (2^15..2^18 points). The register stage kernel reads it back
  // non-temporally (ordinary loads measured slower there); the LDS-tile stage kernel with
  // ordinary loads (10368 points 257 -> 225 us, 165375 points 410 -> 357 per 256 MB batch)
  A.cached_out = 1;
  B.cached_in = B.reg == 0 ? 1 : 0;
  // the complex part of a real-data plan in tile kernels: both stages take the tile variants
  // with the maps in their HBM accesses (finish() drops them where a stage has none)
  if (p->fuse && A.reg == 0 && B.reg == 0 && p->istride == 1 && p->ostride == 1 &&
      p->idist == N && p->odist == N)
    A.stage_io = B.stage_io = 3;
  p->stages.push_back(A);
  p->stages.push_back(B);
  p->stage_passes.push_back(split);
  p->stage_passes.push_back(s - split);
  // equal chunks: a short last chunk is a latency-bound launch pair of its own (10 to 25 us
  // at 10368..165375 points)
  int64_t cmax = std::max<int64_t>(1, std::min<int64_t>(p->howmany, chunk_bytes() / (16 * N)));
  cmax = std::min<int64_t>(cmax, 65535);  // grid.y
  const int64_t nchunk = (std::max<int64_t>(1, p->howmany) + cmax - 1) / cmax;
  p->chunk_h = (std::max<int64_t>(1, p->howmany) + nchunk - 1) / nchunk;
  return true;
}

// A stage left without a tile kernel (a batch smaller than the tile, a measured schedule
// without one): above kFuseTileMax points the stage kernel reads it butterfly by butterfly, as
// it reads every complex stage of that size, and the real-data plan runs the element maps.
void untiled(StageArgs &st) {
  st.tile = 0;
  if ((st.stage_io == 2 && st.L > kFuseTileMax) || st.stage_io == 3) st.stage_io = 0;
}

mfft_plan *finish(mfft_plan *p, int split, int split2 = 0) {
  if (!build_stages(p, split, split2) || hipGetDevice(&p->device) != hipSuccess ||
      build_twiddles(p) != hipSuccess) {
    mfft_destroy_plan(p);
    return nullptr;
  }
  // stages with a fixed-schedule tile kernel run it (two-stage column plans size their stages
  // per chunk and look theirs up at run time)
  for (auto &st : p->stages) {
    st.tile = p->columns ? 0 : mfft::tile_lookup(st);
    if (!st.tile && st.stage_io == 3) {  // no map variant: the stage's plain tile kernel
      st.stage_io = 0;
      st.tile = mfft::tile_lookup(st);
    }
    if (!st.tile) untiled(st);
  }
  // the maps of a two-stage plan fuse only when both stages carry them
  if (p->stages.size() == 2 && (p->stages[0].stage_io == 3) != (p->stages[1].stage_io == 3))
    for (auto &st : p->stages)
      if (st.stage_io == 3) {
        st.stage_io = 0;
        st.tile = mfft::tile_lookup(st);
        if (!st.tile) untiled(st);
      }
  if (p->stages.size() == 3) p->stages[1].tile = 0;  // the middle stage needs grid.z
  if (p->stages.size() > 1) {
    const size_t per = (p->stages.size() == 3 ? 2 : 1) * (size_t)p->n;  // intermediates
    if (hipMalloc(&p->d_scratch, (size_t)p->chunk_h * per * sizeof(double2)) != hipSuccess) {
      mfft_destroy_plan(p);
      return nullptr;
    }
  }
  return p;
}

bool valid_problem(int64_t n, int sign, int64_t howmany) {
  return n >= 2 && n <= (int64_t(1) << 24) && smooth13(n) && (sign == -1 || sign == 1) &&
         howmany >= 0;
}

int64_t count_partials(const mfft_plan *p) {
  if (p->howmany == 0 || p->stages.empty()) return 0;
  const StageArgs &last = p->stages.back();
  const int64_t nx = (last.nsub + last.G - 1) / last.G;
  if (p->stages.size() == 1) return nx;
  return nx * p->howmany;  // one row of nx workgroups per transform, over all chunks
}

// The intermediate of a two-stage plan: the plan's own buffer (chunks of chunk_h transforms),
// or a caller's region of `elems` complex values, in equal chunks of as many whole
// transforms as fit (at most 65535, the grid's y extent).
struct Work {
  double2 *ptr;
  int64_t chunk;
};

int work_for(const mfft_plan *p, const mfft::Workspace *ws, Work &w) {
  if (!ws) {
    if (p->stages.size() > 1 && !p->d_scratch) return MFFT_ERR_WORKSPACE;
    w = {p->d_scratch, p->chunk_h};
    return MFFT_SUCCESS;
  }
  const int64_t hm = std::max<int64_t>(1, p->howmany);
  const int64_t per = p->stages.size() == 3 ? 2 * p->n : p->n;  // intermediates per transform
  const int64_t cmax = std::min<int64_t>({hm, ws->elems / per, 65535});
  if (cmax < 1 || !ws->ptr) return MFFT_ERR_WORKSPACE;
  const int64_t nchunk = (hm + cmax - 1) / cmax;
  w = {ws->ptr, (hm + nchunk - 1) / nchunk};
  return MFFT_SUCCESS;
}

// pre_s / pre_k: optional scalar factor of the input weights (StageArgs::pre_s, pre_k).
int run(const mfft_plan *p, const double2 *in, double2 *out, const double *pre,
        const double *post, double *sumsq, hipStream_t st, const mfft::Workspace *ws = nullptr,
        const double *pre_s = nullptr, double pre_k = 1.0) {
  if (p->howmany == 0) return MFFT_SUCCESS;
  if (!in || !out) return MFFT_ERR_INVALID;
  const bool fwd = p->sign < 0;
  if (p->stages.size() == 1) {
    StageArgs a = p->stages[0];
    a.pre = pre;
    a.pre_s = pre_s;
    a.pre_k = pre_k;
    a.post = post;
    a.sumsq = sumsq;
    // a call with input weights is the adjoint of the differentiable path, whose input (the
    // forward's output) was just written non-temporally: non-temporal loads and stores there
    // (MI355X, (512, 4096) forward + backward 35.3 us with the plain policy, 33.3 with this)
    if (a.reg == 1 && pre) a.cached_in = a.cached_out = 0;
    if (p->outer == 1)
      return mfft::launch_stage(fwd, a, 1, in, out, p->d_tw, st) == hipSuccess ? MFFT_SUCCESS
                                                                              : MFFT_ERR_HIP;
    a.in_h = p->outer_in;  // grid.y walks the outer batch
    a.out_h = p->outer_out;
    for (int64_t o0 = 0; o0 < p->outer; o0 += 65535) {
      const int64_t oc = std::min<int64_t>(65535, p->outer - o0);
      if (mfft::launch_stage(fwd, a, oc, in + o0 * p->outer_in, out + o0 * p->outer_out,
                             p->d_tw, st) != hipSuccess)
        return MFFT_ERR_HIP;
    }
    return MFFT_SUCCESS;
  }
  Work w{};
  if (const int rc = work_for(p, ws, w)) return rc;
  if (p->stages.size() == 3) {  // x -> T -> U -> y, chunks of w.chunk transforms
    const int64_t N = p->n, P3 = p->stages[2].L;
    const int64_t nxc = (p->stages[2].nsub + p->stages[2].G - 1) / p->stages[2].G;
    double2 *T = w.ptr, *U = w.ptr + w.chunk * N;
    for (int64_t h0 = 0; h0 < p->howmany; h0 += w.chunk) {
      const int64_t hc = std::min(w.chunk, p->howmany - h0);
      StageArgs A = p->stages[0], B = p->stages[1], C = p->stages[2];
      A.pre = pre;
      A.pre_s = pre_s;
      A.pre_k = pre_k;
      C.post = post;
      C.sumsq = sumsq ? sumsq + h0 * nxc : nullptr;
      if (mfft::launch_stage(fwd, A, hc, in + h0 * p->idist, T, p->d_tw, st) != hipSuccess ||
          mfft::launch_stage(fwd, B, P3, T, U, p->d_tw, st, hc) != hipSuccess ||
          mfft::launch_stage(fwd, C, hc, U, out + h0 * p->odist, p->d_tw, st) != hipSuccess)
        return MFFT_ERR_HIP;
    }
    return MFFT_SUCCESS;
  }
  if (p->columns) {
    if (pre || post || sumsq) return MFFT_ERR_UNSUPPORTED;
    const int64_t P = p->stages[0].L, Q = p->stages[1].L;
    for (int64_t h0 = 0; h0 < p->howmany; h0 += w.chunk) {
      const int64_t hc = std::min(w.chunk, p->howmany - h0);
      StageArgs A = p->stages[0], B = p->stages[1];
      A.nsub = B.nsub = hc;
      A.out_e = hc;  // intermediate [q P + e][transform]
      A.out_h = P * hc;
      B.in_e = P * hc;
      B.in_h = hc;
      A.tile = mfft::tile_lookup(A);  // the chunk fixes the tile count
      B.tile = mfft::tile_lookup(B);
      if (mfft::launch_stage(fwd, A, Q, in + h0 * p->idist, w.ptr, p->d_tw, st) != hipSuccess)
        return MFFT_ERR_HIP;
      if (mfft::launch_stage(fwd, B, P, w.ptr, out + h0 * p->odist, p->d_tw, st) != hipSuccess)
        return MFFT_ERR_HIP;
    }
    return MFFT_SUCCESS;
  }
  const int64_t nxb = (p->stages[1].nsub + p->stages[1].G - 1) / p->stages[1].G;
  for (int64_t h0 = 0; h0 < p->howmany; h0 += w.chunk) {
    const int64_t hc = std::min(w.chunk, p->howmany - h0);
    StageArgs A = p->stages[0], B = p->stages[1];
    A.pre = pre;
    A.pre_s = pre_s;
    A.pre_k = pre_k;
    B.post = post;
    B.sumsq = sumsq ? sumsq + h0 * nxb : nullptr;
    if (mfft::launch_stage(fwd, A, hc, in + h0 * p->idist, w.ptr, p->d_tw, st) != hipSuccess)
      return MFFT_ERR_HIP;
    if (mfft::launch_stage(fwd, B, hc, w.ptr, out + h0 * p->odist, p->d_tw, st) != hipSuccess)
      return MFFT_ERR_HIP;
  }
  return MFFT_SUCCESS;
}

// Whether the real-data element maps fuse into a contiguous plan's kernels: tile kernels that
// carry them (a staged single stage, or both stages of a two-stage plan), or register kernels
// in every stage, which take them into their first loads and last stores (the input maps at
// 8 elements per thread, or real input into a 16-per-thread schedule that opens with radix 16).
bool fusable(const mfft_plan *q, int io_in) {
  if (q->istride != 1 || q->ostride != 1 || q->columns || q->stages.empty()) return false;
  // a single stage in a staged tile kernel with maps takes them into its tile copies, two
  // stages in tile kernels with maps into their first loads and last stores
  if (q->stages.size() == 1 && q->stages[0].tile > 0 && q->stages[0].stage_io == 2) return true;
  if (q->stages.size() == 2 && q->stages[0].tile > 0 && q->stages[1].tile > 0 &&
      q->stages[0].stage_io == 3 && q->stages[1].stage_io == 3)
    return true;
  for (const auto &a : q->stages)
    if (a.reg == 0) return false;
  const StageArgs &a0 = q->stages[0];
  return a0.ept == 8 || (io_in == 1 && a0.reg == 1 && a0.pass[0].R == 16);
}

// A contiguous complex plan with the real-data maps fused: input real doubles (io_in 1) or
// Hermitian halves (2), output real parts (io_out 1) or the first n/2 + 1 values (2). The
// arithmetic is the complex plan's on the embedded / extended batch, so the bits equal the
// unfused composition's.
int run_io(const mfft_plan *q, const void *in, void *out, int io_in, int io_out, hipStream_t st,
           const mfft::Workspace *ws) {
  const int64_t n = q->n, nh = n / 2 + 1;
  const int64_t din = io_in == 2 ? nh : n, dout = io_out == 2 ? nh : n;  // per transform
  const size_t bin = io_in == 1 ? sizeof(double) : sizeof(double2);
  const size_t bout = io_out == 1 ? sizeof(double) : sizeof(double2);
  const bool fwd = q->sign < 0;
  if (q->stages.size() == 1) {
    StageArgs a = q->stages[0];
    a.io_in = io_in;
    a.io_out = io_out;
    a.io_n = n;
    a.in_g = din;
    a.out_g = dout;
    return mfft::launch_stage(fwd, a, 1, (const double2 *)in, (double2 *)out, q->d_tw, st) ==
                   hipSuccess
               ? MFFT_SUCCESS
               : MFFT_ERR_HIP;
  }
  Work w{};
  if (const int rc = work_for(q, ws, w)) return rc;
  for (int64_t h0 = 0; h0 < q->howmany; h0 += w.chunk) {
    const int64_t hc = std::min(w.chunk, q->howmany - h0);
    StageArgs A = q->stages[0], B = q->stages[1];
    A.io_in = io_in;
    A.io_n = n;
    A.in_h = din;
    B.io_out = io_out;
    B.io_n = n;
    B.out_h = dout;
    // r2c: the half rows (n/2 + 1 values per transform, so most rows start off a 128-byte
    // line) written with ordinary stores (MI355X, 2^24-point batches: 2^15 / 2^16 / 2^17 /
    // 2^18 points 168 / 169 / 192 / 218 -> 161 / 163 / 176 / 208 us); c2r's real rows keep
    // non-temporal stores (ordinary ones measured equal or up to 2% slower)
    if (io_out == 2 && B.reg != 0) B.cached_out = 1;  // (the tile stages measured equal)
    const auto *src = (const double2 *)((const char *)in + (size_t)h0 * din * bin);
    auto *dst = (double2 *)((char *)out + (size_t)h0 * dout * bout);
    if (mfft::launch_stage(fwd, A, hc, src, w.ptr, q->d_tw, st) != hipSuccess ||
        mfft::launch_stage(fwd, B, hc, w.ptr, dst, q->d_tw, st) != hipSuccess)
      return MFFT_ERR_HIP;
  }
  return MFFT_SUCCESS;
}

// Composite plans: the element maps of mfft_real.hip around the complex parts, or fused
// into the register kernels' loads and stores where every stage has one.
int run_composite(const mfft_plan *p, const void *in, void *out, hipStream_t st,
                  const mfft::Workspace *ws = nullptr) {
  if (!in || !out) return MFFT_ERR_INVALID;
  const int64_t hm = p->howmany;
  if (hm == 0) return MFFT_SUCCESS;
  if (p->kind == 1 && p->fused_io) return run_io(p->part[0], in, out, 1, 2, st, ws);
  if (p->kind == 2 && p->fused_io) return run_io(p->part[0], in, out, 2, 1, st, ws);
  // the unfused real-data compositions stage the whole batch in the plan's own work buffer
  if ((p->kind == 1 || p->kind == 2) && ws) return MFFT_ERR_UNSUPPORTED;
  if ((p->kind == 1 || p->kind == 2) && !p->d_work) return MFFT_ERR_WORKSPACE;
  switch (p->kind) {
    case 1: {  // r2c: embed, forward DFT in place, first n/2 + 1 outputs
      if (mfft::launch_embed_real((const double *)in, p->d_work, p->n, hm, st) != hipSuccess)
        return MFFT_ERR_HIP;
      const int rc = run(p->part[0], p->d_work, p->d_work, nullptr, nullptr, nullptr, st);
      if (rc != MFFT_SUCCESS) return rc;
      return mfft::launch_extract_half(p->d_work, (double2 *)out, p->n, hm, st) == hipSuccess
                 ? MFFT_SUCCESS
                 : MFFT_ERR_HIP;
    }
    case 2: {  // c2r: Hermitian extension, backward DFT in place, real parts
      if (mfft::launch_herm_ext((const double2 *)in, p->d_work, p->n, hm, st) != hipSuccess)
        return MFFT_ERR_HIP;
      // a single register kernel (16 elements per thread, where the Hermitian input map would
      // spill) still takes the real-part map into its stores
      const mfft_plan *q = p->part[0];
      if (q->stages.size() == 1 && q->stages[0].reg == 1)
        return run_io(q, p->d_work, out, 0, 1, st, nullptr);
      const int rc = run(p->part[0], p->d_work, p->d_work, nullptr, nullptr, nullptr, st);
      if (rc != MFFT_SUCCESS) return rc;
      return mfft::launch_real_part(p->d_work, (double *)out, p->n, hm, st) == hipSuccess
                 ? MFFT_SUCCESS
                 : MFFT_ERR_HIP;
    }
    case 3: {  // 2-D: axis 0 (columns, stride n1) of every array, then axis 1 (rows)
      if (p->split2d)  // arrays of 16384 points: both axes in one workgroup's registers
        return mfft::launch_split2d(p->sign < 0, p->split2d, p->part[0]->stages[0],
                                    p->part[1]->stages[0], (const double2 *)in, (double2 *)out,
                                    p->part[0]->d_tw, p->part[1]->d_tw, hm, st) == hipSuccess
                   ? MFFT_SUCCESS
                   : MFFT_ERR_HIP;
      if (p->fused2d)  // small arrays: both axes in one pass through LDS
        return mfft::launch_tile2d(p->sign < 0, p->fused2d, p->part[0]->stages[0],
                                   p->part[1]->stages[0], (const double2 *)in, (double2 *)out,
                                   p->part[0]->d_tw, p->part[1]->d_tw, hm, st) == hipSuccess
                   ? MFFT_SUCCESS
                   : MFFT_ERR_HIP;
      const int64_t sz = p->n0 * p->n1;
      // one call over every array when the axis-0 part carries the outer batch
      const int64_t calls = p->part[0]->outer > 1 ? 1 : hm;
      for (int64_t h = 0; h < calls; ++h) {
        const int rc = run(p->part[0], (const double2 *)in + h * sz, (double2 *)out + h * sz,
                           nullptr, nullptr, nullptr, st, ws);
        if (rc != MFFT_SUCCESS) return rc;
      }
      return run(p->part[1], (const double2 *)out, (double2 *)out, nullptr, nullptr, nullptr, st,
                 ws);
    }
    default:
      return MFFT_ERR_INVALID;
  }
}

// Bytes of the input and output arrays of a composite plan (host-pointer execution).
void composite_bytes(const mfft_plan *p, size_t &in_b, size_t &out_b) {
  const size_t hm = (size_t)p->howmany, n = (size_t)p->n, nh = (size_t)(p->n / 2 + 1);
  if (p->kind == 1) {
    in_b = hm * n * sizeof(double);
    out_b = hm * nh * sizeof(double2);
  } else if (p->kind == 2) {
    in_b = hm * nh * sizeof(double2);
    out_b = hm * n * sizeof(double);
  } else {
    in_b = out_b = hm * n * sizeof(double2);
  }
}

mfft_plan *composite_1d(int kind, int64_t n, int64_t howmany, unsigned flags) {
  if (howmany < 0) return nullptr;
  mfft_plan *inner = mfft_plan_dft_1d(n, kind == 1 ? MFFT_FORWARD : MFFT_BACKWARD,
                                      std::max<int64_t>(howmany, 1), flags | kFuseRadices);
  if (!inner) return nullptr;
  auto *p = new mfft_plan;
  p->kind = kind;
  p->n = n;
  p->sign = kind == 1 ? -1 : 1;
  p->howmany = howmany;
  p->part[0] = inner;
  p->radices = inner->radices;
  p->stage_passes = inner->stage_passes;
  p->fused_io = fusable(inner, kind == 1 ? 1 : 2);
  if (howmany > 0 && !p->fused_io) {
    if (hipMalloc(&p->d_work, (size_t)howmany * n * sizeof(double2)) != hipSuccess) {
      mfft_destroy_plan(p);
      return nullptr;
    }
  }
  return p;
}

// ------------------------------------------------------------------ cost model
//
// The reference's operation counts (LeanFFTW/Gen/OpCount.lean:23-79): a butterfly of radix R
// costs its scalar loads (2R input parts and, when twiddled, 2(R-1) twiddle parts), its
// arithmetic by kind (from the generated butterflies: additions, multiplications, fused
// multiply-adds) and its 2R scalar stores, each with a weight; pass t of a plan runs N / R
// butterflies, twiddled from the second pass on; a plan costs the sum of its passes.

struct CostWeights {
  int64_t add, mul, fma, load, store;
};

int64_t bfly_cost(const CostWeights &W, bool fwd, int R, bool tw) {
  const int *k = mfft::kBflyOpKinds[fwd ? 1 : 0][tw ? 1 : 0][R];
  const int64_t loads = 2 * R + (tw ? 2 * (R - 1) : 0);
  return loads * W.load + k[0] * W.add + k[1] * W.mul + k[2] * W.fma + 2 * (int64_t)R * W.store;
}

int64_t stockham_cost(const CostWeights &W, bool fwd, int64_t N, const std::vector<int> &rs) {
  int64_t c = 0;
  for (size_t t = 0; t < rs.size(); ++t) c += (N / rs[t]) * bfly_cost(W, fwd, rs[t], t >= 1);
  return c;
}

// The GPU's cost of a schedule in nanoseconds per transform on MI355X, for ordering and
// pruning the race (the measurement decides): every kernel stage moves the transform through
// HBM once (32 N bytes at the 6.5 TB/s a streaming kernel reaches), every further pass is an
// LDS round trip (32 N bytes at about 60 TB/s over 256 CUs), arithmetic runs at the FP64
// vector peak (78.6 TFLOP/s, an fma two flops) and twiddles come through the caches
// (16 bytes each at about 20 TB/s).
double gpu_cost(bool fwd, int64_t N, const std::vector<int> &rs, int split) {
  const double bytes = 32.0 * (double)N;
  const int stages = (split > 0 && split < (int)rs.size()) ? 2 : 1;
  double tw = 0.0;
  for (size_t t = 1; t < rs.size(); ++t) tw += (double)(N / rs[t]) * (rs[t] - 1);
  const double flops = (double)stockham_cost(CostWeights{1, 1, 2, 0, 0}, fwd, N, rs);
  return stages * bytes / 6.5e3 + ((double)rs.size() - stages) * bytes / 60e3 + flops / 78.6e3 +
         tw * 16.0 / 20e3;
}

// ------------------------------------------------------------------ autotuner (MFFT_MEASURE)
//
// The reference's measuring planner, its race (LeanFFTW/Asm/Planner.lean:176-205, :224-286):
// enumerate candidates, keep those within 1.5x of the least modelled cost among candidates of
// the same kind (here: one or two kernel stages), time them in order of cost, drop a
// candidate whose first measurement is 1.3x slower than the best so far, then re-time the
// four fastest interleaved and keep the fastest. A candidate is a radix sequence, a stage
// split, and for each stage the kernel: one of the precompiled tile kernels that fit its
// shape (tile G sub-problems on NT threads), or none (the register or stage kernel). The
// planner's own schedule (ESTIMATE: the hand-measured radices and tile tables) is always a
// candidate, so MEASURE keeps it unless something timed faster. Candidates run on scratch
// buffers of up to 256 MB of the plan's layout (the real batch up to that size); the winner is cached per (device, size,
// direction, batch class, layout) in a JSON-lines wisdom file.

constexpr double kPruneCut = 1.5;
constexpr double kRaceCut = 1.3;

struct Cand {
  std::vector<int> rs;
  int split = 0;
  // per stage: -1 the planner's own choice, 0 no tile kernel, else the tile kernel of G
  // sub-problems on NT threads (tg, tn)
  int tg[2] = {-1, -1}, tn[2] = {-1, -1};
  double cost = 0.0, us = 1e300;
  bool seed = false;
};

bool same_cand(const Cand &a, const Cand &b) {
  return a.rs == b.rs && a.split == b.split && a.tg[0] == b.tg[0] && a.tg[1] == b.tg[1] &&
         a.tn[0] == b.tn[0] && a.tn[1] == b.tn[1];
}

std::vector<std::vector<int>> seq_candidates(int64_t L) {
  std::vector<std::vector<int>> out;
  auto add = [&](std::vector<int> v) {
    if (v.empty()) return;
    if (std::find(out.begin(), out.end(), v) == out.end()) out.push_back(v);
    std::reverse(v.begin(), v.end());
    if (std::find(out.begin(), out.end(), v) == out.end()) out.push_back(v);
  };
  add(choose_radices(L));
  for (int cap : {4, 8, 16}) {
    std::vector<int> v;
    double c;
    if (best_radices(L, any_radix, v, c, cap)) add(v);
    if (best_radices(L, mid_radix_ok, v, c, cap)) add(v);
    if (best_radices(L, small_radix_ok, v, c, cap)) add(v);
  }
  return out;
}

std::vector<Cand> schedule_candidates(int64_t n, bool allow_big, bool contig) {
  std::vector<Cand> out;
  if (single_stage(n, allow_big, contig))
    for (auto &v : seq_candidates(n)) out.push_back({v, 0});
  if (n > mfft::LDS_ELEMS) {
    std::vector<std::pair<double, int64_t>> splits;
    for (int64_t P = 2; P <= mfft::LDS_ELEMS && P < n; ++P)
      if (n % P == 0 && n / P <= mfft::LDS_ELEMS)
        splits.push_back({std::fabs(std::log((double)P) - std::log((double)(n / P))), P});
    std::sort(splits.begin(), splits.end());
    for (size_t i = 0; i < splits.size() && i < 3; ++i) {
      const int64_t P = splits[i].second;
      auto A = seq_candidates(P), B = seq_candidates(n / P);
      for (size_t a = 0; a < A.size() && a < 3; ++a)
        for (size_t b = 0; b < B.size() && b < 3; ++b) {
          Cand c{A[a], (int)A[a].size()};
          c.rs.insert(c.rs.end(), B[b].begin(), B[b].end());
          out.push_back(c);
        }
    }
  }
  return out;
}

std::string device_key() {
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop{};
  (void)hipGetDeviceProperties(&prop, dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  std::string arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch = arch.substr(0, colon);
  return arch + ":" + std::to_string(ncu);
}

std::string wisdom_file() {
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  if (!g_wisdom_path.empty()) return g_wisdom_path;
  const char *e = std::getenv("MFFT_WISDOM");
  return e ? std::string(e) : std::string();
}

std::string wisdom_key(const mfft_plan *p) {
  int hb = 0;
  while ((int64_t(1) << (hb + 1)) <= p->howmany) ++hb;
  const bool contig = p->istride == 1 && p->ostride == 1 && p->idist == p->n && p->odist == p->n;
  return device_key() + "|" + std::to_string(p->n) + "|" + std::to_string(p->sign) + "|" +
         std::to_string(hb) + "|" + (contig ? "c" : "s") + "|" + (p->allow_big ? "b" : "n");
}

// the integers of a JSON array following `name` in a wisdom line
std::vector<int> json_ints(const std::string &line, const char *name) {
  std::vector<int> v;
  const std::string tag = std::string("\"") + name + "\": [";
  const size_t a = line.find(tag);
  if (a == std::string::npos) return v;
  const size_t b = line.find(']', a);
  if (b == std::string::npos) return v;
  const std::string body = line.substr(a + tag.size(), b - a - tag.size());
  size_t pos = 0;
  while (pos < body.size()) {
    const size_t nx = body.find(',', pos);
    const std::string tok = body.substr(pos, nx == std::string::npos ? std::string::npos : nx - pos);
    if (tok.find_first_not_of(" ") != std::string::npos) v.push_back(std::atoi(tok.c_str()));
    if (nx == std::string::npos) break;
    pos = nx + 1;
  }
  return v;
}

// wisdom lines: {"key": "<key>", "rs": [r0, ...], "split": s, "tg": [G0, G1], "tn": [NT0, NT1],
// "us": t} (tg / tn optional: the planner's own kernels)
bool wisdom_lookup(const std::string &key, int64_t key_n, Cand &c) {
  const std::string path = wisdom_file();
  if (path.empty()) return false;
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  char line[2048];
  bool found = false;
  const std::string want = "\"key\": \"" + key + "\"";
  while (std::fgets(line, sizeof line, f)) {
    const std::string sl(line);
    if (sl.find(want) == std::string::npos) continue;
    const size_t sp = sl.find("\"split\": ");
    if (sp == std::string::npos) continue;
    Cand t;
    t.rs = json_ints(sl, "rs");
    t.split = std::atoi(sl.c_str() + sp + 9);
    const std::vector<int> tg = json_ints(sl, "tg"), tn = json_ints(sl, "tn");
    if (tg.size() == 2 && tn.size() == 2)
      for (int s = 0; s < 2; ++s) {
        t.tg[s] = tg[s];
        t.tn[s] = tn[s];
      }
    int64_t prod = 1;
    bool ok = !t.rs.empty() && t.split >= 0 && t.split < (int)t.rs.size();
    for (int r : t.rs) {
      ok = ok && radix_ok(r);
      prod *= ok ? r : 1;
    }
    if (!ok || prod != key_n) continue;
    c = t;  // the last valid matching line wins
    found = true;
  }
  std::fclose(f);
  return found;
}

void wisdom_store(const std::string &key, const Cand &c, double us) {
  const std::string path = wisdom_file();
  if (path.empty()) return;
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  FILE *f = std::fopen(path.c_str(), "a");
  if (!f) return;
  std::string rs;
  for (size_t i = 0; i < c.rs.size(); ++i) rs += (i ? ", " : "") + std::to_string(c.rs[i]);
  std::fprintf(f,
               "{\"key\": \"%s\", \"rs\": [%s], \"split\": %d, \"tg\": [%d, %d], \"tn\": [%d, %d], "
               "\"us\": %.3f}\n",
               key.c_str(), rs.c_str(), c.split, c.tg[0], c.tg[1], c.tn[0], c.tn[1], us);
  std::fclose(f);
}

mfft_plan *clone_shape(const mfft_plan *p, int64_t howmany) {
  auto *q = new mfft_plan;
  q->n = p->n;
  q->sign = p->sign;
  q->howmany = howmany;
  q->istride = p->istride;
  q->idist = p->idist;
  q->ostride = p->ostride;
  q->odist = p->odist;
  q->allow_big = p->allow_big;
  q->fuse = p->fuse;
  return q;
}

// Give each stage the kernel the candidate names (a tile of that G and NT if one fits, none),
// where it is not the planner's own choice. Column plans pick their tiles per chunk.
void apply_tiles(mfft_plan *q, const Cand &c) {
  if (q->columns) return;
  for (size_t s = 0; s < q->stages.size() && s < 2; ++s) {
    StageArgs &st = q->stages[s];
    if (c.tg[s] < 0) continue;
    if (c.tg[s] == 0) {
      untiled(st);
      continue;
    }
    int tiles[16];
    const int nt = std::min(16, mfft::tile_candidates(st, tiles, 16));
    for (int i = 0; i < nt; ++i) {
      int G = 0, NT = 0;
      if (mfft::tile_shape(tiles[i], &G, &NT) && G == c.tg[s] && NT == c.tn[s]) {
        st.tile = tiles[i];
        break;
      }
    }
  }
}

mfft_plan *build_cand(const mfft_plan *shape, const Cand &c, int64_t howmany) {
  mfft_plan *q = clone_shape(shape, howmany);
  q->radices = c.rs;
  q = finish(q, c.split);
  if (q) apply_tiles(q, c);
  return q;
}

// The planner's own schedule (ESTIMATE) of a plan whose shape is set: radices and split(s).
bool estimate_schedule(const mfft_plan *p, std::vector<int> &rs, int &split, int &split2);

// Expand a schedule into one candidate per combination of its stages' kernels: every tile
// kernel that fits the stage, and no tile; the planner's own combination is marked.
void expand_kernels(const mfft_plan *shape, const Cand &base, int64_t hm, std::vector<Cand> &out) {
  mfft_plan *q = build_cand(shape, base, hm);
  if (!q) return;
  std::vector<std::pair<int, int>> opts[2];  // (G, NT) per option; (0, 0) no tile
  int own[2][2] = {{-1, -1}, {-1, -1}};
  const size_t ns = q->columns ? 0 : std::min<size_t>(2, q->stages.size());
  for (size_t s = 0; s < ns; ++s) {
    const StageArgs &st = q->stages[s];
    opts[s].push_back({0, 0});
    int tiles[16];
    const int nt = std::min(16, mfft::tile_candidates(st, tiles, 16));
    for (int i = 0; i < nt; ++i) {
      int G = 0, NT = 0;
      if (mfft::tile_shape(tiles[i], &G, &NT)) opts[s].push_back({G, NT});
    }
    if (st.tile > 0) (void)mfft::tile_shape(st.tile, &own[s][0], &own[s][1]);
    else own[s][0] = own[s][1] = 0;
  }
  mfft_destroy_plan(q);
  if (ns == 0) {
    out.push_back(base);
    return;
  }
  const size_t n1 = ns > 1 ? opts[1].size() : 1;
  for (const auto &o0 : opts[0])
    for (size_t j = 0; j < n1; ++j) {
      Cand c = base;
      c.tg[0] = o0.first;
      c.tn[0] = o0.second;
      if (ns > 1) {
        c.tg[1] = opts[1][j].first;
        c.tn[1] = opts[1][j].second;
      }
      c.seed = base.seed && c.tg[0] == own[0][0] && c.tn[0] == own[0][1] &&
               (ns < 2 || (c.tg[1] == own[1][0] && c.tn[1] == own[1][1]));
      bool dup = false;
      for (const Cand &d : out) dup = dup || same_cand(d, c);
      if (!dup) out.push_back(c);
    }
}

// Time per run (us): the fastest of `reps` timings, each of `k` back-to-back runs (so that
// launch gaps average out, as a batch of calls runs); after one untimed run when `warm`.
// 1e300 if the plan cannot run.
double time_plan(const mfft_plan *q, const double2 *din, double2 *dout, hipStream_t st,
                 hipEvent_t e0, hipEvent_t e1, int reps, int k, bool warm) {
  if (warm && run(q, din, dout, nullptr, nullptr, nullptr, st) != MFFT_SUCCESS) return 1e300;
  double t = 1e300;
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < k; ++i)
      if (run(q, din, dout, nullptr, nullptr, nullptr, st) != MFFT_SUCCESS) return 1e300;
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) t = std::min(t, (double)ms * 1e3 / k);
  }
  return t;
}

// Back-to-back runs per timing: about 300 us of work, 1 to 32 runs.
int runs_per_timing(double us) {
  return us >= 300.0 ? 1 : (int)std::min(32.0, std::ceil(300.0 / std::max(us, 1.0)));
}

// Fill `n` complex values at `d` with uniform[-0.5, 0.5) data from the reference's input LCG
// (scripts/NativeBench.lean:31-37): a 1 MiB-element block made on the host, then doubled on the
// device. The race times candidates on such data, not on zeros: on MI355X a plan's time on
// zeros differs from its time on real data by 5-10%, and not by the same amount for every
// plan (32768 points: 8 x 8 x 4 | 8 x 8 x 2 is the fastest on random data and the slowest of
// four schedules on zeros).
bool fill_uniform(double2 *d, int64_t n) {
  const int64_t blk = std::min<int64_t>(n, int64_t(1) << 20);
  std::vector<double2> h((size_t)blk);
  uint64_t st = 0x9E3779B97F4A7C15ull;
  auto next = [&st] {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) / 9007199254740992.0 - 0.5;
  };
  for (auto &v : h) {
    v.x = next();
    v.y = next();
  }
  if (hipMemcpy(d, h.data(), (size_t)blk * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
    return false;
  for (int64_t done = blk; done < n;) {
    const int64_t c = std::min(done, n - done);
    if (hipMemcpy(d + done, d, (size_t)c * sizeof(double2), hipMemcpyDeviceToDevice) != hipSuccess)
      return false;
    done += c;
  }
  return true;
}

// Race the candidates on scratch buffers; returns false (and leaves `best` alone) when the
// layout's span is too large to tune on scratch, or nothing could be timed.
bool measure(const mfft_plan *shape, Cand &best) {
  const int64_t n = shape->n;
  // the plan's own batch up to 256 MB (2^24 elements): tile and XCD-remap choices depend on
  // how many sub-problems a launch has, so a smaller sample can prefer a plan that is slower
  // at the real batch (165375 points: 272 us ESTIMATE, 301 us for the choice timed at 64 MB)
  const int64_t hm = std::max<int64_t>(1, std::min<int64_t>(shape->howmany, (int64_t(1) << 24) / n));
  auto span = [&](int64_t st, int64_t d) { return std::llabs(st) * (n - 1) + std::llabs(d) * (hm - 1) + 1; };
  const int64_t si = span(shape->istride, shape->idist), so = span(shape->ostride, shape->odist);
  if (si > (int64_t(1) << 25) || so > (int64_t(1) << 25)) return false;
  if (shape->istride < 0 || shape->ostride < 0 || shape->idist < 0 || shape->odist < 0) return false;
  // candidates: the planner's own schedule first, then the cost-model schedules, each with
  // every combination of its stages' kernels
  std::vector<Cand> scheds;
  {
    Cand seed;
    int split2 = 0;
    const bool est = estimate_schedule(shape, seed.rs, seed.split, split2);
    if (split2 > 0) return false;  // three stages: the planner's own schedule, not raced
    if (est) {
      seed.seed = true;
      scheds.push_back(seed);
    }
    for (const Cand &c : schedule_candidates(n, shape->allow_big,
                                             shape->istride == 1 && shape->ostride == 1)) {
      if (!shape->allow_big && !all_mid(c.rs)) continue;  // ragged classes: the mid kernel
      bool dup = false;
      for (const Cand &d : scheds) dup = dup || (d.rs == c.rs && d.split == c.split);
      if (!dup) scheds.push_back(c);
    }
  }
  std::vector<Cand> cands;
  for (const Cand &c : scheds) expand_kernels(shape, c, hm, cands);
  if (cands.empty()) return false;
  // prune: within 1.5x of the least modelled cost of the same number of stages (the planner's
  // own schedule always stays); then race in order of cost
  double least[3] = {1e300, 1e300, 1e300};
  auto nstage = [](const Cand &c) { return c.split > 0 && c.split < (int)c.rs.size() ? 2 : 1; };
  for (Cand &c : cands) {
    c.cost = gpu_cost(shape->sign < 0, n, c.rs, c.split);
    least[nstage(c)] = std::min(least[nstage(c)], c.cost);
  }
  std::vector<Cand> order;
  for (const Cand &c : cands)
    if (c.seed || c.cost <= kPruneCut * least[nstage(c)]) order.push_back(c);
  std::stable_sort(order.begin(), order.end(),
                   [](const Cand &a, const Cand &b) { return a.cost < b.cost; });

  double2 *din = nullptr, *dout = nullptr;
  if (hipMalloc(&din, si * sizeof(double2)) != hipSuccess) return false;
  if (hipMalloc(&dout, so * sizeof(double2)) != hipSuccess) {
    (void)hipFree(din);
    return false;
  }
  if (!fill_uniform(din, si)) {
    (void)hipFree(din);
    (void)hipFree(dout);
    return false;
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipStreamCreate(&st);
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  double best_us = 1e300;
  for (Cand &c : order) {
    mfft_plan *q = build_cand(shape, c, hm);
    if (!q) continue;
    const double t0 = time_plan(q, din, dout, st, e0, e1, 1, 1, true);
    const int k = t0 < 1e299 ? runs_per_timing(t0) : 1;
    const double t1 = t0 < 1e299 ? time_plan(q, din, dout, st, e0, e1, 1, k, false) : t0;
    c.us = t1;
    if (t1 < 1e299 && t1 <= kRaceCut * best_us)
      c.us = std::min(t1, time_plan(q, din, dout, st, e0, e1, 2, k, false));
    best_us = std::min(best_us, c.us);
    mfft_destroy_plan(q);
  }
  // the four fastest again, interleaved, three rounds
  std::stable_sort(order.begin(), order.end(),
                   [](const Cand &a, const Cand &b) { return a.us < b.us; });
  std::vector<std::pair<Cand, mfft_plan *>> top;
  for (const Cand &c : order) {
    if ((int)top.size() == 4 || c.us >= 1e299) break;
    if (mfft_plan *q = build_cand(shape, c, hm)) top.push_back({c, q});
  }
  std::vector<double> ts(top.size(), 1e300);
  for (int round = 0; round < 3; ++round)
    for (size_t i = 0; i < top.size(); ++i)
      ts[i] = std::min(ts[i], time_plan(top[i].second, din, dout, st, e0, e1, 1,
                                        runs_per_timing(top[i].first.us), round == 0));
  size_t bi = 0;
  for (size_t i = 1; i < top.size(); ++i)
    if (ts[i] < ts[bi]) bi = i;
  const bool ok = !top.empty() && ts[bi] < 1e299;
  if (ok) best = top[bi].first;
  const double us = ok ? ts[bi] : 1e300;
  for (auto &tq : top) mfft_destroy_plan(tq.second);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(din);
  (void)hipFree(dout);
  if (!ok) return false;
  wisdom_store(wisdom_key(shape), best, us);
  return true;
}

}  // namespace

namespace {

// The planner's own schedule (ESTIMATE) of a plan whose shape is set: the hand-measured
// radices of the register-resident sizes, else the cost-model radices of one stage, of the
// most balanced split into two, or (no two-way split fits the LDS) of three.
bool estimate_schedule(const mfft_plan *p, std::vector<int> &rs, int &split, int &split2) {
  const int64_t n = p->n;
  const bool contig = p->istride == 1 && p->ostride == 1;
  split = split2 = 0;
  // columns too long for a tile of 4 (64-byte rows) run as two stages of wider rows rather
  // than one stage of narrower rows (MI355X, 2-D columns: 4096 points two stages 274-304 us
  // against 868). 2048-point columns take one register stage of 4-wide rows, whose tiles the
  // XCD remap keeps beside the neighbours they share lines with, and which carries the batch
  // of arrays in one launch (2-D 2048 x 2048 x 4: 295 us as two column stages of 32 x 64 per
  // array, 252-253 in one; before the remap one stage measured 316-322 against 285-287)
  const bool col_two = columns_layout(p) && 4 * n > mfft::BIG_ELEMS;
  if (single_stage(n, p->allow_big, contig) && !col_two) {
    rs = p->allow_big && contig && reg_size(n) ? reg_radices(n, p->fuse)
                                               : choose_radices(n, !p->allow_big);
    return true;
  }
  int64_t P;
  if (!choose_split(n, P)) {
    int64_t P1, P2;
    if (!p->allow_big || columns_layout(p) || !choose_split3(n, P1, P2)) return false;
    const std::vector<int> a = choose_radices(P1), b = choose_radices(P2),
                           c = choose_radices(n / P1 / P2);
    rs = a;
    rs.insert(rs.end(), b.begin(), b.end());
    rs.insert(rs.end(), c.begin(), c.end());
    split = (int)a.size();
    split2 = (int)(a.size() + b.size());
    return true;
  }
  std::vector<int> a = choose_radices(P, !p->allow_big), b = choose_radices(n / P, !p->allow_big);
  rs = a;
  rs.insert(rs.end(), b.begin(), b.end());
  split = (int)a.size();
  return true;
}

}  // namespace

mfft_plan *plan_many_internal(int64_t n, int sign, int64_t howmany, int64_t istride,
                              int64_t idist, int64_t ostride, int64_t odist, unsigned flags,
                              bool allow_big) {
  if (!valid_problem(n, sign, howmany) || istride == 0 || ostride == 0) return nullptr;
  auto *p = new mfft_plan;
  p->n = n;
  p->sign = sign;
  p->howmany = howmany;
  p->istride = istride;
  p->idist = idist;
  p->ostride = ostride;
  p->odist = odist;
  p->allow_big = allow_big;
  p->fuse = (flags & kFuseRadices) != 0;
  if (flags & MFFT_MEASURE) {
    // A stale or hand-edited wisdom line that does not describe a valid schedule for this
    // size is ignored (and a tile it names that does not fit, the planner's own kernel is
    // kept); the default schedule is used when nothing can be measured.
    Cand c;
    if (wisdom_lookup(wisdom_key(p), n, c) || measure(p, c)) {
      if (mfft_plan *q = build_cand(p, c, howmany)) {
        delete p;
        return q;
      }
    }
  }
  int split = 0, split2 = 0;
  if (!estimate_schedule(p, p->radices, split, split2)) {
    delete p;
    return nullptr;
  }
  return finish(p, split, split2);
}

int mfft::execute_in(const mfft_plan *p, const void *in, void *out, const mfft::Workspace *ws,
                     hipStream_t st) {
  if (!p) return MFFT_ERR_INVALID;
  if (p->kind != 0) {
    if (p->kind != 3 && in == out) return MFFT_ERR_UNSUPPORTED;  // different element types
    return run_composite(p, in, out, st, ws);
  }
  return run(p, (const double2 *)in, (double2 *)out, nullptr, nullptr, nullptr, st, ws);
}

int mfft::plan_device(const mfft_plan *p, int *dev) {
  if (!p || !dev) return MFFT_ERR_INVALID;
  *dev = p->kind == 0 ? p->device : p->part[0]->device;  // composites: their parts' device
  return MFFT_SUCCESS;
}

int64_t mfft::workspace_elems(const mfft_plan *p) {
  if (!p) return 0;
  if (p->kind == 0) return p->stages.size() == 3 ? 2 * p->n : p->stages.size() > 1 ? p->n : 0;
  int64_t e = 0;
  for (const mfft_plan *q : p->part) e = std::max(e, workspace_elems(q));
  return e;
}

int mfft::export_class(const mfft_plan *p, mfft::RClass *c, std::vector<double2> *tw) {
  if (!p || p->howmany != 1 || p->stages.empty() || p->stages.size() > 2) return 1;
  for (const auto &st : p->stages)
    if (st.ept != 8) return 1;  // the ragged kernel runs 8 elements per thread only
  if (!all_mid(p->radices)) return 1;  // and holds the mid radix set
  const int64_t base = (int64_t)tw->size();
  *c = mfft::RClass{};
  c->N = (int)p->n;
  c->nstage = (int)p->stages.size();
  for (int s = 0; s < c->nstage; ++s) {
    const StageArgs &a = p->stages[s];
    c->L[s] = a.L;
    c->npass[s] = a.npass;
    c->G[s] = a.G;
    for (int q = 0; q < a.npass; ++q) {
      c->pass[s][q] = a.pass[q];
      c->pass[s][q].twoff = a.pass[q].twoff >= 0 ? a.pass[q].twoff + base : -1;
    }
  }
  if (c->nstage == 1) {  // a ragged tile holds LDS_ELEMS / N transforms of the class
    c->G[0] = std::max(1, mfft::LDS_ELEMS / c->L[0]);
    for (int q = 0; q < c->npass[0]; ++q) c->pass[0][q].m_G = mfft::magic32(c->G[0]);
  }
  c->P = c->nstage == 2 ? p->stages[0].L : 0;
  c->Q = c->nstage == 2 ? p->stages[1].L : 0;
  tw->insert(tw->end(), p->h_tw.begin(), p->h_tw.end());
  return 0;
}

extern "C" {

mfft_plan *mfft_plan_with_radices(int64_t n, int sign, int64_t howmany, const int *radices,
                                  int nradix, int stage_split) {
  if (!valid_problem(n, sign, howmany) || !radices || nradix < 1) return nullptr;
  int64_t prod = 1;
  for (int i = 0; i < nradix; ++i) {
    if (!radix_ok(radices[i])) return nullptr;
    prod *= radices[i];
  }
  if (prod != n) return nullptr;
  auto *p = new mfft_plan;
  p->n = n;
  p->sign = sign;
  p->howmany = howmany;
  p->istride = p->ostride = 1;
  p->idist = p->odist = n;
  p->radices.assign(radices, radices + nradix);
  return finish(p, stage_split);
}

mfft_plan *mfft_plan_with_stages(int64_t n, int sign, int64_t howmany, const int *radices,
                                 int nradix, const int *passes, int nstage) {
  if (!passes || nstage < 1 || nstage > 3) return nullptr;
  int sum = 0;
  for (int i = 0; i < nstage; ++i) {
    if (passes[i] < 1) return nullptr;
    sum += passes[i];
  }
  if (sum != nradix) return nullptr;
  if (nstage < 3) return mfft_plan_with_radices(n, sign, howmany, radices, nradix,
                                                nstage == 2 ? passes[0] : 0);
  if (!valid_problem(n, sign, howmany) || !radices) return nullptr;
  int64_t prod = 1;
  for (int i = 0; i < nradix; ++i) {
    if (!radix_ok(radices[i])) return nullptr;
    prod *= radices[i];
  }
  if (prod != n) return nullptr;
  auto *p = new mfft_plan;
  p->n = n;
  p->sign = sign;
  p->howmany = howmany;
  p->istride = p->ostride = 1;
  p->idist = p->odist = n;
  p->radices.assign(radices, radices + nradix);
  return finish(p, passes[0], passes[0] + passes[1]);
}

mfft_plan *mfft_plan_many(int64_t n, int sign, int64_t howmany, int64_t istride, int64_t idist,
                          int64_t ostride, int64_t odist, unsigned flags) {
  return plan_many_internal(n, sign, howmany, istride, idist, ostride, odist, flags, true);
}

int mfft_plan_schedule(int64_t n, int *radices, int max, int *split) {
  if (!valid_problem(n, -1, 1)) return -1;
  mfft_plan shape;
  shape.n = n;
  shape.idist = shape.odist = n;
  std::vector<int> rs;
  int sp = 0, sp2 = 0;
  if (!estimate_schedule(&shape, rs, sp, sp2)) return -1;
  for (int i = 0; i < (int)rs.size() && i < max; ++i) radices[i] = rs[i];
  if (split) *split = sp;
  return (int)rs.size();
}

mfft_plan *mfft_plan_create(int64_t n, int sign, int64_t howmany, int64_t istride,
                            int64_t idist, int64_t ostride, int64_t odist, unsigned flags) {
  return mfft_plan_many(n, sign, howmany, istride, idist, ostride, odist, flags);
}

void mfft_plan_destroy(mfft_plan *p) { mfft_destroy_plan(p); }

mfft_plan *mfft_plan_dft_1d(int64_t n, int sign, int64_t howmany, unsigned flags) {
  return mfft_plan_many(n, sign, howmany, 1, n, 1, n, flags);
}

mfft_plan *mfft_plan_spec(const double *spec, size_t len, unsigned flags) {
  // [kind, sign, rank, vrank, inRe, inIm, outRe, outIm, (n,is,os) x rank, (n,is,os) x vrank]
  if (!spec || len < 8) return nullptr;
  auto gi = [&](size_t i) { return (int64_t)spec[i]; };
  const int64_t kind = gi(0), sign = gi(1), rank = gi(2), vrank = gi(3);
  if (vrank < 0 || vrank > 1 || rank < 1 || rank > 2) return nullptr;
  if (len < (size_t)(8 + 3 * (rank + vrank))) return nullptr;
  if (kind == 1 || rank == 2) {
    // contiguous layouts only (else NULL: the caller keeps its own path)
    const int64_t inRe = gi(4), inIm = gi(5), outRe = gi(6), outIm = gi(7);
    const int64_t hm = vrank ? gi(8 + 3 * rank) : 1;
    mfft_plan *p = nullptr;
    if (kind == 1 && rank == 1) {
      const int64_t n = gi(8), is = gi(9), os = gi(10), nh = n / 2 + 1;
      const bool fwd = sign < 0;
      const int64_t vis = vrank ? gi(12) : (fwd ? n : 2 * nh);
      const int64_t vos = vrank ? gi(13) : (fwd ? 2 * nh : n);
      const bool ok = fwd ? (is == 1 && os == 2 && vis == n && vos == 2 * nh && outIm == outRe + 1)
                          : (is == 2 && os == 1 && vis == 2 * nh && vos == n && inIm == inRe + 1);
      if (!ok) return nullptr;
      p = fwd ? mfft_plan_r2c(n, hm, flags) : mfft_plan_c2r(n, hm, flags);
    } else if (kind == 0 && rank == 2) {
      const int64_t n0 = gi(8), n1 = gi(11);
      const bool ok = inIm == inRe + 1 && outIm == outRe + 1 && gi(9) == 2 * n1 &&
                      gi(10) == 2 * n1 && gi(12) == 2 && gi(13) == 2 &&
                      (!vrank || (gi(15) == 2 * n0 * n1 && gi(16) == 2 * n0 * n1));
      if (!ok) return nullptr;
      p = mfft_plan_dft_2d(n0, n1, sign < 0 ? -1 : 1, hm, flags);
    }
    if (p) {
      p->spec_in_off = inRe;
      p->spec_out_off = outRe;
    }
    return p;
  }
  if (kind != 0) return nullptr;
  if (len < (size_t)(8 + 3 * (rank + vrank))) return nullptr;
  const int64_t inRe = gi(4), inIm = gi(5), outRe = gi(6), outIm = gi(7);
  if (inIm != inRe + 1 || outIm != outRe + 1) return nullptr;  // interleaved only
  const int64_t n = gi(8), is = gi(9), os = gi(10);
  int64_t hm = 1, vis = 0, vos = 0;
  if (vrank == 1) {
    hm = gi(11);
    vis = gi(12);
    vos = gi(13);
  }
  if ((is % 2) || (os % 2) || (vis % 2) || (vos % 2)) return nullptr;
  mfft_plan *p = mfft_plan_many(n, sign < 0 ? -1 : 1, hm, is / 2, vis / 2, os / 2, vos / 2, flags);
  if (p) {
    p->spec_in_off = inRe;
    p->spec_out_off = outRe;
  }
  return p;
}

int mfft_execute(const mfft_plan *p, const void *in, void *out, void *stream) {
  return mfft::execute_in(p, in, out, nullptr, (hipStream_t)stream);
}

mfft_plan *mfft_plan_r2c(int64_t n, int64_t howmany, unsigned flags) {
  return composite_1d(1, n, howmany, flags);
}

mfft_plan *mfft_plan_c2r(int64_t n, int64_t howmany, unsigned flags) {
  return composite_1d(2, n, howmany, flags);
}

mfft_plan *mfft_plan_dft_2d(int64_t n0, int64_t n1, int sign, int64_t howmany, unsigned flags) {
  if (n0 < 2 || n1 < 2 || howmany < 0 || (sign != -1 && sign != 1)) return nullptr;
  mfft_plan *col = mfft_plan_many(n0, sign, n1, n1, 1, n1, 1, flags);
  mfft_plan *row = col ? mfft_plan_many(n1, sign, n0 * std::max<int64_t>(howmany, 1), 1, n1, 1,
                                        n1, flags)
                       : nullptr;
  if (!col || !row) {
    mfft_destroy_plan(col);
    mfft_destroy_plan(row);
    return nullptr;
  }
  // the axis-0 part's output is the 2-D transform's intermediate, which the axis-1 part reads
  // back at once: ordinary stores (MI355X, 2^24 elements per size: 256 x 256 186 -> 170 us,
  // 512 x 512 196 -> 185, 1024 x 1024 232 -> 223, 4096 x 4096 292 -> 258; 2048 x 2048, whose
  // columns take two stages, 287 -> 297)
  for (auto &st : col->stages) st.cached_out = 1;
  if (col->stages.size() == 1 && howmany > 1 && col->stages[0].reg != 1) {
    col->outer = howmany;  // one launch over the columns of every array
    col->outer_in = col->outer_out = n0 * n1;
  }
  auto *p = new mfft_plan;
  p->kind = 3;
  p->n = n0 * n1;
  p->n0 = n0;
  p->n1 = n1;
  p->sign = sign;
  p->howmany = howmany;
  p->part[0] = col;
  p->part[1] = row;
  p->radices = row->radices;
  p->stage_passes = row->stage_passes;
  if (col->stages.size() == 1 && row->stages.size() == 1 && n0 * n1 <= 4096)
    p->fused2d = mfft::tile2d_lookup(n0, n1, col->stages[0], row->stages[0]);
  if (col->stages.size() == 1 && row->stages.size() == 1)
    p->split2d = mfft::split2d_lookup(n0, n1, col->stages[0], row->stages[0]);
  return p;
}

int mfft_plan_kind(const mfft_plan *p) { return p ? p->kind : -1; }

int mfft_execute_weighted(const mfft_plan *p, const void *in, void *out, const double *pre,
                          const double *post, void *stream) {
  if (!p) return MFFT_ERR_INVALID;
  if (p->kind != 0) return MFFT_ERR_UNSUPPORTED;
  if ((pre || post) && (p->istride != 1 || p->ostride != 1)) return MFFT_ERR_UNSUPPORTED;
  return run(p, (const double2 *)in, (double2 *)out, pre, post, nullptr, (hipStream_t)stream);
}

int mfft_execute_weighted_scaled(const mfft_plan *p, const void *in, void *out,
                                 const double *pre, const double *scale, double factor,
                                 const double *post, void *stream) {
  if (!p || !pre || !scale) return MFFT_ERR_INVALID;
  if (p->kind != 0) return MFFT_ERR_UNSUPPORTED;
  if (p->istride != 1 || p->ostride != 1) return MFFT_ERR_UNSUPPORTED;
  return run(p, (const double2 *)in, (double2 *)out, pre, post, nullptr, (hipStream_t)stream,
             nullptr, scale, factor);
}

int64_t mfft_plan_partials(const mfft_plan *p) {
  return p && p->kind == 0 ? count_partials(p) : -1;
}

int mfft_execute_weighted_sumsq(const mfft_plan *p, const void *in, void *out, const double *pre,
                                const double *post, double *partials, void *stream) {
  if (!p || !partials) return MFFT_ERR_INVALID;
  if (p->kind != 0) return MFFT_ERR_UNSUPPORTED;
  if ((pre || post) && (p->istride != 1 || p->ostride != 1)) return MFFT_ERR_UNSUPPORTED;
  return run(p, (const double2 *)in, (double2 *)out, pre, post, partials, (hipStream_t)stream);
}

int mfft_execute_spec(const mfft_plan *p, const double *in, double *out, void *stream) {
  if (!p || !in || !out) return MFFT_ERR_INVALID;
  return mfft_execute(p, in + p->spec_in_off, out + p->spec_out_off, stream);
}

int mfft_execute_host(const mfft_plan *p, const double *in, double *out) {
  if (!p || !in || !out) return MFFT_ERR_INVALID;
  if (p->kind != 0) {  // contiguous composite layouts
    size_t ib = 0, ob = 0;
    composite_bytes(p, ib, ob);
    void *di = nullptr, *dout = nullptr;
    if (hipMalloc(&di, ib ? ib : 16) != hipSuccess) return MFFT_ERR_ALLOC;
    if (hipMalloc(&dout, ob ? ob : 16) != hipSuccess) {
      (void)hipFree(di);
      return MFFT_ERR_ALLOC;
    }
    int rc = hipMemcpy(di, in, ib, hipMemcpyHostToDevice) == hipSuccess ? MFFT_SUCCESS
                                                                          : MFFT_ERR_HIP;
    if (rc == MFFT_SUCCESS) rc = run_composite(p, di, dout, nullptr);
    if (rc == MFFT_SUCCESS && (hipDeviceSynchronize() != hipSuccess ||
                               hipMemcpy(out, dout, ob, hipMemcpyDeviceToHost) != hipSuccess))
      rc = MFFT_ERR_HIP;
    (void)hipFree(di);
    (void)hipFree(dout);
    return rc;
  }
  // span of elements touched on each side
  auto span = [&](int64_t stride, int64_t dist) {
    const int64_t a = std::llabs(stride) * (p->n - 1) + std::llabs(dist) * (p->howmany - 1) + 1;
    return a;
  };
  if (p->istride < 0 || p->ostride < 0 || p->idist < 0 || p->odist < 0)
    return MFFT_ERR_UNSUPPORTED;
  const int64_t ni = span(p->istride, p->idist), no = span(p->ostride, p->odist);
  double2 *di = nullptr, *dout = nullptr;
  if (hipMalloc(&di, ni * sizeof(double2)) != hipSuccess) return MFFT_ERR_ALLOC;
  if (hipMalloc(&dout, no * sizeof(double2)) != hipSuccess) {
    (void)hipFree(di);
    return MFFT_ERR_ALLOC;
  }
  int rc = MFFT_SUCCESS;
  if (hipMemcpy(di, in, ni * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dout, out, no * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
    rc = MFFT_ERR_HIP;
  if (rc == MFFT_SUCCESS) rc = run(p, di, dout, nullptr, nullptr, nullptr, nullptr);
  if (rc == MFFT_SUCCESS && (hipDeviceSynchronize() != hipSuccess ||
                             hipMemcpy(out, dout, no * sizeof(double2), hipMemcpyDeviceToHost) !=
                                 hipSuccess))
    rc = MFFT_ERR_HIP;
  (void)hipFree(di);
  (void)hipFree(dout);
  return rc;
}

void mfft_destroy_plan(mfft_plan *p) {
  if (!p) return;
  if (p->d_tw) (void)hipFree(p->d_tw);
  if (p->d_scratch) (void)hipFree(p->d_scratch);
  if (p->d_work) (void)hipFree(p->d_work);
  mfft_destroy_plan(p->part[0]);
  mfft_destroy_plan(p->part[1]);
  delete p;
}

int mfft_plan_radices(const mfft_plan *p, int *radices, int max) {
  if (!p) return -1;
  const int n = (int)p->radices.size();
  for (int i = 0; i < n && i < max; ++i) radices[i] = p->radices[i];
  return n;
}

int mfft_plan_stages(const mfft_plan *p, int *passes, int max) {
  if (!p) return -1;
  const int n = (int)p->stage_passes.size();
  for (int i = 0; i < n && i < max; ++i) passes[i] = p->stage_passes[i];
  return n;
}

void mfft_cos_sin_2pi(uint64_t k, uint64_t n, double *c, double *s) {
  double cc, ss;
  mfft::cos_sin_2pi(k, n, cc, ss);
  if (c) *c = cc;
  if (s) *s = ss;
}

double mfft_tan_2pi(uint64_t num, uint64_t den) { return mfft::tan_2pi(num, den); }

int mfft_pass_twiddles(int fwd, int64_t N, int R, int64_t ns, double *out) {
  if (!out || N < 1 || R < 1 || ns < 1 || N % (ns * R)) return MFFT_ERR_INVALID;
  std::vector<double> tab;
  mfft::cos_sin_2pi_table((uint64_t)N, tab);
  const int64_t step = N / (ns * R);
  size_t o = 0;
  for (int r = 1; r < R; ++r) {
    for (int64_t k = 0; k < ns; ++k) out[o++] = tab[2 * (r * k * step)];
    for (int64_t k = 0; k < ns; ++k) {
      const double s = tab[2 * (r * k * step) + 1];
      out[o++] = fwd ? (s == 0.0 ? 0.0 : -s) : s;
    }
  }
  return MFFT_SUCCESS;
}

int64_t mfft_stockham_cost(int fwd, int64_t n, const int *radices, int nradix,
                           const int64_t *weights) {
  if (!radices || !weights || nradix < 1 || n < 1) return -1;
  std::vector<int> rs(radices, radices + nradix);
  int64_t prod = 1;
  for (int r : rs) {
    if (!radix_ok(r)) return -1;
    prod *= r;
  }
  if (prod != n) return -1;
  const CostWeights W{weights[0], weights[1], weights[2], weights[3], weights[4]};
  return stockham_cost(W, fwd != 0, n, rs);
}

double mfft_schedule_cost(int fwd, int64_t n, const int *radices, int nradix, int split) {
  if (!radices || nradix < 1 || n < 1) return -1.0;
  std::vector<int> rs(radices, radices + nradix);
  int64_t prod = 1;
  for (int r : rs) {
    if (!radix_ok(r)) return -1.0;
    prod *= r;
  }
  if (prod != n || split < 0 || split >= nradix) return -1.0;
  return gpu_cost(fwd != 0, n, rs, split);
}

int mfft_set_threads(size_t n) {
  g_threads = n;
  return MFFT_SUCCESS;
}

int mfft_set_wisdom_path(const char *path) {
  std::lock_guard<std::mutex> lk(g_wisdom_mu);
  g_wisdom_path = path ? path : "";
  return MFFT_SUCCESS;
}

const char *mfft_error_string(int code) {
  switch (code) {
    case MFFT_SUCCESS: return "success";
    case MFFT_ERR_INVALID: return "invalid argument";
    case MFFT_ERR_UNSUPPORTED: return "unsupported problem";
    case MFFT_ERR_HIP: return "HIP runtime error";
    case MFFT_ERR_ALLOC: return "allocation failed";
    case MFFT_ERR_WORKSPACE:
      return "no intermediate for a two-stage plan (the workspace cannot hold one transform)";
    default: return "unknown error";
  }
}

const char *mfft_version(void) { return "mfft 0.1 (gfx950)"; }

}  // extern "C"
