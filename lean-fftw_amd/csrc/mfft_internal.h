// Internal declarations shared by the host library and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

struct mfft_plan;

namespace mfft {

// ---------------------------------------------------------------- twiddles (mfft_trig.cpp)
void cos_sin_2pi(uint64_t k, uint64_t n, double &c, double &s);
double tan_2pi(uint64_t num, uint64_t den);
void cos_sin_2pi_table(uint64_t n, std::vector<double> &cs);  // 2n doubles: c0,s0,c1,s1,...

// ---------------------------------------------------------------- schedule
constexpr int MAX_PASSES = 12;      // passes per kernel stage
constexpr int WG_THREADS = 256;     // threads per workgroup of the stage kernels
constexpr int LDS_ELEMS = 2048;     // complex doubles per LDS buffer (two buffers: 64 KiB)

// One kernel stage: a group of consecutive Stockham passes run on independent
// sub-problems ("sub-transforms") held in LDS.
//
// Sub-problem s (0 <= s < nsub) is (h, u) = (s / U, s % U): transform h of the batch,
// part u of it. Element i (< L) of the sub-problem is read from
//     in [h * idist + u * in_u + i * in_i]
// and, after the stage's passes (the Stockham passes of the full plan restricted to
// this sub-problem), element i is written to
//     out[h * odist + u * out_u + i * out_i].
// Pass p of the stage has radix R[p] and local ns nsl[p] (the product of the stage's
// earlier radices). Its twiddle for butterfly input r (1 <= r < R) and local index k
// (= j mod nsl) is tw[twoff[p] + (r-1) * twns[p] + kb + k * kmul] with kb = u * kb_u:
// twns[p] is the pass's global ns (the table row length) and kmul, kb map the local
// index to the global one. twoff[p] < 0 marks the untwiddled first pass of a plan.
struct StageArgs {
  int L;
  int npass;
  int R[MAX_PASSES];
  int nsl[MAX_PASSES];
  int64_t twoff[MAX_PASSES];
  int64_t twns[MAX_PASSES];
  int64_t kmul;
  int64_t kb_u;
  int64_t U;
  int64_t in_u, in_i, out_u, out_i;
  int64_t idist, odist;
  int64_t nsub;
  int G;        // sub-problems per workgroup
  int gfast_in;  // 1: consecutive threads walk the sub-problem index on load
  int gfast_out; // same for the store
  // optional real weights (nullptr = none): input element i of every sub-problem is
  // multiplied by pre[pre_off(u, i)] before the first pass, output element i by
  // post[...] after the last; offsets use the same (u, i) -> element map as the data
  // within one transform: e = u * in_u' + i * in_i' with the strides below.
  const double *pre;
  const double *post;
  int64_t pre_u, pre_i, post_u, post_i;
};

// A size class of a ragged batch: the stages of one plan (at most two), with pass
// parameters as in StageArgs and twiddle offsets into the ragged batch's shared table.
struct RClass {
  int N, nstage, P, Q;
  int L[2], npass[2], G[2];
  int R[2][MAX_PASSES];
  int nsl[2][MAX_PASSES];
  int64_t twoff[2][MAX_PASSES];
  int64_t twns[2][MAX_PASSES];
};

// Describe a single-transform plan as a ragged size class, appending its twiddle table
// to `tw` (offsets in the class are relative to the start of `tw`). Returns 0 on success.
int export_class(const struct ::mfft_plan *p, RClass *c, std::vector<double2> *tw);

// Launch one stage; returns a hipError_t.
hipError_t launch_stage(bool fwd, const StageArgs &a, const double2 *in, double2 *out,
                        const double2 *tw, hipStream_t stream);

}  // namespace mfft
