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
// Radix sets of the kernel variants. A kernel's register allocation is set by the largest
// butterflies it contains, so stages run the smallest set holding their radices: SMALL
// (2, 3, 4, 5, 6, 8: about 88 VGPRs at 8 elements per thread), MID (2 ... 13, enough for
// every 13-smooth size: about 110, four waves per SIMD), ALL (up to 16, with 14, 15 and 16:
// about 134, three waves); POW2 is the set of the large-tile variant with 16 elements per
// thread.
enum RadixSet { RSET_ALL = 0, RSET_SMALL = 1, RSET_POW2 = 2, RSET_MID = 3 };
__host__ __device__ constexpr bool small_radix(int r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 6 || r == 8;
}
__host__ __device__ constexpr bool mid_radix(int r) { return r >= 2 && r <= 13; }
__host__ __device__ constexpr bool in_set(int set, int r) {
  return set == RSET_ALL || (set == RSET_SMALL ? small_radix(r)
                             : set == RSET_MID ? mid_radix(r)
                                               : (r == 2 || r == 4 || r == 8 || r == 16));
}

constexpr int MAX_PASSES = 12;      // passes per kernel stage
constexpr int WG_THREADS = 256;     // threads per workgroup of the stage kernels
constexpr int LDS_ELEMS = 2048;     // complex doubles per LDS buffer, before padding
// upper bound of a workgroup's dynamic LDS: LDS_ELEMS elements padded by one element per 16
// plus one per sub-problem (see mfft_pass.h); the launches request only what their stage
// needs
constexpr int kLdsBytes = (LDS_ELEMS + LDS_ELEMS / 16 + LDS_ELEMS / 2) * 16;
// the large-tile variants: up to BIG_ELEMS points per workgroup, 16 per thread
constexpr int BIG_ELEMS = 8192;
constexpr int kLdsBytesBig = (BIG_ELEMS + BIG_ELEMS / 16 + 1) * 16;
// dynamic LDS limit of the stage kernels: a tile of up to 8192 points of sub-problems padded
// one element per 16 plus one each (at most 139,520 bytes), below the 160 KiB of a gfx950 CU
// with room for the kernels' static LDS
constexpr int kLdsBytesMax = 144 * 1024;

// One Stockham pass inside a kernel stage. Local index k of butterfly j is j mod nsl (nsl:
// product of the stage's earlier radices); its twiddle for input r (1 <= r < R) is
// tw[twoff + (r-1) * twns + kb + k * kmul], twns being the pass's global ns (the table row
// length) and kb the sub-problem's base index (see StageArgs). twoff < 0: untwiddled (the
// first pass of a plan). m_* are ceil(2^32 / d) for d = L/R, nsl, G (0 when d = 1).
struct PassArgs {
  int R, nsl;
  uint32_t m_LR, m_nsl, m_G;
  int64_t twoff, twns, kmul;
};

// One kernel stage: consecutive Stockham passes applied to independent sub-problems of
// length L, G sub-problems per workgroup, grid (ceil(nsub / G), transforms).
//
// Element e of sub-problem g (global index within its transform) of transform h is read
// from in[h * in_h + g * in_g + e * in_e] and, after the passes, written to
// out[h * out_h + g * out_g + e * out_e]. Twiddle base kb = g * kb_u + h * kb_h. Optional real
// weights: pre[g * pre_g + e * pre_e] multiplies the input element, post[...] the output.
struct StageArgs {
  int L, npass, G;
  int threads;  // workgroup size
  int ept;      // elements per thread: 8 (G*L <= LDS_ELEMS, 256 threads) or 16 (large tiles,
                // G*L <= BIG_ELEMS, up to 512 threads)
  int gfast_in, gfast_out;  // 1: consecutive threads walk g (g adjacent in memory)
  int stage_io;             // 1: coalesced tile copies HBM -> LDS and back (contiguous small
                            // L); 2: the same in the complex part of a real-data plan, whose
                            // tile kernel takes the real-data maps into its copies
  int reg;                  // 1: register-resident kernel (mfft_reg.hip), one transform of
                            // L = 32..16384 points, L / 8 or L / 16 threads per transform;
                            // 2: register-resident stage kernel, G sub-problems of L / 8
                            // threads each
  uint32_t m_L;             // ceil(2^32 / L), for the tile copies
  PassArgs pass[MAX_PASSES];
  int64_t nsub;
  int64_t in_g, in_e, in_h, out_g, out_e, out_h;
  int64_t kb_u;
  int64_t kb_h;  // twiddle base of workgroup row h: kb = g * kb_u + h * kb_h
  // grid.z (the stage kernel only; the middle stage of a three-stage plan, where grid.y walks a
  // second sub-problem index): element offsets in_z, out_z per z
  int64_t in_z, out_z;
  // real-data element maps fused into the first loads / last stores (register kernels only):
  // io_in 0 complex, 1 real doubles, 2 Hermitian half; io_out 0 complex, 1 real parts,
  // 2 the first io_n / 2 + 1 elements; io_n the full transform length
  int io_in, io_out;
  int64_t io_n;
  // 1: ordinary (cached) loads / stores on that side instead of non-temporal ones (set by
  // the planner where they measured faster: a two-stage intermediate, register-kernel outputs)
  int cached_in, cached_out;
  const double *pre;
  const double *post;
  int64_t pre_g, pre_e, post_g, post_e;
  // optional scalar factor of the input weights: with pre_s set, the input weight is
  // (pre_k * pre_s[0]) * pre[...] (pre_s a device value, pre_k a host constant)
  const double *pre_s;
  double pre_k;
  // optional: each workgroup writes the sum of |out|^2 of its stores to
  // sumsq[blockIdx.y * gridDim.x + blockIdx.x] (last stage of a plan only)
  double *sumsq;
  // 1 + index of a fixed-schedule tile kernel (mfft_tile.hip) for this stage, 0 for none; it
  // runs the stage when no weights, partials or element maps are asked for
  int tile;
  // 1: grid.x is padded to a multiple of 8 and workgroups are remapped so that each of the 8
  // XCDs (which take workgroups round-robin in launch order) runs one contiguous eighth of the
  // sub-problems of every row: stages whose twiddles depend on the sub-problem then read one
  // eighth of the table per XCD's L2
  int xcd;
};

// A size class of a ragged batch: the stages of one plan (at most two), with pass
// parameters as in StageArgs and twiddle offsets into the ragged batch's shared table.
struct RClass {
  int N, nstage, P, Q;
  int L[2], npass[2], G[2];
  PassArgs pass[2][MAX_PASSES];
};

// Describe a single-transform plan as a ragged size class, appending its twiddle table
// to `tw` (offsets in the class are relative to the start of `tw`). Returns 0 on success.
int export_class(const struct ::mfft_plan *p, RClass *c, std::vector<double2> *tw);

// The register-resident kernel: power-of-two n in [32, 8192] with radices in {2, 4, 8}, or
// n = 16384 as radix 4 only.
bool reg_supported(int64_t n, const int *radices, int nradix);
// Register-resident stage kernels (mfft_reg.hip): the sub-problems per workgroup of the
// stage schedule for length L and these radices (the preferred one if its G <= gmax, else the
// largest G <= gmax; 0 if there is none); and the launch of a stage whose G is one of them
// (sum-of-squares stages run the stage kernel).
int reg_stage_g(int L, const int *radices, int nradix, int gmax);
hipError_t launch_reg_stage(bool fwd, const StageArgs &a, int64_t ntrans, const double2 *in,
                            double2 *out, const double2 *tw, hipStream_t stream);
int reg_elems(int64_t n);  // elements per thread of the register-resident kernel at size n
// Split register kernels (mfft_split.hip): one transform per workgroup as column and row
// transforms inside the workgroup, for the radix sequences in its table (non-null if there is
// one); launch_split returns false when the stage is not one it runs (element maps, strides,
// several transforms per workgroup), else launches it and sets *err.
const void *split_schedule(int64_t n, const int *radices, int nradix);
bool launch_split(bool fwd, const StageArgs &a, const double2 *in, double2 *out,
                  const double2 *tw, hipStream_t stream, hipError_t *err);
hipError_t launch_reg(bool fwd, const StageArgs &a, const double2 *in, double2 *out,
                      const double2 *tw, hipStream_t stream);

// Element maps of the composite transforms (mfft_real.hip): embedding of a real batch
// (howmany x n doubles) as complex, the first n/2 + 1 of each transform, the Hermitian
// extension of half-complex transforms, the real parts.
hipError_t launch_embed_real(const double *x, double2 *work, int64_t n, int64_t howmany,
                             hipStream_t st);
hipError_t launch_extract_half(const double2 *work, double2 *out, int64_t n, int64_t howmany,
                               hipStream_t st);
hipError_t launch_herm_ext(const double2 *in, double2 *work, int64_t n, int64_t howmany,
                           hipStream_t st);
hipError_t launch_real_part(const double2 *work, double *out, int64_t n, int64_t howmany,
                            hipStream_t st);

// Fixed-schedule tile kernels (mfft_tile.hip): the kernel for a stage whose shape, radices,
// lane orders and tile copies match one (1 + its index, 0 if none), and its launch.
int tile_lookup(const StageArgs &a);
bool tile_fits(const StageArgs &a, int tile);
int tile_candidates(const StageArgs &a, int *tiles, int max);
bool tile_shape(int tile, int *G, int *NT);
int xcd_remap(const StageArgs &a, int G);
hipError_t launch_tile(bool fwd, const StageArgs &a, int64_t ntrans, const double2 *in,
                       double2 *out, const double2 *tw, hipStream_t stream);

// Single-pass 2-D transforms of small arrays (mfft_tile2d.hip): the fused kernel for an
// n0 x n1 plan whose axis-0 (col) and axis-1 (row) parts are single-stage with these stages'
// radices (1 + its index, 0 if none), and its launch over `howmany` contiguous arrays.
int tile2d_lookup(int64_t n0, int64_t n1, const StageArgs &col, const StageArgs &row);
hipError_t launch_tile2d(bool fwd, int idx, const StageArgs &col, const StageArgs &row,
                         const double2 *in, double2 *out, const double2 *tw0,
                         const double2 *tw1, int64_t howmany, hipStream_t stream);

// Single-pass 2-D transforms of arrays that fill one workgroup's registers (mfft_split.hip:
// 128 x 128): the kernel for an n0 x n1 plan whose parts are single stages of its radices
// (nullptr if none), and its launch over `arrays` contiguous arrays.
const void *split2d_lookup(int64_t n0, int64_t n1, const StageArgs &col, const StageArgs &row);
hipError_t launch_split2d(bool fwd, const void *sched, const StageArgs &col, const StageArgs &row,
                          const double2 *in, double2 *out, const double2 *tw0,
                          const double2 *tw1, int64_t arrays, hipStream_t stream);

// A caller-provided intermediate for two-stage plans: `elems` complex values at `ptr`.
struct Workspace {
  double2 *ptr;
  int64_t elems;
};
// Execute a plan (any kind) with its own intermediate (ws == nullptr) or with `ws`, in chunks
// of whole transforms; MFFT_ERR_WORKSPACE when neither is available or `ws` cannot hold one
// transform. Returns an MFFT_* code.
int execute_in(const ::mfft_plan *p, const void *in, void *out, const Workspace *ws,
               hipStream_t st);
// Complex values of intermediate one transform of the plan needs (0: single stage).
int64_t workspace_elems(const ::mfft_plan *p);
// The device a plan was made on.
int plan_device(const ::mfft_plan *p, int *dev);

// Launch one stage over `ntrans` rows of workgroups (grid.y) and `nz` layers (grid.z, the stage
// kernel only); returns a hipError_t.
hipError_t launch_stage(bool fwd, const StageArgs &a, int64_t ntrans, const double2 *in,
                        double2 *out, const double2 *tw, hipStream_t stream, int64_t nz = 1);

inline uint32_t magic32(int64_t d) {
  return d <= 1 ? 0u : (uint32_t)(((uint64_t(1) << 32) + (uint64_t)d - 1) / (uint64_t)d);
}

}  // namespace mfft

// Plan without the C ABI's defaults: `allow_big` = false keeps every stage in the
// 256-thread layout (ragged size classes).
mfft_plan *plan_many_internal(int64_t n, int sign, int64_t howmany, int64_t istride,
                              int64_t idist, int64_t ostride, int64_t odist, unsigned flags,
                              bool allow_big);
