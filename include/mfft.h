/* mfft: MI355X (gfx950) double-precision complex FFT -- the C ABI.
 *
 * This is the drop-in boundary for the reference's fast complex path. Every entry point
 * below names the reference interface it replaces (paths relative to the reference
 * repository). The reference binds its engine from C through these:
 *
 *   lean_object *lfa_plan_measure(size_t N, uint8_t fwd, size_t in, size_t out);
 *        LeanFFTW/Asm/Planner.lean:291-302, declared bench/backends/bench-leanfftw.c:225
 *   void lfa_call(const void *record, const double *in, double *out);
 *        LeanFFTW/Asm/Kernels.lean:1203-1204, declared bench/backends/bench-leanfftw.c:226
 *   lean_object *lfa_set_threads(size_t n);          LeanFFTW/Asm/Planner.lean:305-306
 *   lfftw_generic_plan(FloatArray spec) / lfftw_generic_exec(plan, in, out)
 *        LeanFFTW/FFI.lean:50-58 (spec layout :13-21)
 *
 * Conventions (docs/CONTRACT.md:9-17 of the reference): forward is sign -1,
 * Y[k] = sum_j x[j] exp(-2 pi i j k / N); backward is sign +1; nothing is normalised.
 * Data is interleaved complex binary64 (re, im). Sizes: 2 <= N <= 2^24 with prime factors
 * <= 13, as the reference's fast engine (bench-leanfftw.c:253-268): one kernel stage up to a
 * few thousand points, two up to 2048 x 2048 where N splits so, three otherwise (a 2^24-point
 * transform is 256 MiB). Column layouts (|dist| < |stride|) take one or two stages only.
 *
 * Differences from the reference, by design:
 *   - Device pointers: the execute calls take device (HBM) pointers and a HIP stream and
 *     enqueue asynchronously. mfft_execute_host is the host-pointer convenience.
 *   - Planning never reads or writes user arrays (lfa_plan_measure overwrites them).
 *   - Errors are return codes (lfa_call has none).
 */
#ifndef MFFT_H
#define MFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFFT_FORWARD (-1)
#define MFFT_BACKWARD (+1)

/* planner flags */
#define MFFT_ESTIMATE 0u  /* heuristic radix/tile choice, no timing */
#define MFFT_MEASURE 1u   /* autotune by timing candidate schedules on scratch buffers */

/* return codes */
#define MFFT_SUCCESS 0
#define MFFT_ERR_INVALID 1     /* bad argument (null plan, misaligned or null pointer) */
#define MFFT_ERR_UNSUPPORTED 2 /* size/layout outside the engine's domain */
#define MFFT_ERR_HIP 3         /* a HIP runtime call failed */
#define MFFT_ERR_ALLOC 4       /* device or host allocation failed */
#define MFFT_ERR_WORKSPACE 5   /* no intermediate available for a two-stage plan */

typedef struct mfft_plan mfft_plan;
typedef struct mfft_ragged mfft_ragged;

/* Plan `howmany` transforms of size n with arbitrary element strides and batch
 * distances (in complex elements), on the current HIP device. Returns NULL if the
 * problem is unsupported (replaces the empty plan array of lfa_plan_measure,
 * Planner.lean:297-302, and the `none` of lfftw_generic_plan, FFI.lean:50-52). */
mfft_plan *mfft_plan_many(int64_t n, int sign, int64_t howmany, int64_t istride,
                          int64_t idist, int64_t ostride, int64_t odist, unsigned flags);

/* The same entry point under the name of the reference-side plan record constructor a
 * caller may already bind (lfa_plan_measure's replacement, Planner.lean:291-302). Planning
 * never reads or writes caller arrays (MEASURE times candidate schedules on buffers of its
 * own), so there is no in-place-measure flag. */
mfft_plan *mfft_plan_create(int64_t n, int sign, int64_t howmany, int64_t istride,
                            int64_t idist, int64_t ostride, int64_t odist, unsigned flags);

/* Contiguous batch: istride = ostride = 1, idist = odist = n.
 * Replaces lfa_plan_measure(N, fwd, in, out) (with howmany = 1). */
mfft_plan *mfft_plan_dft_1d(int64_t n, int sign, int64_t howmany, unsigned flags);

/* Real-input forward transform: howmany contiguous real arrays of n doubles -> howmany
 * contiguous arrays of n/2 + 1 interleaved complex values, F_k = sum_j x_j w^(jk), k <= n/2.
 * Computed as the reference's generic engine does (LineOp.r2c, Problem.lean:128-129,
 * :144-146): the complex forward DFT of (x_j, 0), first n/2 + 1 outputs. in != out.
 * Replaces the rdft2 forward rank-1 case of planProblemCore (Problem.lean:227-234). */
mfft_plan *mfft_plan_r2c(int64_t n, int64_t howmany, unsigned flags);

/* Real-output backward transform (FFTW's c2r, unnormalised): howmany arrays of n/2 + 1
 * complex values -> howmany real arrays of n doubles, the real part of the backward DFT of
 * the Hermitian extension (hermExt, Problem.lean:138-140; LineOp.c2r :147-149). in != out.
 * Replaces the rdft2 backward rank-1 case of planProblemCore (Problem.lean:235-238). */
mfft_plan *mfft_plan_c2r(int64_t n, int64_t howmany, unsigned flags);

/* 2-D complex DFT of howmany contiguous row-major n0 x n1 arrays: the line DFT along
 * axis 0 (length n0, stride n1), then along axis 1 (length n1, contiguous), the order of the
 * reference's per-axis ops (planProblemCore .dft, Problem.lean:222-224; alongAxis
 * :165-182). in == out allowed. Replaces the dft rank-2 case of lfftw_generic_plan. */
mfft_plan *mfft_plan_dft_2d(int64_t n0, int64_t n1, int sign, int64_t howmany, unsigned flags);

/* 0 = complex 1-D, 1 = r2c, 2 = c2r, 3 = 2-D complex; -1 for NULL. The composite kinds
 * (1..3) reject the weighted and sum-of-squares entry points with MFFT_ERR_UNSUPPORTED. */
int mfft_plan_kind(const mfft_plan *p);

/* Plan from the reference's generic problem spec (FFI.lean:13-21):
 * [kind, sign, rank, vrank, inRe, inIm, outRe, outIm, (n,is,os) x rank, (n,is,os) x vrank]
 * in units of doubles. Supported, vrank 0 or 1: kind 0 (complex DFT) rank 1, interleaved
 * (inIm = inRe + 1, outIm = outRe + 1, even strides); kind 0 rank 2 and kind 1 (rdft2)
 * rank 1 in contiguous interleaved layouts (mfft_plan_dft_2d, mfft_plan_r2c,
 * mfft_plan_c2r). Offsets are applied by the caller through the pointers given to
 * mfft_execute_spec. Returns NULL otherwise.
 * Replaces lfftw_generic_plan (FFI.lean:50-52). */
mfft_plan *mfft_plan_spec(const double *spec, size_t len, unsigned flags);

/* Execute on device pointers (in == out allowed), enqueued on `stream`
 * (a hipStream_t; NULL = the default stream). Replaces lfa_call (Kernels.lean:1203-1204).
 *
 * Concurrency: a plan owns its device intermediate (two-stage plans) and work buffer
 * (unfused real-data plans), so executions of ONE plan must not overlap: calls on one
 * stream are ordered, but one plan must not run on two streams at once (make a plan per
 * stream, as the Python binding's caches do). Distinct plans run concurrently; planning is
 * thread-safe. */
int mfft_execute(const mfft_plan *p, const void *in, void *out, void *stream);

/* As mfft_execute, but the intermediate of two-stage plans is the library's static device
 * scratch -- a file-scope array of MFFT_STATIC_SCRATCH_BYTES (64 MiB) per GPU, held until the
 * process ends -- instead of the plan's own buffer; a batch runs in chunks of as many whole
 * transforms as fit. Allocates nothing (no hipMalloc / hipFree, no allocating library
 * routine). For callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation. MFFT_ERR_WORKSPACE if one transform's
 * intermediate exceeds the scratch (n above 4,194,304 points); MFFT_ERR_UNSUPPORTED for
 * real-data plans whose maps do not fuse into their kernels (they stage the whole batch in a
 * work buffer). Every call on a GPU shares that GPU's scratch: calls that overlap on one GPU
 * (two streams, two threads) race on it. Nothing in it needs to start at zero (stage A
 * writes every element stage B reads). No counterpart in the reference. */
#define MFFT_STATIC_SCRATCH_BYTES (64u << 20)
int mfft_execute_static(const mfft_plan *p, const void *in, void *out, void *stream);

/* Execute with host pointers: copy in, transform, copy out, synchronise. For the
 * libbench2-shaped backend (bench-leanfftw.c:281-291) and host-array callers. */
int mfft_execute_host(const mfft_plan *p, const double *in, double *out);

/* Execute a plan made by mfft_plan_spec: `in`/`out` point at the doubles the spec's
 * offsets are relative to (device pointers). Replaces lfftw_generic_exec (FFI.lean:56-58). */
int mfft_execute_spec(const mfft_plan *p, const double *in, double *out, void *stream);

void mfft_destroy_plan(mfft_plan *p);
/* Alias of mfft_destroy_plan (replaces the adapter's lean_dec(aplan),
 * bench-leanfftw.c:293-297). */
void mfft_plan_destroy(mfft_plan *p);

/* Plan description: the radix sequence R_0..R_{s-1} of the Stockham schedule (the
 * reference's `rs`, Stockham.lean:112) and the number of passes fused into each kernel
 * stage. Return the count written (or needed, if max is too small). */
int mfft_plan_radices(const mfft_plan *p, int *radices, int max);
int mfft_plan_stages(const mfft_plan *p, int *passes_per_stage, int max);

/* The schedule mfft_plan_dft_1d would choose for size n, without allocating anything (host
 * only): the radix sequence (return value: its length, -1 if unsupported) and the number of
 * passes of the first kernel stage (*split, 0 for one stage). */
int mfft_plan_schedule(int64_t n, int *radices, int max, int *split);

/* Force a radix sequence (tests and tuning): like mfft_plan_many with ESTIMATE, but the
 * Stockham schedule is `radices` (product n, each radix in {2..16} with prime factors
 * <= 13) and, if stage_split > 0, the first kernel stage runs passes [0, stage_split). */
mfft_plan *mfft_plan_with_radices(int64_t n, int sign, int64_t howmany, const int *radices,
                                  int nradix, int stage_split);

/* Force a radix sequence and its cut into kernel stages (tests and tuning): `passes` lists the
 * number of passes of each of the nstage (1..3) stages, summing to nradix; contiguous batches.
 * NULL if a stage's sub-problem length exceeds the LDS (2048 points) or the schedule is
 * invalid. */
mfft_plan *mfft_plan_with_stages(int64_t n, int sign, int64_t howmany, const int *radices,
                                 int nradix, const int *passes, int nstage);

/* Ragged batch: `count` transforms of sizes sizes[i], packed back to back (transform i
 * starts at element sum_{j<i} sizes[j]), one direction, executed as two persistent launches
 * from size-sorted work queues: every stage-A and single-stage tile, then every stage-B tile
 * (the launch boundary orders the stages; a HIP graph replays both as one unit), every
 * workgroup taking tiles from one shared queue per launch. in != out. The plan's queue heads
 * are device state: executes of one ragged plan must not overlap (one stream, or ordered).
 * No counterpart in the reference (it has no batched fast path, bench-leanfftw.c:257). */
mfft_ragged *mfft_plan_ragged(int64_t count, const int64_t *sizes, int sign, unsigned flags);
int mfft_execute_ragged(const mfft_ragged *p, const void *in, void *out, void *stream);
/* The schedule a ragged batch uses for its transforms of size n: the radix sequence (return
 * value: its length, -1 if n is not in the batch) and the number of passes of the first
 * kernel stage (*split, 0 for a single stage). */
int mfft_ragged_class_plan(const mfft_ragged *p, int64_t n, int *radices, int max, int *split);
/* Synchronise `stream`: MFFT_SUCCESS when the enqueued launches completed, else MFFT_ERR_HIP. */
int mfft_ragged_status(const mfft_ragged *p, void *stream);
/* Host only (tests): the work queues of a ragged batch of `count` transforms, built from
 * explicit class shapes -- per class, ascending distinct sizes, six values: n, stages (1 or
 * 2), P, Q (n = P Q), tile width of stage A (or of the single stage), of stage B -- with
 * `nqueue` queues per launch (1..64; the batch plans use one). Writes five values
 * per tile (class, stage, count, base, sub0) for at most max_tiles tiles, the first *n_first
 * of them the first launch's; qoff (2 (nqueue + 1) values, may be NULL) receives the queue
 * boundaries, queue q of the first launch at tiles [qoff[q], qoff[q + 1]) and of the second
 * at [qoff[nqueue + 1 + q], qoff[nqueue + 2 + q]); offs receives the transform offsets
 * single-stage tiles list (base indexes them). Returns the number of tiles, -1 for
 * inconsistent shapes or a size without a class. No counterpart in the reference. */
int64_t mfft_ragged_queue(int64_t count, const int64_t *sizes, int64_t nclass,
                          const int64_t *shapes, int64_t nqueue, int64_t *tiles,
                          int64_t max_tiles, int64_t *n_first, int64_t *qoff, int64_t *offs,
                          int64_t max_offs);
void mfft_destroy_ragged(mfft_ragged *p);

/* Fused weighted transform for the differentiable path: out = diag(post) F (diag(pre) in),
 * with optional real weights pre[n] and post[n] (NULL = 1) shared by every transform of
 * the batch. The adjoint of a forward plan is the backward plan with pre and post
 * exchanged. No counterpart in the reference. */
int mfft_execute_weighted(const mfft_plan *p, const void *in, void *out, const double *pre,
                          const double *post, void *stream);

/* As mfft_execute_weighted with input weights pre[n] (required) scaled by a device scalar:
 * element e of every transform is multiplied by (factor * scale[0]) * pre[e], the weight
 * formed in the kernel in that order, so the result is bit-identical to passing the vector
 * (factor * scale[0]) * pre computed elementwise on the device. For the adjoint of the
 * differentiable loss, whose weight 2 g w holds the upstream gradient g: a value on the
 * device that the host never reads, so no elementwise launch or host sync is needed to form
 * it. No counterpart in the reference. */
int mfft_execute_weighted_scaled(const mfft_plan *p, const void *in, void *out,
                                 const double *pre, const double *scale, double factor,
                                 const double *post, void *stream);

/* As mfft_execute_weighted, and every workgroup of the last kernel stage also writes the sum
 * of |out|^2 over the elements it stored to partials[i], 0 <= i < mfft_plan_partials(p):
 * the sum of the partials is sum |out|^2 (a deterministic fixed-order reduction). For the
 * loss of the differentiable path; no counterpart in the reference. */
int64_t mfft_plan_partials(const mfft_plan *p);
int mfft_execute_weighted_sumsq(const mfft_plan *p, const void *in, void *out, const double *pre,
                                const double *post, double *partials, void *stream);

/* The reference's twiddle routine (LeanFFTW/Core/Trig.lean:126, :90-92), exposed for
 * tests: (cos, sin)(2 pi k / n) and tan(2 pi num / den) as binary64. */
void mfft_cos_sin_2pi(uint64_t k, uint64_t n, double *c, double *s);
double mfft_tan_2pi(uint64_t num, uint64_t den);

/* The reference's pass table (LeanFFTW/Core/FastPlan.lean:27-37): for r in [1, R), ns
 * real parts then ns imaginary parts (2 (R-1) ns doubles). */
int mfft_pass_twiddles(int fwd, int64_t N, int R, int64_t ns, double *out);

/* The reference's operation count of a Stockham schedule (stockhamCost,
 * LeanFFTW/Gen/OpCount.lean:55-69): per pass t, N / R_t butterflies (twiddled from the second
 * pass on), each costing weights[0] per addition, [1] per multiplication, [2] per fused
 * multiply-add of its generated code, [3] per scalar load (2R inputs, 2(R-1) twiddle parts)
 * and [4] per scalar store (2R). -1 for an invalid schedule. */
int64_t mfft_stockham_cost(int fwd, int64_t n, const int *radices, int nradix,
                           const int64_t *weights);

/* The measuring planner's model of a schedule's GPU time, nanoseconds per transform: HBM
 * passes (kernel stages, split > 0 for two), LDS passes, arithmetic and twiddle traffic; it
 * orders and prunes the race (LeanFFTW/Asm/Planner.lean:185-205), the timings decide. */
double mfft_schedule_cost(int fwd, int64_t n, const int *radices, int nradix, int split);

/* Accepted for source compatibility with lfa_set_threads (Planner.lean:305-306); the GPU
 * grid replaces the CPU worker pool, so this only records the value. */
int mfft_set_threads(size_t n);

/* Autotuner cache ("wisdom"), JSON lines keyed by (device, n, sign, batch class). */
int mfft_set_wisdom_path(const char *path);

const char *mfft_error_string(int code);
const char *mfft_version(void);

#ifdef __cplusplus
}
#endif
#endif
