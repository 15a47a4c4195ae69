/* Oracle: the binary64 reference transform (test infrastructure only).
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark load this
 * code; the shipped library never does.
 *
 * Restatement of refTransform / stockhamWith / passWith / bflyInputs
 * (LeanFFTW/Gen/Stockham.lean:77-115). Pass t of a plan rs has ns = R_0 ... R_{t-1};
 * for j < N/R, k = j mod ns, b = j / ns, butterfly j reads z[j + r*N/R] (r < R) and,
 * for t >= 1, the twiddles tw[2*r*ns + k] (re) and tw[(2r+1)*ns + k] (im) of its pass
 * table (FastPlan.lean:27-37), and output r goes to z'[b*ns*R + k + r*ns].
 * The butterflies are the generated straight-line kernels in bfly_gen.c.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef void (*bfly_fn)(const double *restrict, double *restrict);
bfly_fn oracle_bfly(int fwd, int R, int tw);
int oracle_has_radix(int R);

/* One Stockham pass on interleaved data, src -> dst. */
static void pass(bfly_fn k, int R, int64_t ns, int64_t N, const double *tw, const double *src,
                 double *dst)
{
     double in[4 * 64], out[2 * 64];
     const int64_t NR = N / R;
     for (int64_t j = 0; j < NR; ++j) {
          const int64_t kk = j % ns, b = j / ns;
          for (int r = 0; r < R; ++r) {
               in[2 * r] = src[2 * (j + r * NR)];
               in[2 * r + 1] = src[2 * (j + r * NR) + 1];
          }
          if (tw)
               for (int r = 0; r < R - 1; ++r) {
                    in[2 * R + 2 * r] = tw[2 * r * ns + kk];
                    in[2 * R + 2 * r + 1] = tw[(2 * r + 1) * ns + kk];
               }
          k(in, out);
          for (int r = 0; r < R; ++r) {
               const int64_t c = b * ns * R + kk + r * ns;
               dst[2 * c] = out[2 * r];
               dst[2 * c + 1] = out[2 * r + 1];
          }
     }
}

/* refTransform over `howmany` transforms of size N, batch stride N (complex elements).
 * tw: the plan's pass tables for passes 1..npass-1, concatenated (planTwiddles,
 * FastPlan.lean:40-47). work: 4N doubles of scratch. Returns 0, or -1 for a radix the
 * oracle does not carry or a plan whose product is not N. */
int oracle_ref_transform(int fwd, int64_t N, int npass, const int *rs, const double *tw,
                         int64_t howmany, const double *in, double *out, double *work)
{
     int64_t prod = 1;
     for (int t = 0; t < npass; ++t) {
          if (!oracle_has_radix(rs[t])) return -1;
          prod *= rs[t];
     }
     if (prod != N || npass < 1) return -1;
     for (int64_t h = 0; h < howmany; ++h) {
          const double *src = in + 2 * N * h;
          double *bufs[2] = {work, work + 2 * N};
          const double *twp = tw;
          int64_t ns = 1;
          for (int t = 0; t < npass; ++t) {
               const int R = rs[t];
               double *dst = (t == npass - 1) ? out + 2 * N * h : bufs[t & 1];
               pass(oracle_bfly(fwd, R, t >= 1), R, ns, N, t >= 1 ? twp : NULL, src, dst);
               if (t >= 1) twp += 2 * (int64_t)(R - 1) * ns;
               src = dst;
               ns *= R;
          }
     }
     return 0;
}

/* The CPU-baseline timing loop: the same one-transform plan applied `reps` times to the
 * same input and output (as libbench2's doit loops the reference's lfa_call,
 * bench/backends/bench-leanfftw.c:281-291), so that the timing holds no per-call overhead
 * of the caller. */
int oracle_ref_repeat(int fwd, int64_t N, int npass, const int *rs, const double *tw,
                      int64_t reps, const double *in, double *out, double *work)
{
     for (int64_t r = 0; r < reps; ++r)
          if (oracle_ref_transform(fwd, N, npass, rs, tw, 1, in, out, work) != 0) return -1;
     return 0;
}

/* One butterfly, for the tests: s has n_inputs(R, tw) doubles, y gets 2R. */
int oracle_butterfly(int fwd, int R, int tw, const double *s, double *y)
{
     bfly_fn k = oracle_bfly(fwd, R, tw);
     if (!k) return -1;
     k(s, y);
     return 0;
}
