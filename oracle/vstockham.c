/* Oracle, CPU-baseline timing path (test infrastructure only): refTransform on four
 * transforms at once, one per lane of a four-double vector.
 *
 * Only the CPU-baseline leg of the benchmark and the tests load this code. The passes are
 * oracle_ref_transform's (LeanFFTW/Gen/Stockham.lean:77-115, refstockham.c), with the
 * generated butterflies of vbfly_gen.c inlined in a pass loop per radix; every lane runs the
 * scalar oracle's operations in the same order, so the results are its bits (a CPU test
 * checks this). The four transforms are moved into and out of the lane-interleaved layout
 * (element e of lane l at z[(2 e + c) * 4 + l], c = 0 real, 1 imaginary) inside each call.
 */
#include <stdint.h>
#include <stddef.h>

typedef double v4d __attribute__((vector_size(32)));
typedef void (*vpass_fn)(int64_t, int64_t, const double *, const v4d *restrict, v4d *restrict);
vpass_fn oracle_vpass(int fwd, int R, int tw);
int oracle_has_radix(int R);

/* Four transforms of size N (batch stride N complex values) from in to out. tw: the plan's
 * pass tables for passes 1..npass-1, as oracle_ref_transform. work: 24 N doubles, 32-byte
 * aligned. Returns 0, or -1 for a radix the oracle does not carry or a plan whose product is
 * not N. */
int oracle_ref_transform_x4(int fwd, int64_t N, int npass, const int *rs, const double *tw,
                            const double *in, double *out, double *work)
{
     int64_t prod = 1;
     for (int t = 0; t < npass; ++t) {
          if (!oracle_has_radix(rs[t]) || !oracle_vpass(fwd, rs[t], t >= 1)) return -1;
          prod *= rs[t];
     }
     if (prod != N || npass < 1 || ((uintptr_t)work & 31)) return -1;
     v4d *bufs[3] = {(v4d *)work, (v4d *)work + 2 * N, (v4d *)work + 4 * N};
     v4d *z = bufs[2];
     for (int64_t e = 0; e < 2 * N; ++e)
          z[e] = (v4d){in[e], in[2 * N + e], in[4 * N + e], in[6 * N + e]};
     const v4d *src = z;
     const double *twp = tw;
     int64_t ns = 1;
     for (int t = 0; t < npass; ++t) {
          const int R = rs[t];
          v4d *dst = bufs[t & 1];
          oracle_vpass(fwd, R, t >= 1)(ns, N, t >= 1 ? twp : NULL, src, dst);
          if (t >= 1) twp += 2 * (int64_t)(R - 1) * ns;
          src = dst;
          ns *= R;
     }
     for (int64_t e = 0; e < 2 * N; ++e) {
          const v4d v = src[e];
          out[e] = v[0];
          out[2 * N + e] = v[1];
          out[4 * N + e] = v[2];
          out[6 * N + e] = v[3];
     }
     return 0;
}

/* The timing loop of the CPU baseline: the four-transform call `reps` times on the same
 * input and output, as oracle_ref_repeat does for one transform. */
int oracle_ref_repeat_x4(int fwd, int64_t N, int npass, const int *rs, const double *tw,
                         int64_t reps, const double *in, double *out, double *work)
{
     for (int64_t r = 0; r < reps; ++r)
          if (oracle_ref_transform_x4(fwd, N, npass, rs, tw, in, out, work) != 0) return -1;
     return 0;
}
