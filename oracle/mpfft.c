/* Oracle: a quad-precision DFT, the accuracy reference (test infrastructure only).
 *
 * The reference measures accuracy with libbench2's multiple-precision FFT (FFTW 3.3.10,
 * not vendored; driven by tools/accuracy.py:21-51) and, in its independent cross-check,
 * against long-double FFTW (docs/REPRODUCTION.md:92-103). This file is this build's own
 * replacement: a mixed-radix decimation-in-time DFT in IEEE binary128 (__float128,
 * 113-bit significand) with twiddles from libquadmath's cosq/sinq, accurate to about
 * 1e-32 relative for the sizes used here. It is checked against mpmath at 200 bits by
 * the tests.
 *
 *   X[k] = sum_j x[j] exp(sign * 2 pi i j k / N),   sign = -1 forward, +1 backward.
 */
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>

typedef __float128 q;

static int64_t smallest_factor(int64_t n)
{
     for (int64_t d = 2; d * d <= n; ++d)
          if (n % d == 0) return d;
     return n;
}

/* y[k] = DFT_n(x[0], x[s], ..., x[(n-1)s]); w[m] = exp(sign 2 pi i m / N), wstep = N/n. */
static void dft_rec(int64_t n, const q *xr, const q *xi, int64_t s, q *yr, q *yi, const q *wr,
                    const q *wi, int64_t wstep, int64_t N)
{
     if (n == 1) {
          yr[0] = xr[0];
          yi[0] = xi[0];
          return;
     }
     const int64_t p = smallest_factor(n), m = n / p;
     q *tr = malloc(sizeof(q) * n), *ti = malloc(sizeof(q) * n);
     for (int64_t j1 = 0; j1 < p; ++j1)
          dft_rec(m, xr + j1 * s, xi + j1 * s, s * p, tr + j1 * m, ti + j1 * m, wr, wi, wstep * p, N);
     for (int64_t k = 0; k < n; ++k) {
          q ar = 0, ai = 0;
          const int64_t km = k % m;
          for (int64_t j1 = 0; j1 < p; ++j1) {
               const int64_t e = ((j1 * k) % n) * wstep % N;
               const q cr = wr[e], ci = wi[e], vr = tr[j1 * m + km], vi = ti[j1 * m + km];
               ar += cr * vr - ci * vi;
               ai += cr * vi + ci * vr;
          }
          yr[k] = ar;
          yi[k] = ai;
     }
     free(tr);
     free(ti);
}

static void twiddles(int sign, int64_t N, q *wr, q *wi)
{
     const q tau = 2 * M_PIq;
     for (int64_t m = 0; m < N; ++m) {
          const q a = tau * (q)m / (q)N;
          wr[m] = cosq(a);
          wi[m] = sign * sinq(a);
     }
}

/* Exact-to-binary128 DFT of `howmany` interleaved transforms; the result is returned
 * as double-double pairs: hi (rounded to double) and lo = (exact - hi) rounded. */
int oracle_mp_dft(int sign, int64_t N, int64_t howmany, const double *in, double *hi, double *lo)
{
     q *wr = malloc(sizeof(q) * N), *wi = malloc(sizeof(q) * N);
     q *xr = malloc(sizeof(q) * N), *xi = malloc(sizeof(q) * N);
     q *yr = malloc(sizeof(q) * N), *yi = malloc(sizeof(q) * N);
     if (!wr || !wi || !xr || !xi || !yr || !yi) return -1;
     twiddles(sign, N, wr, wi);
     for (int64_t h = 0; h < howmany; ++h) {
          for (int64_t j = 0; j < N; ++j) {
               xr[j] = in[2 * (h * N + j)];
               xi[j] = in[2 * (h * N + j) + 1];
          }
          dft_rec(N, xr, xi, 1, yr, yi, wr, wi, 1, N);
          for (int64_t k = 0; k < N; ++k) {
               const double a = (double)yr[k], b = (double)yi[k];
               hi[2 * (h * N + k)] = a;
               hi[2 * (h * N + k) + 1] = b;
               lo[2 * (h * N + k)] = (double)(yr[k] - (q)a);
               lo[2 * (h * N + k) + 1] = (double)(yi[k] - (q)b);
          }
     }
     free(wr); free(wi); free(xr); free(xi); free(yr); free(yi);
     return 0;
}

/* Relative L2 error ||y - Y||_2 / ||Y||_2 of each of `howmany` transforms y (interleaved
 * doubles) against the binary128 DFT Y of `in`, accumulated in binary128. */
int oracle_rel_l2(int sign, int64_t N, int64_t howmany, const double *in, const double *y,
                  double *err)
{
     q *wr = malloc(sizeof(q) * N), *wi = malloc(sizeof(q) * N);
     q *xr = malloc(sizeof(q) * N), *xi = malloc(sizeof(q) * N);
     q *yr = malloc(sizeof(q) * N), *yi = malloc(sizeof(q) * N);
     if (!wr || !wi || !xr || !xi || !yr || !yi) return -1;
     twiddles(sign, N, wr, wi);
     for (int64_t h = 0; h < howmany; ++h) {
          for (int64_t j = 0; j < N; ++j) {
               xr[j] = in[2 * (h * N + j)];
               xi[j] = in[2 * (h * N + j) + 1];
          }
          dft_rec(N, xr, xi, 1, yr, yi, wr, wi, 1, N);
          q num = 0, den = 0;
          for (int64_t k = 0; k < N; ++k) {
               const q dr = (q)y[2 * (h * N + k)] - yr[k], di = (q)y[2 * (h * N + k) + 1] - yi[k];
               num += dr * dr + di * di;
               den += yr[k] * yr[k] + yi[k] * yi[k];
          }
          err[h] = den > 0 ? (double)sqrtq(num / den) : (double)sqrtq(num);
     }
     free(wr); free(wi); free(xr); free(xi); free(yr); free(yi);
     return 0;
}
