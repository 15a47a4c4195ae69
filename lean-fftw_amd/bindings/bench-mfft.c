/* libbench2 backend for the MI355X library (libmfft.so): the GPU counterpart of the
 * reference's adapter bench/backends/bench-leanfftw.c, with the same shape and the same
 * problem domain, for the reference's benchmark protocol (tools/bench.py).
 *
 * Build beside the reference's other backends (FFTW 3.3.10's libbench2; not part of this
 * repository, so this file is not compiled here):
 *
 *   cc -O2 -Ilibbench2 -I$MFFT/include -I/opt/rocm/include -D__HIP_PLATFORM_AMD__ \
 *      bench-mfft.c libbench2/libbench2.a -L$MFFT/lean-fftw_amd/_build -lmfft \
 *      -L/opt/rocm/lib -lamdhip64 -lm -o bench-mfft
 *
 * Problems of the fast engine's domain (complex, rank 1, no batch, unit strides, 2 <= n <=
 * 2^24 with prime factors <= 13: fast_ok, bench-leanfftw.c:253-268) are planned with
 * mfft_plan_dft_1d under MFFT_MEASURE (lfa_plan_measure's counterpart, :234-249, which plans
 * on the caller's arrays; this one plans on buffers of its own). Other problems go through the
 * generic spec (mfft_plan_spec, the counterpart of lfftw_generic_plan, :159-171, whose spec
 * layout it builds the same way); what the library does not support exits as the reference's
 * adapter does for its generic engine.
 *
 * The arrays stay on the GPU between the transforms of one doit(iter): each doit copies the
 * input to the device once, runs `iter` transforms there (as the CPU engine keeps its data in
 * cache between lfa_call's), and copies the output back, so libbench2's times include one
 * host-to-device and one device-to-host copy per doit. The device-resident throughput of the
 * same kernels is what bench.py of this repository reports.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "bench-user.h"
#include "mfft.h"

static mfft_plan *gplan = 0;
static int spec_plan = 0;          /* 1: planned from the generic spec */
static void *din = 0, *dout = 0;   /* device copies of the problem's arrays */
static size_t in_bytes, out_bytes;
static int inplace = 0;
static unsigned plan_flags = MFFT_MEASURE;

BEGIN_BENCH_DOC
BENCH_DOC("name", "mfft")
BENCH_DOC("author", "MI355X (gfx950) FP64 FFT; plan/execute of the Lean FFTW fast path")
BENCH_DOC("notes", "arrays resident on the GPU within each doit; -oestimate plans without timing")
END_BENCH_DOC

static void check(hipError_t e, const char *what)
{
     if (e != hipSuccess) {
          fprintf(stderr, "mfft backend: %s: %s\n", what, hipGetErrorString(e));
          bench_exit(1);
     }
}

static int fast_ok(const bench_problem *p)
{
     size_t n;
     if (p->kind != PROBLEM_COMPLEX || p->split) return 0;
     if (p->sz->rnk != 1 || p->vecsz->rnk != 0) return 0;
     if (p->sz->dims[0].is != 1 || p->sz->dims[0].os != 1) return 0;
     n = (size_t) p->sz->dims[0].n;
     if (n < 2 || n > ((size_t) 1 << 24)) return 0;
     {
          static const size_t ps[] = { 2, 3, 5, 7, 11, 13 };
          for (size_t i = 0; i < sizeof ps / sizeof ps[0]; ++i)
               while (n % ps[i] == 0) n /= ps[i];
     }
     return n == 1;
}

int can_do(bench_problem *p)
{
     return p->kind == PROBLEM_COMPLEX || p->kind == PROBLEM_REAL;
}

/* the generic problem spec of the reference (LeanFFTW/FFI.lean:13-21), in doubles:
 * [kind, sign, rank, vrank, inRe, inIm, outRe, outIm, (n, is, os) x rank, (n, is, os) x vrank] */
static size_t build_spec(const bench_problem *p, double *buf)
{
     size_t k = 0;
     int i;
     const int interleaved = !p->split;
     /* strides in doubles: complex sides interleaved (2 per element); the real side of a
      * real-data problem (the input forward, the output backward) 1 per element */
     const int real_in = p->kind == PROBLEM_REAL && p->sign < 0;
     const int real_out = p->kind == PROBLEM_REAL && p->sign > 0;
     const double si = (interleaved && !real_in) ? 2.0 : 1.0;
     const double so = (interleaved && !real_out) ? 2.0 : 1.0;
     buf[k++] = p->kind == PROBLEM_COMPLEX ? 0.0 : 1.0;
     buf[k++] = p->sign;
     buf[k++] = p->sz->rnk;
     buf[k++] = p->vecsz->rnk;
     buf[k++] = 0;                  /* offsets are applied through the pointers */
     buf[k++] = interleaved ? 1 : 0;
     buf[k++] = 0;
     buf[k++] = interleaved ? 1 : 0;
     for (i = 0; i < p->sz->rnk; ++i) {
          buf[k++] = p->sz->dims[i].n;
          buf[k++] = p->sz->dims[i].is * si;
          buf[k++] = p->sz->dims[i].os * so;
     }
     for (i = 0; i < p->vecsz->rnk; ++i) {
          buf[k++] = p->vecsz->dims[i].n;
          buf[k++] = p->vecsz->dims[i].is * si;
          buf[k++] = p->vecsz->dims[i].os * so;
     }
     return k;
}

void setup(bench_problem *p)
{
     inplace = p->in == p->out;
     in_bytes = p->iphyssz * (p->kind == PROBLEM_COMPLEX ? sizeof(bench_complex) : sizeof(bench_real));
     out_bytes = p->ophyssz * (p->kind == PROBLEM_COMPLEX ? sizeof(bench_complex) : sizeof(bench_real));
     if (fast_ok(p)) {
          gplan = mfft_plan_dft_1d((int64_t) p->sz->dims[0].n,
                                   p->sign < 0 ? MFFT_FORWARD : MFFT_BACKWARD, 1, plan_flags);
          spec_plan = 0;
     }
     if (!gplan) {
          double spec[8 + 3 * 8];
          size_t len;
          if (p->sz->rnk + p->vecsz->rnk > 8) bench_exit(1);
          len = build_spec(p, spec);
          gplan = mfft_plan_spec(spec, len, plan_flags);
          spec_plan = 1;
     }
     if (!gplan) {                  /* NULL: unsupported, as the reference's empty plan */
          fprintf(stderr, "mfft: problem not supported\n");
          bench_exit(1);
     }
     check(hipMalloc(&din, in_bytes), "hipMalloc");
     dout = din;
     if (!inplace) check(hipMalloc(&dout, out_bytes), "hipMalloc");
}

void doit(int iter, bench_problem *p)
{
     int it, rc = MFFT_SUCCESS;
     check(hipMemcpy(din, p->in, in_bytes, hipMemcpyHostToDevice), "copy in");
     for (it = 0; it < iter && rc == MFFT_SUCCESS; ++it)
          rc = spec_plan ? mfft_execute_spec(gplan, (const double *) din, (double *) dout, NULL)
                         : mfft_execute(gplan, din, dout, NULL);
     if (rc != MFFT_SUCCESS) {
          fprintf(stderr, "mfft: %s\n", mfft_error_string(rc));
          bench_exit(1);
     }
     check(hipMemcpy(p->out, dout, out_bytes, hipMemcpyDeviceToHost), "copy out");
}

void done(bench_problem *p)
{
     UNUSED(p);
     if (gplan) mfft_destroy_plan(gplan);
     gplan = 0;
     if (din) (void) hipFree(din);
     if (dout && dout != din) (void) hipFree(dout);
     din = dout = 0;
}

void main_init(int *argc, char ***argv)
{
     UNUSED(argc);
     UNUSED(argv);
}

void initial_cleanup(void) {}
void final_cleanup(void) {}
void cleanup(void) {}

void useropt(const char *arg)
{
     if (!strcmp(arg, "estimate")) plan_flags = MFFT_ESTIMATE;
     else if (!strncmp(arg, "wisdom=", 7)) mfft_set_wisdom_path(arg + 7);
     else fprintf(stderr, "mfft: unknown user option: %s\n", arg);
}
