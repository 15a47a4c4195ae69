/* C side of LeanFFTW/GPU.lean: Lean external objects around libmfft plans (include/mfft.h).
 * Compiled with the reference's Lean toolchain (leanc) and linked with libmfft.so; no Lean
 * toolchain is available to this repository, so it is not compiled here.
 *
 * Conventions as the reference's own externs (LeanFFTW/Asm/Plan.lean:136-174): IO results
 * built with lean_io_result_mk_ok / _mk_error; `@&` arguments borrowed, others owned. */
#include <lean/lean.h>

#include "mfft.h"

static lean_external_class *plan_class = NULL;

static void plan_finalize(void *p) { mfft_destroy_plan((mfft_plan *) p); }
static void plan_foreach(void *p, b_lean_obj_arg f) { (void) p; (void) f; }

static lean_obj_res io_error(int rc)
{
     return lean_io_result_mk_error(lean_mk_io_user_error(lean_mk_string(mfft_error_string(rc))));
}

/* Option Plan: none (a boxed 0) when the library returns NULL, as the reference's empty plan
 * array / `none` */
static lean_obj_res some_plan(mfft_plan *p)
{
     lean_object *o;
     if (!p) return lean_io_result_mk_ok(lean_box(0));
     if (!plan_class) plan_class = lean_register_external_class(plan_finalize, plan_foreach);
     o = lean_alloc_ctor(1, 1, 0);
     lean_ctor_set(o, 0, lean_alloc_external(plan_class, p));
     return lean_io_result_mk_ok(o);
}

LEAN_EXPORT lean_obj_res lfftw_gpu_plan(size_t n, uint8_t fwd, size_t howmany, lean_obj_arg w)
{
     (void) w;
     return some_plan(mfft_plan_dft_1d((int64_t) n, fwd ? MFFT_FORWARD : MFFT_BACKWARD,
                                       (int64_t) howmany, MFFT_MEASURE));
}

LEAN_EXPORT lean_obj_res lfftw_gpu_plan_spec(b_lean_obj_arg spec, lean_obj_arg w)
{
     (void) w;
     return some_plan(mfft_plan_spec(lean_float_array_cptr(spec), lean_sarray_size(spec),
                                     MFFT_ESTIMATE));
}

LEAN_EXPORT lean_obj_res lfftw_gpu_exec(b_lean_obj_arg plan, b_lean_obj_arg in, lean_obj_arg out,
                                        lean_obj_arg w)
{
     mfft_plan *p = (mfft_plan *) lean_get_external_data(plan);
     int rc;
     (void) w;
     if (!lean_is_exclusive(out)) out = lean_copy_float_array(out);
     rc = mfft_execute_host(p, lean_float_array_cptr(in), lean_float_array_cptr(out));
     if (rc != MFFT_SUCCESS) {
          lean_dec(out);
          return io_error(rc);
     }
     return lean_io_result_mk_ok(out);
}

LEAN_EXPORT lean_obj_res lfftw_gpu_radices(b_lean_obj_arg plan, lean_obj_arg w)
{
     mfft_plan *p = (mfft_plan *) lean_get_external_data(plan);
     int rs[64];
     int k = mfft_plan_radices(p, rs, 64), i;
     lean_object *arr;
     (void) w;
     if (k < 0 || k > 64) return io_error(MFFT_ERR_INVALID);
     arr = lean_alloc_array(0, (size_t) k);
     for (i = 0; i < k; ++i) arr = lean_array_push(arr, lean_usize_to_nat((size_t) rs[i]));
     return lean_io_result_mk_ok(arr);
}
