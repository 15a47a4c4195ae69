/-
The GPU plan kind for the reference's Lean API: `@[extern]` bindings to libmfft.so (the
MI355X library of this repository, include/mfft.h) through the C shim lfftw_gpu_shim.c.

This file goes into the reference's tree as LeanFFTW/GPU.lean (imported by LeanFFTW.lean);
its package links lfftw_gpu_shim.o and libmfft.so (`moreLinkArgs` in lakefile.lean). No Lean
toolchain is available to this repository, so it is not compiled here.

The shapes follow the reference's own externs: `IO` results and owned `FloatArray`s as in
LeanFFTW/Asm/Plan.lean:136-174 (lfa_run, lfa_addr, ...), `Option` plans for unsupported
problems as lfftw_generic_plan (LeanFFTW/FFI.lean:50-58), the problem spec of
LeanFFTW/FFI.lean:13-21.
-/
namespace LeanFFTW.GPU

opaque PlanPointed : NonemptyType
/-- A plan of libmfft (a batch of transforms on the GPU), freed when the last reference goes. -/
def Plan : Type := PlanPointed.type
instance : Nonempty Plan := PlanPointed.property

/-- `lfa_plan_measure`'s GPU counterpart: `howmany` contiguous transforms of `n` points in
direction `fwd` (1 forward, 0 backward), planned by the measuring planner on buffers of its
own (the caller's arrays are never touched). `none` when `n` is outside the engine's domain
(2 ≤ n ≤ 2^24, prime factors ≤ 13). -/
@[extern "lfftw_gpu_plan"]
opaque plan (n : USize) (fwd : UInt8) (howmany : USize) : IO (Option Plan)

/-- `lfftw_generic_plan`'s GPU counterpart: a plan from the generic problem spec, `none` for
the problems libmfft does not take (the caller keeps its own path). -/
@[extern "lfftw_gpu_plan_spec"]
opaque planSpec (spec : @& FloatArray) : IO (Option Plan)

/-- `out := transform(inp)` on host arrays (2·n·howmany doubles each, interleaved re/im):
copy to the GPU, transform, copy back. `out` is consumed and returned, as the reference's
`execFFI` (LeanFFTW/FFI.lean:54-58). -/
@[extern "lfftw_gpu_exec"]
opaque exec (p : @& Plan) (inp : @& FloatArray) (out : FloatArray) : IO FloatArray

/-- The radix sequence the plan's Stockham schedule uses (the reference's `rs`). -/
@[extern "lfftw_gpu_radices"]
opaque radices (p : @& Plan) : IO (Array Nat)

/-- A plan of the Native engine's kind (`LeanFFTW/Native/Exec.lean:29-95`): `make?` and
`exec` with the same signatures, backed by the GPU. -/
structure NativePlan where
  n : Nat
  fwd : Bool
  gpu : Plan

def NativePlan.make? (n : Nat) (fwd : Bool) : IO (Option NativePlan) := do
  match ← plan n.toUSize (if fwd then 1 else 0) 1 with
  | some p => return some ⟨n, fwd, p⟩
  | none => return none

def NativePlan.exec (p : NativePlan) (x : FloatArray) : IO FloatArray :=
  exec p.gpu x (FloatArray.mk (Array.mkArray (2 * p.n) 0.0))

end LeanFFTW.GPU
