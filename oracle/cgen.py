"""Oracle: C emission of the symbolic butterflies (test infrastructure only).

Writes straight-line C for butterflyK fwd R tw (LeanFFTW/Gen/Stockham.lean:65-66) from
the CSE'd kernels of oracle.symgen, with constants as exact hexadecimal literals
(Cst.cLiteral, LeanFFTW/Gen/SExpr.lean:47-52) and every fused operation through C99
fma(). The file is compiled with -ffp-contract=off so that the C compiler neither
fuses nor reorders anything: each operation is the binary64 operation of
floatArith (LeanFFTW/Gen/Emit.lean:46-55).
"""

from __future__ import annotations

from . import symgen, trig

# Radices the oracle carries: the reference's Stockham radices
# (LeanFFTW/Asm/Kernels.lean:1345) plus the single-codelet radices of small plans.
RADICES = (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 24, 32, 36, 64)


def c_literal(x: float) -> str:
    b = trig.float_to_bits(x)
    if b & 0x7FFFFFFFFFFFFFFF == 0:
        return "0.0"
    return x.hex()


def emit_kernel(fwd: bool, R: int, tw: bool) -> str:
    n_in = symgen.n_inputs(R, tw)
    ops, outs = symgen.cse_kernel(n_in, symgen.butterfly(fwd, R, tw))
    name = f"bf_{'f' if fwd else 'b'}{R}{'t' if tw else 'n'}"
    lines = [f"static void {name}(const double *restrict s, double *restrict y) {{"]

    def ref(slot: int) -> str:
        return f"s[{slot}]" if slot < n_in else f"v{slot}"

    for i, op in enumerate(ops):
        slot = n_in + i
        k = op[0]
        if k == "zero":
            rhs = "0.0"
        elif k == "cst":
            rhs = c_literal(trig.cst_value(op[1]))
        elif k == "add":
            rhs = f"{ref(op[1])} + {ref(op[2])}"
        elif k == "sub":
            rhs = f"{ref(op[1])} - {ref(op[2])}"
        elif k == "mul":
            rhs = f"{ref(op[1])} * {ref(op[2])}"
        elif k == "neg":
            rhs = f"-{ref(op[1])}"
        elif k == "fma":
            rhs = f"fma({ref(op[1])}, {ref(op[2])}, {ref(op[3])})"
        elif k == "fms":
            rhs = f"fma({ref(op[1])}, {ref(op[2])}, -{ref(op[3])})"
        elif k == "fnma":
            rhs = f"fma(-{ref(op[1])}, {ref(op[2])}, {ref(op[3])})"
        else:
            raise ValueError(op)
        lines.append(f"  const double v{slot} = {rhs};")
    for i, s in enumerate(outs):
        lines.append(f"  y[{i}] = {ref(s)};")
    lines.append("}")
    return "\n".join(lines)


def emit_vkernel(fwd: bool, R: int, tw: bool) -> str:
    """The same butterfly on v4d vectors (four transforms, one per lane): every operation of
    emit_kernel applied lane by lane -- vector add, subtract, multiply and negate are the
    binary64 operations of each lane, and fma is the lane-wise fused multiply-add -- so each
    lane rounds exactly as the scalar kernel. Followed by its Stockham pass with the kernel
    inlined (the CPU-baseline timing path; the scalar kernels stay the checker)."""
    n_in = symgen.n_inputs(R, tw)
    ops, outs = symgen.cse_kernel(n_in, symgen.butterfly(fwd, R, tw))
    tag = f"{'f' if fwd else 'b'}{R}{'t' if tw else 'n'}"
    lines = [f"static inline void vbf_{tag}(const v4d *restrict s, v4d *restrict y) {{"]

    def ref(slot: int) -> str:
        return f"s[{slot}]" if slot < n_in else f"v{slot}"

    for i, op in enumerate(ops):
        slot = n_in + i
        k = op[0]
        if k == "zero":
            rhs = "VB(0.0)"
        elif k == "cst":
            rhs = f"VB({c_literal(trig.cst_value(op[1]))})"
        elif k == "add":
            rhs = f"{ref(op[1])} + {ref(op[2])}"
        elif k == "sub":
            rhs = f"{ref(op[1])} - {ref(op[2])}"
        elif k == "mul":
            rhs = f"{ref(op[1])} * {ref(op[2])}"
        elif k == "neg":
            rhs = f"-{ref(op[1])}"
        elif k == "fma":
            rhs = f"VFMA({ref(op[1])}, {ref(op[2])}, {ref(op[3])})"
        elif k == "fms":
            rhs = f"VFMA({ref(op[1])}, {ref(op[2])}, -{ref(op[3])})"
        elif k == "fnma":
            rhs = f"VFMA(-{ref(op[1])}, {ref(op[2])}, {ref(op[3])})"
        else:
            raise ValueError(op)
        lines.append(f"  const v4d v{slot} = {rhs};")
    for i, s_ in enumerate(outs):
        lines.append(f"  y[{i}] = {ref(s_)};")
    lines.append("}")
    # the pass (Stockham.lean:88-98) over lane-interleaved data: element e of the four
    # transforms at z[2 e] (re) and z[2 e + 1] (im); twiddles broadcast to the lanes
    lines.append(f"static void vpass_{tag}(int64_t ns, int64_t N, const double *tw, "
                 "const v4d *restrict src, v4d *restrict dst) {")
    lines.append(f"  const int64_t NR = N / {R};")
    lines.append(f"  v4d in[{n_in}], out[{2 * R}];")
    lines.append("  for (int64_t b = 0; b < NR / ns; ++b)  /* butterfly j = b ns + kk */")
    lines.append("  for (int64_t kk = 0; kk < ns; ++kk) {")
    lines.append("    const int64_t j = b * ns + kk;")
    lines.append(f"    for (int r = 0; r < {R}; ++r) {{")
    lines.append("      in[2 * r] = src[2 * (j + r * NR)];")
    lines.append("      in[2 * r + 1] = src[2 * (j + r * NR) + 1];")
    lines.append("    }")
    if tw:
        lines.append(f"    for (int r = 0; r < {R - 1}; ++r) {{")
        lines.append(f"      in[{2 * R} + 2 * r] = VB(tw[2 * r * ns + kk]);")
        lines.append(f"      in[{2 * R} + 2 * r + 1] = VB(tw[(2 * r + 1) * ns + kk]);")
        lines.append("    }")
    lines.append(f"    vbf_{tag}(in, out);")
    lines.append(f"    for (int r = 0; r < {R}; ++r) {{")
    lines.append(f"      const int64_t c = b * ns * {R} + kk + r * ns;")
    lines.append("      dst[2 * c] = out[2 * r];")
    lines.append("      dst[2 * c + 1] = out[2 * r + 1];")
    lines.append("    }")
    lines.append("  }")
    lines.append("}")
    return "\n".join(lines)


def emit_vector_all(radices=RADICES) -> str:
    parts = [
        "/* Generated by oracle/cgen.py: the oracle's butterflies on four lanes (CPU-baseline",
        " * timing path). Do not edit. */",
        "#include <immintrin.h>",
        "#include <stdint.h>",
        "#include <stddef.h>",
        "typedef double v4d __attribute__((vector_size(32)));",
        "#define VB(x) ((v4d){(x), (x), (x), (x)})",
        "#define VFMA(a, b, c) ((v4d)_mm256_fmadd_pd((__m256d)(a), (__m256d)(b), (__m256d)(c)))",
        "typedef void (*vpass_fn)(int64_t, int64_t, const double *, const v4d *restrict,"
        " v4d *restrict);",
    ]
    entries = []
    for R in radices:
        for fwd in (True, False):
            for tw in (False, True):
                parts.append(emit_vkernel(fwd, R, tw))
                entries.append((R, fwd, tw))
    parts.append("vpass_fn oracle_vpass(int fwd, int R, int tw) {")
    parts.append("  switch (R * 4 + (fwd ? 2 : 0) + (tw ? 1 : 0)) {")
    for R, fwd, tw in entries:
        key = R * 4 + (2 if fwd else 0) + (1 if tw else 0)
        parts.append(f"  case {key}: return vpass_{'f' if fwd else 'b'}{R}{'t' if tw else 'n'};")
    parts.append("  default: return NULL;\n  }\n}")
    return "\n".join(parts) + "\n"


def emit_all(radices=RADICES) -> str:
    parts = [
        "/* Generated by oracle/cgen.py: the oracle's butterflies. Do not edit. */",
        "#include <math.h>",
        "#include <stddef.h>",
        "typedef void (*bfly_fn)(const double *restrict, double *restrict);",
    ]
    entries = []
    for R in radices:
        for fwd in (True, False):
            for tw in (False, True):
                parts.append(emit_kernel(fwd, R, tw))
                entries.append((R, fwd, tw))
    parts.append("bfly_fn oracle_bfly(int fwd, int R, int tw) {")
    parts.append("  switch (R * 4 + (fwd ? 2 : 0) + (tw ? 1 : 0)) {")
    for R, fwd, tw in entries:
        key = R * 4 + (2 if fwd else 0) + (1 if tw else 0)
        parts.append(f"  case {key}: return bf_{'f' if fwd else 'b'}{R}{'t' if tw else 'n'};")
    parts.append("  default: return NULL;\n  }\n}")
    parts.append("int oracle_has_radix(int R) {\n  switch (R) {")
    for R in radices:
        parts.append(f"  case {R}:")
    parts.append("    return 1;\n  default: return 0;\n  }\n}")
    return "\n".join(parts) + "\n"
