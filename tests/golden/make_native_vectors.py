"""Write tests/golden/native_bfly.json: input/output vectors of the reference's own
printed butterflies.

LeanFFTW/Native/Kernels.lean (generated by the reference from butterflyK, see its
header) holds the straight-line binary64 text of bfly{2,4,8,16}{n,t}{f,b}: untwiddled
and twiddled radix-2^k butterflies, forward and backward, constants as exact bit
patterns. This script reads that file as text (it is not run: there is no Lean here),
evaluates each function line by line in binary64 -- +, -, * in IEEE double, Float.fma
through C's fma() (libm, single rounding), `cst b` as the double with bits b -- on
seeded inputs, and stores only the input and output bit patterns. No text of the
reference is kept.

Slot layout of a butterfly (LeanFFTW/Gen/Stockham.lean:36-42): x_r in slots 2r, 2r+1;
twiddle w_r (r >= 1) in slots 2R + 2(r-1), +1.

Run from the repository root:  python tests/golden/make_native_vectors.py
"""

from __future__ import annotations

import ctypes
import json
import os
import re
import struct
import sys

SRC = "/root/reference/LeanFFTW/Native/Kernels.lean"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native_bfly.json")

_libm = ctypes.CDLL("libm.so.6")
_libm.fma.argtypes = [ctypes.c_double] * 3
_libm.fma.restype = ctypes.c_double


def f64(bits: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


HDR = re.compile(r"def bfly(\d+)([nt])([fb]) ")
RD_X = re.compile(r"let i(\d+) := rd x \(2 \* \(src \+ (\d+) \* sstr\)( \+ 1)?\)")
RD_Y = re.compile(r"let i(\d+) := rd y \(2 \* \(c \+ (\d+) \* ns\)( \+ 1)?\)")
RD_T = re.compile(r"let i(\d+) := rd tw \((\d+) \* ns \+ q\)")
WR = re.compile(r"let y := wr y \(2 \* \((?:dst|c) \+ (\d+) \* (?:dstr|ns)\)( \+ 1)?\) ([it]\d+)")
BIN = re.compile(r"let t(\d+) := ([it]\d+) ([-+*]) ([it]\d+)$")
FMA = re.compile(r"let t(\d+) := Float\.fma (\(-)?([it]\d+)\)? ([it]\d+) (\(-)?([it]\d+)\)?$")
CST = re.compile(r"let t(\d+) := cst (\d+)$")
NEG = re.compile(r"let t(\d+) := -([it]\d+)$")


def parse(text: str):
    funcs = {}
    cur = None
    for raw in text.splitlines():
        line = raw.strip()
        m = HDR.search(line)
        if m:
            cur = (int(m.group(1)), m.group(2) == "t", m.group(3) == "f")
            funcs[cur] = []
            continue
        if cur is None or not line.startswith("let "):
            if line == "" or line.startswith("@[") or line.startswith("def "):
                pass
            continue
        funcs[cur].append(line)
    return funcs


def evaluate(R: int, tw: bool, lines, s):
    env = {}
    y = [None] * (2 * R)
    for line in lines:
        m = RD_X.match(line) or RD_Y.match(line)
        if m:
            env["i" + m.group(1)] = s[2 * int(m.group(2)) + (1 if m.group(3) else 0)]
            continue
        m = RD_T.match(line)
        if m:
            # tw[(2(r-1)) * ns + q] is Re w_r, tw[(2(r-1)+1) * ns + q] is Im w_r
            env["i" + m.group(1)] = s[2 * R + int(m.group(2))]
            continue
        m = WR.match(line)
        if m:
            y[2 * int(m.group(1)) + (1 if m.group(2) else 0)] = env[m.group(3)]
            continue
        m = BIN.match(line)
        if m:
            a, b = env[m.group(2)], env[m.group(4)]
            env["t" + m.group(1)] = a + b if m.group(3) == "+" else a - b if m.group(3) == "-" else a * b
            continue
        m = FMA.match(line)
        if m:
            a, b, c = env[m.group(3)], env[m.group(4)], env[m.group(6)]
            if m.group(2):
                a = -a
            if m.group(5):
                c = -c
            env["t" + m.group(1)] = _libm.fma(a, b, c)
            continue
        m = CST.match(line)
        if m:
            env["t" + m.group(1)] = f64(int(m.group(2)))
            continue
        m = NEG.match(line)
        if m:
            env["t" + m.group(1)] = -env[m.group(2)]
            continue
        if line.startswith("let y := y") or line == "y":
            continue
        raise ValueError(f"unparsed line: {line}")
    assert all(v is not None for v in y)
    return y


def lcg(seed: int, n: int):
    M = (1 << 64) - 1
    s = seed & M
    out = []
    for _ in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & M
        out.append((s >> 11) / 9007199254740992.0 - 0.5)
    return out


def main():
    if not os.path.exists(SRC):
        sys.exit(f"{SRC} not found: the fixtures are generated where the reference is mounted")
    funcs = parse(open(SRC).read())
    assert len(funcs) == 16, sorted(funcs)
    cases = []
    for (R, tw, fwd), lines in sorted(funcs.items()):
        n_in = 4 * R - 2 if tw else 2 * R
        for rnd in range(4):
            s = lcg(0x9E3779B97F4A7C15 ^ (R * 131 + tw * 7 + fwd * 3 + rnd * 1000003), n_in)
            if tw:
                # unit-modulus twiddles, as in the pass tables
                for r in range(R - 1):
                    re, im = s[2 * R + 2 * r], s[2 * R + 2 * r + 1]
                    h = (re * re + im * im) ** 0.5
                    s[2 * R + 2 * r], s[2 * R + 2 * r + 1] = re / h, im / h
            y = evaluate(R, tw, lines, s)
            cases.append({
                "R": R, "tw": tw, "fwd": fwd,
                "in": [bits(v) for v in s], "out": [bits(v) for v in y],
            })
    with open(OUT, "w") as f:
        json.dump({"source": "LeanFFTW/Native/Kernels.lean (evaluated)", "cases": cases}, f)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
