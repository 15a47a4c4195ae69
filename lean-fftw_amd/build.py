"""Build libmfft.so for gfx950 in-tree (lean-fftw_amd/_build/).

    python lean-fftw_amd/build.py

Steps: regenerate csrc/bfly_gen.h from codegen.py, compile the HIP kernels and the host
library with hipcc, link one shared library. Every HIP translation unit is compiled with
-ffp-contract=off (the butterflies spell out every fused operation) and with packed FP32
instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libmfft.so")
ARCH = os.environ.get("MFFT_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]

SOURCES = [
    ("mfft_kernels.hip", True),
    ("mfft_ragged.hip", True),
    ("mfft_plan.cpp", False),
    ("mfft_trig.cpp", False),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    sys.path.insert(0, HERE)
    try:
        import codegen  # noqa: WPS433
    finally:
        sys.path.pop(0)
    codegen.main()
    os.makedirs(OUT, exist_ok=True)
    objs = []
    cc = hipcc()
    for src, device in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OUT, src.rsplit(".", 1)[0] + ".o")
        flags = HIP_FLAGS if device else HOST_FLAGS
        cmd = [cc, *flags, "-c", path, "-o", obj, "-I", CSRC]
        if device:
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
