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
    ("mfft_reg.hip", True),
    ("mfft_split.hip", True),
    ("mfft_tile.hip", True),
    ("mfft_tile2d.hip", True),
    *((f"mfft_tile_gen{i}.hip", True) for i in range(8)),
    ("mfft_real.hip", True),
    ("mfft_ragged.hip", True),
    ("mfft_static.hip", True),
    ("mfft_plan.cpp", False),
    ("mfft_trig.cpp", False),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def _stale(obj: str, deps: list[str]) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = False, force: bool = True) -> str:
    """Compile every translation unit (in parallel) and link. With force=False, objects newer
    than their source, the shared headers and this script are reused."""
    sys.path.insert(0, HERE)
    try:
        import codegen  # noqa: WPS433
    finally:
        sys.path.pop(0)
    codegen.main()
    os.makedirs(OUT, exist_ok=True)
    cc = hipcc()
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers += [os.path.join(HERE, "..", "include", "mfft.h"), os.path.abspath(__file__)]
    jobs, objs = [], []
    for src, device in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OUT, src.rsplit(".", 1)[0] + ".o")
        flags = HIP_FLAGS if device else HOST_FLAGS
        cmd = [cc, *flags, "-c", path, "-o", obj, "-I", CSRC]
        if device:
            cmd[1:1] = ["-x", "hip"]
        objs.append(obj)
        if force or _stale(obj, [path, *headers]):
            jobs.append(cmd)
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    failed = [p.args for p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, force="--incremental" not in sys.argv))
