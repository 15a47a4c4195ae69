"""GPU runs of the libbench2-shaped command line (lean-fftw_amd/bench2.py): its verify tests
(linearity, impulse, time shift) on the transforms the protocol covers, and a speed line in
the format the reference's protocol runner parses (tools/bench.py:45-49)."""

import importlib.util
import os
import re

import pytest

from conftest import ROOT, gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():  # pragma: no cover - CPU-only containers
    pytest.skip("needs a GPU", allow_module_level=True)

import torch  # noqa: E402


def _bench2():
    spec = importlib.util.spec_from_file_location(
        "mfft_bench2", os.path.join(ROOT, "lean-fftw_amd", "bench2.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("prob", ["ocf2", "ocf1024", "icb1000", "ocf262144", "ocb165375",
                                  "ocf64x48", "icb32x32", "ocf1024x1024", "orf4096", "orb1000",
                                  "orf7", "orb2", "orb65536"])
def test_verify(lf, prob):
    b = _bench2()
    e_lin, e_imp, e_shift = b.verify(lf, torch, b.parse_problem(prob))
    assert e_lin < 1e-13 and e_imp < 1e-13 and e_shift < 1e-13, (e_lin, e_imp, e_shift)


def test_speed_line(lf):
    b = _bench2()
    line = b.speed(lf, torch, b.parse_problem("ocf4096"), time_min=0.002, reps=2)
    m = re.search(r"^Problem: ocf4096, setup:\s*([0-9.e+-]+)\s*(ns|us|ms|s),\s*"
                  r"time:\s*([0-9.e+-]+)\s*(ns|us|ms|s), ``mflops'': ([0-9.]+)$", line)
    assert m, line
    assert float(m.group(3)) > 0 and float(m.group(5)) > 0
