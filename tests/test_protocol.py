"""The protocol's statistics (lean-fftw_amd/protocol.py) against the reference's published
analysis of its own results (results/emr-i12/results.json, kept as the fixture
tests/golden/emr_i12_results.json; summary.txt and analysis.json of the same run): medians,
G and the bootstrap 95% intervals, every digit printed."""

import json
import os

from conftest import GOLDEN, ROOT


def load_protocol():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "mfft_protocol", os.path.join(ROOT, "lean-fftw_amd", "protocol.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


# results/emr-i12/summary.txt of the reference
PUBLISHED = {
    "fftw": (0.7497, 0.7340, 0.7720),
    "fftw-exhaustive": (0.7264, 0.7074, 0.7446),
    "fftw-measure": (0.6615, 0.6442, 0.6771),
    "fftw-patient": (0.7061, 0.6912, 0.7268),
    "lean-i11": (0.9290, 0.9102, 0.9568),
}


def test_reproduces_published_g_and_intervals():
    P = load_protocol()
    rows = json.load(open(os.path.join(GOLDEN, "emr_i12_results.json")))
    out = P.analyze(rows, "lean-i12", {"fftw": ["fftw-measure", "fftw-patient", "fftw-exhaustive"]})
    got = {c: (round(v["G"], 4), round(v["ci95"][0], 4), round(v["ci95"][1], 4))
           for c, v in out["vs"].items()}
    assert got == PUBLISHED
    # per-size medians of the summary table (4 significant digits)
    med = out["medians"]
    assert f"{med['lean-i12'][1024]:.4g}" == "1.075e-06"
    assert f"{med['fftw'][131072]:.4g}" == "0.0005522"
    assert f"{med['fftw-patient'][2]:.4g}" == "3.74e-09"
    assert len(out["sizes"]) == 18
    text = P.report(out, "lean-i12")
    assert "lean-i12 / fftw: G = 0.7497  95% CI [0.7340, 0.7720]" in text


def test_g_closed_forms():
    """G of an implementation against itself scaled by c at every size is 1/c, and the
    interval of identical runs collapses to G."""
    P = load_protocol()
    rows = []
    for n in (8, 64, 1024):
        for r in range(5):
            rows.append({"impl": "a", "n": n, "run": r, "time": 1e-6 * n * (1 + 0.01 * r)})
            rows.append({"impl": "b", "n": n, "run": r, "time": 2e-6 * n * (1 + 0.01 * r)})
            rows.append({"impl": "c", "n": n, "run": r, "time": 3e-6 * n})
            rows.append({"impl": "d", "n": n, "run": r, "time": 1.5e-6 * n})
    out = P.analyze(rows, "c", {"ab": ["a", "b"]}, B=200)
    assert abs(out["vs"]["d"]["G"] - 2.0) < 1e-12
    assert out["vs"]["d"]["ci95"] == [2.0, 2.0]
    assert abs(out["vs"]["ab"]["G"] - 3.0 / 1.02) < 1e-12  # min of the members' medians
