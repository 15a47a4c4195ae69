"""Copy the reference's published numbers for this path into small JSON fixtures.

Data only (no source text). Sources, all under results/ of the reference:
  repro-2026-11-04/kernel-constants.txt  -> kernel_constants.json
      (name, k, n, binary64 bits) of every constant of butterflyK fwd R tw, R <= 64,
      in the order the reference's dump printed them (scripts/ErrorBound.lean:64-83)
  repro-2026-11-04/twiddles.jsonl        -> twiddle_tau.json
      per-N max |error| of the base table against long-double cos/sin (3 digits)
  repro-2026-11-04/bounds.jsonl          -> bounds.json
      per plan: rounding depth a, constant depth b, magMax (scripts/ErrorBound.lean:49-59)
  emr-i12/accuracy/accuracy.json         -> accuracy_ref.json["pow2"]   (lean-i12)
  spr-nonpow2-i9/accuracy.json           -> accuracy_ref.json["nonpow2"] (every engine run)
  repro-2026-11-04/xcheck-patient.jsonl  -> accuracy_ref.json["xcheck"] (lean_l2, lean_R)

Run from the repository root:  python tests/golden/make_reference_fixtures.py
"""

from __future__ import annotations

import json
import os
import sys

REF = "/root/reference/results"
HERE = os.path.dirname(os.path.abspath(__file__))


def jsonl(path):
    out = []
    for line in open(path):
        line = line.strip()
        if line:
            out.append(json.loads(line))
    return out


def main():
    if not os.path.isdir(REF):
        sys.exit("reference results not mounted")
    consts = []
    for line in open(f"{REF}/repro-2026-11-04/kernel-constants.txt"):
        nm, k, n, b = line.split()
        consts.append([nm, int(k), int(n), int(b)])
    json.dump(consts, open(f"{HERE}/kernel_constants.json", "w"))

    tau = [[d["N"], d["tau_abs"]] for d in jsonl(f"{REF}/repro-2026-11-04/twiddles.jsonl") if "N" in d]
    json.dump(tau, open(f"{HERE}/twiddle_tau.json", "w"))

    bounds = [
        {"N": d["N"], "R": d["R"], "a": d["a"], "b": d["b"], "magMax": d["magMax"]}
        for d in jsonl(f"{REF}/repro-2026-11-04/bounds.jsonl")
    ]
    json.dump(bounds, open(f"{HERE}/bounds.json", "w"), indent=0)

    acc = {}
    a = json.load(open(f"{REF}/emr-i12/accuracy/accuracy.json"))["results"]["lean-i12"]
    acc["pow2"] = {n: {"fwd_l2": v["fwd_l2"], "bwd_l2": v["bwd_l2"]} for n, v in a.items()}
    b = json.load(open(f"{REF}/spr-nonpow2-i9/accuracy.json"))["results"]
    acc["nonpow2"] = {
        eng: {n: {"fwd_l2": v["fwd_l2"], "bwd_l2": v["bwd_l2"]} for n, v in tab.items()}
        for eng, tab in b.items() if eng.startswith("lean")
    }
    acc["xcheck"] = {
        str(d["N"]): {"lean_l2": d["lean_l2"], "lean_R": d["lean_R"]}
        for d in jsonl(f"{REF}/repro-2026-11-04/xcheck-patient.jsonl")
    }
    json.dump(acc, open(f"{HERE}/accuracy_ref.json", "w"), indent=1)
    print("fixtures written")


if __name__ == "__main__":
    main()
