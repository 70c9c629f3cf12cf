"""ORACLE — test infrastructure only.

Runs the reference program (through oracle/run_reference.py) on the
fixture cases of tests/golden and compares its Res64 and residue hash with the
exact values, one JSON line per case, then a summary line.

    python oracle/compare_reference.py              # the default case list
    python oracle/compare_reference.py --quick      # the n <= 2^17 cases only

The cases: whole PRP/LL tests of the known-answer fixtures with p > 8
(kat.json) and config 1 (c1.json, n = 512), one exponent per length 2^3..2^17 at its
fixture checkpoints (sweep.json), one per length 2^18..2^23 (mid.json), the
n = 2^20 PRP checkpoints (c2.json: 500, 1000, 2000) and the n = 2^24 LL
checkpoint at 100 (c3.json).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(os.path.dirname(HERE), "tests", "golden")
sys.path.insert(0, HERE)

import run_reference  # noqa: E402


def cases(quick: bool):
    def load(name):
        with open(os.path.join(GOLDEN, name)) as f:
            return json.load(f)["cases"]
    out = []
    # exponents from 8 up (the reference's carry shifts a word by 64 for
    # the 0-bit digits of p < 8, mersenne2_gpu.cpp:1982-1993)
    for c in [c for c in load("kat.json") if c["p"] > 8] + load("c1.json"):
        k = max(int(x) for x in c["checkpoints"])
        out.append((c["p"], c["mode"], k, c["checkpoints"][str(k)]))
    for c in load("sweep.json"):
        if c["p"] > 7 and 3 <= c["ln"] <= 17:
            k = max(int(x) for x in c["checkpoints"])
            out.append((c["p"], c["mode"], k, c["checkpoints"][str(k)]))
    if quick:
        return out
    for c in load("mid.json"):
        k = min(int(x) for x in c["checkpoints"])
        out.append((c["p"], c["mode"], k, c["checkpoints"][str(k)]))
    for c in load("c2.json"):
        for k in sorted(int(x) for x in c["checkpoints"]):
            out.append((c["p"], c["mode"], k, c["checkpoints"][str(k)]))
    for c in load("c3.json"):
        out.append((c["p"], c["mode"], 100, c["checkpoints"]["100"]))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args(argv)
    total = match = 0
    for p, mode, k, want in cases(a.quick):
        full = k == (p if mode == "prp" else p - 2)
        r = run_reference.run(p, mode, 0 if full else k)
        ok = r["res64"] == want["res64"] and ("sha256" not in want or r["sha256"] == want["sha256"])
        total += 1
        match += ok
        print(json.dumps(dict(r, want_res64=want["res64"], match=ok)), flush=True)
    print(json.dumps({"cases": total, "match": match, "mismatch": total - match}), flush=True)
    return 0 if match == total else 1


if __name__ == "__main__":
    sys.exit(main())
