"""Writes the golden Res64 fixtures under tests/golden/.

Every value comes from exact integer arithmetic (tests/golden/exact.py: GMP or
Python integers), never from the transform code under test. Usage:
    python tests/golden/gen_fixtures.py kat sweep c1 c4      # seconds to a minute
    python tests/golden/gen_fixtures.py c2                   # a few minutes (GMP)
    python tests/golden/gen_fixtures.py c3                   # ~20-30 minutes (GMP)
    python tests/golden/gen_fixtures.py c5                   # a few minutes (GMP)
    python tests/golden/gen_fixtures.py c5big                # ~10 minutes on 8 cores (GMP)
    python tests/golden/gen_fixtures.py mid                  # a few minutes (GMP)
"""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from exact import run  # noqa: E402
from mersenne2gpu_amd import workloads as W  # noqa: E402


def total_iters(p: int, mode: str) -> int:
    # PRP: p squarings; LL: p - 2 steps (mersenne2_gpu.cpp:4780)
    return p if mode == "prp" else max(p - 2, 0)


def case(p, mode, ks, keep=()):
    t0 = time.time()
    res = run(p, mode, ks, keep)
    print("  p=%d %s k=%s (%.1fs)" % (p, mode, ",".join(map(str, sorted(res))), time.time() - t0), flush=True)
    return {"p": p, "mode": mode, "ln": W.transformsize(p),
            "checkpoints": {str(k): v for k, v in sorted(res.items())}}


def write(name, obj):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", path)


def gen_kat():
    # Verdict known answers of the reference's own self-tests
    # (gf61...fused.cpp:536-539, ..._powonly_blockcarry.cpp:1319-1322) and the
    # README's example exponents (README.md:55-58), plus small LL cases.
    cases = []
    for p in (3, 5, 7, 13, 17, 19, 31, 61, 89, 107, 127, 11, 23, 29, 521, 607, 1279, 2203, 2281, 9941):
        for mode in ("prp", "ll"):
            n = total_iters(p, mode)
            ks = sorted(set([k for k in (1, 2, 10, 100) if k <= n] + [n]))
            c = case(p, mode, ks, keep=[n])
            c["final"] = n
            cases.append(c)
    write("kat.json", {"desc": "small exponents, full tests; Res64 + full residue at the end", "cases": cases})


def gen_sweep():
    # Both ends of the exponent range of every transform length from 2^3 to 2^17.
    cases = []
    for ln in range(3, 18):
        lo, hi = W.exponent_range(ln)
        for p in sorted(set((max(lo, 3), hi))):
            for mode in ("prp", "ll"):
                n = total_iters(p, mode)
                ks = [k for k in (1, 5, 50, 200) if k <= n]
                cases.append(case(p, mode, ks))
    write("sweep.json", {"desc": "Res64 at 1/5/50/200 iterations at both ends of every N from 8 to 2^17",
                         "cases": cases})


def gen_c1():
    # Plain PRP test at n = 512: seeded p in [10499, 20983], plus 19937 (prime)
    # and 20983 (composite); full test with the full residue at the end.
    ps = [W.draw_uniform(s, *W.C1_RANGE) for s in (1, 2, 3)] + [19937, 20983]
    cases = []
    for p in ps:
        n = total_iters(p, "prp")
        c = case(p, "prp", [100, 500, 1000, 2000, n], keep=[2000, n])
        c["final"] = n
        cases.append(c)
    c = case(19937, "ll", [100, 1000, 19935], keep=[19935])
    c["final"] = 19935
    cases.append(c)
    write("c1.json", {"desc": "n=512 plain PRP full tests (seeds 1-3, 19937, 20983) and LL 19937",
                      "seeds": [1, 2, 3], "cases": cases})


def gen_c2():
    p = W.draw_uniform(1, *W.C2_RANGE)
    write("c2.json", {"desc": "n=2^20 PRP, seed 1: Res64 and SHA-256 of the residue at 500/1000/2000",
                      "seed": 1, "cases": [case(p, "prp", [1, 2, 10, 100, 500, 1000, 2000])]})


def gen_c3():
    p = W.draw_uniform(1, *W.C3_RANGE)
    write("c3.json", {"desc": "n=2^24 LL, seed 1: Res64 and SHA-256 of the residue at 100/1000",
                      "seed": 1, "cases": [case(p, "ll", [1, 2, 10, 100, 1000])]})


def gen_mid():
    # one exponent per transform length 2^18 .. 2^23 (the middle of its range),
    # PRP and LL alternately
    cases = []
    for ln in range(18, 24):
        lo, hi = W.exponent_range(ln)
        p = W.prime_near((lo + hi) // 2, lo, hi)
        cases.append(case(p, "prp" if ln % 2 == 0 else "ll", [2, 10, 50]))
    write("mid.json", {"desc": "one exponent per length n = 2^18..2^23: Res64 at 2/10/50", "cases": cases})


def gen_c4():
    # 512 log-uniform exponents over n = 2^8..2^16 (seed 1). Every test gets its
    # Res64 at iteration min(p, 2000); tests with p <= 60000 also get the final Res64.
    ps = W.batch_exponents(1)
    cases = []
    for p in ps:
        n = total_iters(p, "prp")
        ks = sorted(set([min(n, 2000)] + ([n] if p <= 60000 else [])))
        c = case(p, "prp", ks)
        c["final"] = n
        cases.append(c)
    write("c4.json", {"desc": "512 batched PRP tests, seed 1, log-uniform over n=2^8..2^16",
                      "seed": 1, "cases": cases})


def gen_c5():
    # Serving stream (seed 1, 2000 requests, 2000-iteration budget): exact Res64
    # of every request with n <= 2^14 among the first 400, and of the first 6
    # requests with n = 2^16.
    reqs = W.serve_stream(1)
    cases = []
    seen = set()
    n16 = 0
    for i, (_, p, mode, k) in enumerate(reqs):
        ln = W.transformsize(p)
        take = (i < 400 and ln <= 14) or (ln == 16 and n16 < 6)
        if not take or (p, mode, k) in seen:
            continue
        n16 += ln == 16
        seen.add((p, mode, k))
        cases.append(case(p, mode, [k]))
    write("c5.json", {"desc": "serving stream seed 1: exact Res64 of requests with n <= 2^14 among the first "
                              "400, and of the first six with n = 2^16, at their iteration budget",
                      "seed": 1, "cases": cases})


# Requests with n >= 2^17 in the serving test stream run this many iterations
# (the rest keep the stream's 2000), so that exact values stay cheap to make.
C5_BIG_LN, C5_BIG_ITERS = 17, 100


def _case_args(a):
    return case(*a)


def gen_c5big():
    # Adds to c5.json the exact Res64 of every request with n >= 2^17 among the
    # first 400 of the serving stream (seed 1), at C5_BIG_ITERS iterations.
    import multiprocessing
    reqs = W.serve_stream(1)
    todo, seen = [], set()
    for i, (_, p, mode, _) in enumerate(reqs[:400]):
        if W.transformsize(p) >= C5_BIG_LN and (p, mode) not in seen:
            seen.add((p, mode))
            todo.append((p, mode, [C5_BIG_ITERS]))
    with multiprocessing.Pool(min(8, os.cpu_count() or 1)) as pool:
        big = pool.map(_case_args, todo, chunksize=1)
    path = os.path.join(HERE, "c5.json")
    with open(path) as f:
        fx = json.load(f)
    have = {(c["p"], c["mode"]) for c in fx["cases"]}
    fx["cases"] += [c for c in big if (c["p"], c["mode"]) not in have]
    fx["big_ln"], fx["big_iters"] = C5_BIG_LN, C5_BIG_ITERS
    fx["desc"] = ("serving stream seed 1: exact Res64 of requests with n <= 2^14 among the first 400, and of the "
                  "first six with n = 2^16, at their iteration budget; of every request with n >= 2^17 among "
                  "the first 400 at %d iterations" % C5_BIG_ITERS)
    write("c5.json", fx)


if __name__ == "__main__":
    todo = sys.argv[1:] or ["kat", "sweep", "c1", "c4"]
    for name in todo:
        print("==", name, flush=True)
        globals()["gen_" + name]()
