"""Batched throughput: 512 independent PRP tests against the plain form.

The workload: 512 PRP base-3 tests whose exponents are drawn from a seed
log-uniformly over the exponent ranges of n = 2^8 .. 2^16 (5381 .. 2424827),
all resident at once and advanced together (m2_batch: tests of one transform
length share every launch, lengths run on their own streams, multi-pass
lengths replay captured graphs).

The plain form is one test per run, one after another, through the same
engine (m2_test). Its time is measured per transform length on one test and
multiplied out: a plain test of exponent p takes p iterations at its length's
per-iteration time, which does not depend on the iteration.

    python -m mersenne2gpu_amd.batch_bench --full      # every test to completion
    python -m mersenne2gpu_amd.batch_bench --cap 2000  # 2000 iterations of each test

Prints one JSON line. Res64 values are checked against tests/golden/c4.json
where it holds the stopping iteration (iteration 2000 of every test; the
final residue of every test with p <= 60000).
"""
from __future__ import annotations

import argparse
import collections
import json
import os
import sys
import time

import mersenne2gpu_amd as m2
from mersenne2gpu_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plain_us_per_iter(ctx, p: int, mode: str = "prp", target_s: float = 0.25) -> float:
    """Device µs per iteration of one test alone (m2_test, after warm-up)."""
    with m2.Test(ctx, p, mode) as t:
        t.run(64)
        t.sync()
        ms = t.time(64)
        k = int(max(64, min(20000, target_s * 1e3 / max(ms / 64, 1e-6))))
        ms = t.time(k)
    return ms * 1e3 / k


def plain_rates(ctx, lns, mode="prp"):
    out = {}
    for ln in sorted(set(lns)):
        p = W.exponent_range(ln)[0]
        out[ln] = plain_us_per_iter(ctx, p, mode)
        log("plain n=2^%d (p=%d): %.3f us/iter" % (ln, p, out[ln]))
    return out


def load_fixture(name):
    path = os.path.join(ROOT, "tests", "golden", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--full", action="store_true", help="run every test to completion")
    ap.add_argument("--cap", type=int, default=2000, help="iterations per test when not --full")
    ap.add_argument("--chunk", type=int, default=50000, help="iterations per test between progress lines")
    args = ap.parse_args(argv)

    ps = W.batch_exponents(args.seed, args.count)
    stops = [p if args.full else min(p, args.cap) for p in ps]
    lns = [W.transformsize(p) for p in ps]
    by_ln = collections.Counter(lns)
    log("batch: %d tests, lengths %s, %d iterations in total" % (
        len(ps), dict(sorted(by_ln.items())), sum(stops)))

    ctx = m2.Context(0)
    rates = plain_rates(ctx, lns)

    t0 = time.perf_counter()
    b = m2.Batch(ctx, ps, "prp", iters=stops)
    t_create = time.perf_counter() - t0
    # advanced in chunks of --chunk iterations per test, with a progress line
    # after each (the device stays busy: every chunk is enqueued before the wait)
    t0 = time.perf_counter()
    done_iters = 0
    while True:
        b.run(args.chunk)
        b.sync()
        done = sum(min(b.iteration(i), stops[i]) for i in range(len(ps)))
        finished = sum(b.iteration(i) >= stops[i] for i in range(len(ps)))
        log("  %.1f s: %d/%d tests finished, %.1f%% of iterations" % (
            time.perf_counter() - t0, finished, len(ps), 100.0 * done / sum(stops)))
        if done == done_iters or finished == len(ps):
            break
        done_iters = done
    t_run = time.perf_counter() - t0
    log("batched: create %.3f s, run %.3f s" % (t_create, t_run))

    # parity against the exact fixture where it holds the stopping iteration
    fx = load_fixture("c4.json")
    checked = mismatched = 0
    if fx and fx.get("seed") == args.seed:
        byp = {c["p"]: c for c in fx["cases"]}
        for i, p in enumerate(ps):
            c = byp.get(p)
            if c and str(stops[i]) in c["checkpoints"]:
                checked += 1
                if b.res64(i) != c["checkpoints"][str(stops[i])]["res64"]:
                    mismatched += 1
                    log("MISMATCH p=%d at %d" % (p, stops[i]))
    b.close()

    equiv = sum(s / p for s, p in zip(stops, ps))  # completed-test equivalents
    plain_s = sum(s * rates[ln] * 1e-6 for s, ln in zip(stops, lns))
    per_ln = {}
    for ln in sorted(by_ln):
        idx = [i for i in range(len(ps)) if lns[i] == ln]
        per_ln["2^%d" % ln] = {"tests": len(idx), "iterations": sum(stops[i] for i in idx),
                               "plain_us_per_iter": round(rates[ln], 3),
                               "plain_s": round(sum(stops[i] * rates[ln] * 1e-6 for i in idx), 3)}
    out = {
        "metric": "batched tests/s vs plain (512 PRP tests, n = 2^8..2^16)",
        "value": round(equiv / t_run, 3),
        "unit": "tests/s",
        "higher_is_better": True,
        "plain_tests_per_s": round(equiv / plain_s, 3),
        "speedup_vs_plain": round(plain_s / t_run, 2),
        "batched_s": round(t_run, 3),
        "plain_s": round(plain_s, 3),
        "test_equivalents": round(equiv, 3),
        "full": bool(args.full),
        "cap": None if args.full else args.cap,
        "res64_checked": checked,
        "res64_mismatched": mismatched,
        "per_length": per_ln,
        "config": {"workload": "512 PRP tests, exponents log-uniform over n=2^8..2^16, seed %d" % args.seed,
                   "seed": args.seed, "count": args.count},
    }
    print(json.dumps(out))
    ctx.close()
    return 0 if mismatched == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
