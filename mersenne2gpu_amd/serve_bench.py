"""Serving: a Poisson stream of LL/PRP requests through m2_server.

The stream (mersenne2gpu_amd.workloads.serve_stream): 2000 requests from a
seed, 70% PRP, exponents 80% in n = 2^10..2^14, 18% in 2^16..2^20, 2% at
n = 2^22, each with an iteration budget (min(budget, full test)). Requests
are submitted at their arrival times; the server admits them by device
footprint under the memory budget (200 GB), keeps every residue resident,
and runs each transform length's members in slices of 1000 iterations on
that length's stream, starting a new slice (with newly admitted tests)
whenever the previous one ends.

Reported: p50/p99 completion latency (completion - arrival) and tests/s,
beside the plain form: the same arrivals served one test at a time in
arrival order (a FIFO queue). With --plain measured (the default) that form
is run: each request, once it has arrived and the one before it is done, is
an m2_test created, run and read alone, timed by the wall clock; its tests/s
at capacity is the count over the sum of those service times. --plain model
instead gives each test its iterations times its length's measured
single-test µs/iteration.

Parity: every completion whose (p, mode, iterations) is in
tests/golden/c5.json is compared with the exact value; the measured plain
form re-runs every request alone and compares its Res64 with the in-mix one
(the model form re-runs a seeded sample).

    python -m mersenne2gpu_amd.serve_bench [--count 2000] [--rate 40] [--iters 2000] [--plain measured]
"""
from __future__ import annotations

import argparse
import json
import os
import random
import sys
import time

import mersenne2gpu_amd as m2
from mersenne2gpu_amd import workloads as W
from mersenne2gpu_amd.batch_bench import load_fixture, log, plain_rates


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def fifo_latencies(reqs, service_s):
    """Completion - arrival for one-at-a-time service in arrival order."""
    t = 0.0
    out = []
    for (arr, _, _, _), s in zip(reqs, service_s):
        t = max(t, arr) + s
        out.append(t - arr)
    return out, t


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--count", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=40.0, help="mean arrivals per second")
    ap.add_argument("--iters", type=int, default=2000, help="iteration budget per request (0: full tests)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--slice", type=int, default=1000)
    ap.add_argument("--alone-sample", type=int, default=48)
    ap.add_argument("--plain", choices=("measured", "model"), default="measured",
                    help="run the one-at-a-time form (measured) or price it from per-length rates (model)")
    args = ap.parse_args(argv)

    reqs = W.serve_stream(args.seed, args.count, args.rate, args.iters)
    lns = [W.transformsize(p) for _, p, _, _ in reqs]
    log("stream: %d requests over %.1f s, lengths %s" % (
        len(reqs), reqs[-1][0], {ln: lns.count(ln) for ln in sorted(set(lns))}))

    ctx = m2.Context(0)
    rates = plain_rates(ctx, lns) if args.plain == "model" else None

    srv = m2.Server(ctx, hbm_budget_bytes=int(args.budget_gb * 1e9), slice_iters=args.slice)
    ids = {}
    done = {}
    peak_active = peak_bytes = 0
    nxt = 0
    t0 = time.perf_counter()
    while len(done) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt][0] <= now:
            _, p, mode, k = reqs[nxt]
            ids[srv.submit(p, mode, k)] = nxt
            nxt += 1
        busy = srv.step()
        tnow = time.perf_counter() - t0
        for rid, r64, verdict in srv.poll():
            done[ids[rid]] = (tnow, r64, verdict)
        st = srv.stats()
        peak_active = max(peak_active, st["active"])
        peak_bytes = max(peak_bytes, st["bytes_in_use"])
        if not busy and nxt < len(reqs):
            wait = reqs[nxt][0] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.01))
    makespan = time.perf_counter() - t0
    srv.close()
    lat = [done[i][0] - reqs[i][0] for i in range(len(reqs))]
    log("served %d requests in %.3f s (peak %d active, %.2f GB)" % (len(reqs), makespan, peak_active,
                                                                   peak_bytes / 1e9))

    alone = None
    if args.plain == "measured":
        # plain form, run: FIFO in arrival order, one m2_test at a time
        plain_lat, service, alone = [], [], []
        t0 = time.perf_counter()
        for i, (arr, p, mode, k) in enumerate(reqs):
            wait = arr - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            s0 = time.perf_counter()
            with m2.Test(ctx, p, mode) as t:
                t.run(k)
                alone.append(t.res64())
            service.append(time.perf_counter() - s0)
            plain_lat.append(time.perf_counter() - t0 - arr)
            if (i + 1) % 250 == 0:
                log("plain form: %d of %d requests done at %.1f s" % (i + 1, len(reqs), time.perf_counter() - t0))
        plain_end = time.perf_counter() - t0
        log("plain form run: %d requests in %.3f s (service %.3f s)" % (len(reqs), plain_end, sum(service)))
    else:
        # plain form, priced: FIFO, one test at a time, measured per-length rates
        service = [k * rates[ln] * 1e-6 for (_, _, _, k), ln in zip(reqs, lns)]
        plain_lat, plain_end = fifo_latencies(reqs, service)

    # parity: exact fixture, and a sample re-run alone
    fx = load_fixture("c5.json")
    checked = mismatched = 0
    if fx and fx.get("seed") == args.seed:
        want = {(c["p"], c["mode"], int(k)): v["res64"] for c in fx["cases"] for k, v in c["checkpoints"].items()}
        for i, (_, p, mode, k) in enumerate(reqs):
            if (p, mode, k) in want:
                checked += 1
                if done[i][1] != want[(p, mode, k)]:
                    mismatched += 1
                    log("MISMATCH (fixture) p=%d %s k=%d" % (p, mode, k))
    alone_mismatch = 0
    if alone is not None:
        sample = list(range(len(reqs)))
        for i in sample:
            if alone[i] != done[i][1]:
                alone_mismatch += 1
                log("MISMATCH (alone) p=%d %s k=%d" % reqs[i][1:])
    else:
        rnd = random.Random(args.seed)
        sample = rnd.sample(range(len(reqs)), min(args.alone_sample, len(reqs)))
        for i in sample:
            _, p, mode, k = reqs[i]
            with m2.Test(ctx, p, mode) as t:
                t.run(k)
                if t.res64() != done[i][1]:
                    alone_mismatch += 1
                    log("MISMATCH (alone) p=%d %s k=%d" % (p, mode, k))
    ctx.close()

    out = {
        "metric": "served tests/s and p50/p99 completion latency vs plain (2000 Poisson LL/PRP requests)",
        "value": round(len(reqs) / makespan, 3),
        "unit": "tests/s",
        "higher_is_better": True,
        "p50_s": round(pct(lat, 50), 4),
        "p99_s": round(pct(lat, 99), 4),
        "makespan_s": round(makespan, 3),
        "plain": {"tests_per_s": round(len(reqs) / plain_end, 3), "p50_s": round(pct(plain_lat, 50), 4),
                  "p99_s": round(pct(plain_lat, 99), 4), "makespan_s": round(plain_end, 3),
                  "service_s": round(sum(service), 3),
                  "how": ("measured: FIFO in arrival order, one m2_test at a time" if alone is not None else
                          "FIFO one test at a time, measured single-test us/iter per length")},
        "p99_vs_plain": round(pct(plain_lat, 99) / pct(lat, 99), 2),
        "tests_per_s_vs_plain": round(plain_end / makespan, 2),
        "capacity_vs_plain_service": round(sum(service) / makespan, 2),
        "peak_active": peak_active,
        "peak_bytes": peak_bytes,
        "res64_checked_fixture": checked,
        "res64_mismatched_fixture": mismatched,
        "alone_sample": len(sample),
        "alone_mismatched": alone_mismatch,
        "config": {"workload": "%d Poisson requests at %.1f/s, 70%% PRP, n mix 80%% 2^10-2^14 / 18%% 2^16-2^20 / "
                               "2%% 2^22, %s iterations each, %.0f GB budget, %d-iteration slices" % (
                                   args.count, args.rate, args.iters or "full", args.budget_gb, args.slice),
                   "seed": args.seed},
    }
    print(json.dumps(out))
    return 0 if mismatched == 0 and alone_mismatch == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
