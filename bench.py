"""Benchmark: PRP base-3 IBDWT squaring at n = 2^20 on one MI355X.

A "step" is one squaring iteration of one test (x -> x^2 mod 2^p - 1) over the
whole residue, p drawn from seed 1 in [18874379, 36700159] (p = 27899687).
The run first checks Res64 at iterations 500/1000/2000 against the exact
fixture (tests/golden/c2.json), then times `--steps` more iterations after
`--warmup` untimed ones, with the residue resident in HBM throughout.
`--workload ll-n2p24` measures the n = 2^24 Lucas-Lehmer test the same way
(p from seed 1 in [285212677, 553648103], Res64 checked at 100 and 1000).

Prints one JSON line. Extra fields:
  roofline      dominant kernel's algorithmic HBM bytes per launch / its mean
                duration (HIP events on the test's stream) against 8 TB/s;
                traffic = measured HBM bytes per launch from the committed
                rocprofv3 counter summary named in traffic_source;
                traffic_sources_match says whether the kernel sources
                that summary was collected on (a digest it records) are
                the ones this run was built from;
  int_roofline  the integer co-bound: the iteration's algorithmic base products
                (49.5 h per field; 4 v_mad_u64_u32 per M61 product, 1 per M31
                product) per second against the device's v_mad_u64_u32 issue
                rate measured in the same run, and the dominant kernel's VALU
                busy fraction per SIMD from the committed counter summary
                (valu_busy_sources_match: the same digest check);
  config1       the n = 512 workload: whole PRP tests (the seeded exponent
                in [10499, 20983] and 19937) to completion, one at a time,
                final Res64 against the exact fixture, seconds per test;
  config3       the n = 2^24 workload: Lucas-Lehmer, Res64 at 100 and 1000
                against the exact fixture, then µs/iter and each kernel's
                mean duration, the dominant one against HBM peak;
  batched       the metric's batched half: 512 PRP tests (n = 2^8..2^16,
                seed 1) for 2000 iterations each in one m2_batch, tests/s
                against the plain form (one m2_test after another) measured
                on a seeded sample; every Res64 checked;
  served        the metric's served half: the first 400 requests of the
                serving stream (Poisson arrivals, 70% PRP, n = 2^10..2^22,
                2000 iterations each) through m2_server in real time, p50/p99
                completion latency and tests/s against a measured FIFO run
                of the same requests one at a time; every Res64 the same
                alone and in the mix;
  cpu_baseline  the exact CPU restatement of the reference's algorithm
                (oracle/) on all host threads, with the one-thread figure
                beside it (the reference's own OpenCL path needs a CPU OpenCL
                device, which the GPU host does not have).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "µs/iter + HBM GB/s at N=2^20 PRP; batched/served tests/s, p99 vs plain"
HBM_PEAK_GBPS = 8000.0
ELEM_BYTES = 24  # one element of GF(M61^2) x GF(M31^2)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel: str, h: int) -> int:
    """HBM bytes a kernel must move per launch: the 24-byte elements it reads
    and writes once (digits are 16 bytes per element)."""
    per_elem = {"k_centre": 2 * ELEM_BYTES, "k_mid_fwd": 2 * ELEM_BYTES, "k_mid_inv": 2 * ELEM_BYTES,
                "k_outer_fwd": 16 + ELEM_BYTES, "k_outer_inv": ELEM_BYTES + 16, "k_outer_fused": 2 * ELEM_BYTES,
                "k_small": 0}
    return per_elem.get(kernel, 2 * ELEM_BYTES) * h


def load_traffic(kernel: str, workload: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(workload, {}).get(kernel)
    except Exception:
        return None


def load_valu(kernel: str, workload: str):
    """VALU busy fraction per SIMD of a kernel, from the committed rocprofv3
    shader-counter summary (profiles/valu.json), when present."""
    path = os.path.join(ROOT, "profiles", "valu.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get(workload, {}).get(kernel)
    except Exception:
        return None


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"nproc": affinity, "cpu_model": model, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(p: int, mode: str, iters_mt: int, iters_1: int, gpu_res64):
    """Times the oracle (the exact CPU restatement of the reference's
    algorithm) on all host threads for `iters_mt` iterations, then on one
    thread for `iters_1` more, and checks its Res64 against the GPU's at the
    same iteration (gpu_res64(k))."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import Oracle
    o = Oracle(so)
    threads = o.L.m2o_set_threads(0)
    t = o.L.m2o_test_new(p, 1 if mode == "prp" else 0, 64)
    try:
        o.L.m2o_test_run(t, 1)  # first iteration outside the sample (page-in, cache warm-up)
        t0 = time.perf_counter()
        o.L.m2o_test_run(t, iters_mt)
        dt_mt = time.perf_counter() - t0
        o.L.m2o_set_threads(1)
        t0 = time.perf_counter()
        o.L.m2o_test_run(t, iters_1)
        dt_1 = time.perf_counter() - t0
        o.L.m2o_set_threads(threads)
        k = 1 + iters_mt + iters_1
        res = "%016X" % o.L.m2o_test_res64(t)
    finally:
        o.L.m2o_test_free(t)
    want = gpu_res64(k)
    host = host_info()
    return {"value": round(dt_mt / iters_mt * 1e6, 1), "unit": "µs/iter", "cores": threads, "kind": "port",
            "sample": "oracle/m2_oracle.c (exact CPU restatement of the reference's radix-4 IBDWT, OpenMP over "
                      "each stage's independent butterflies), iterations 2..%d of the same %s test on %d host "
                      "threads (%.1f s), then %d..%d on 1 thread; Res64 at %d %s the GPU's" % (
                          1 + iters_mt, mode.upper(), threads, dt_mt, 2 + iters_mt, k, k,
                          "matches" if res == want else "DIFFERS FROM"),
            "single_thread": {"value": round(dt_1 / iters_1 * 1e6, 1), "unit": "µs/iter", "cores": 1},
            "host": host,
            "reason": "the reference's OpenCL path needs a CPU OpenCL device (PoCL), which the GPU host does not "
                      "have, and it enumerates GPU devices only (mersenne2_gpu.cpp:3743-3752); this is the exact "
                      "restatement of the same algorithm on the same host cores",
            "res64_match": res == want}


def batched_leg(ctx, seed: int, cap: int, sample: int):
    """512 PRP tests (exponents log-uniform over n = 2^8..2^16) for
    min(p, cap) iterations each: all in one m2_batch, against the plain form
    (one m2_test after another, each created, run and read alone) timed on a
    seeded sample of the same tests. Wall times include creating the tests
    and reading every Res64 in both forms."""
    import random
    import mersenne2gpu_amd as m2
    from mersenne2gpu_amd import workloads as W
    ps = W.batch_exponents(seed, 512)
    stops = [min(p, cap) for p in ps]
    t0 = time.perf_counter()
    b = m2.Batch(ctx, ps, "prp", iters=stops)
    b.run(0)
    b.sync()
    got = [b.res64(i) for i in range(len(ps))]
    t_batched = time.perf_counter() - t0
    b.close()
    idx = sorted(random.Random(seed).sample(range(len(ps)), sample))
    t0 = time.perf_counter()
    plain_diff = 0
    for i in idx:
        with m2.Test(ctx, ps[i], "prp") as t:
            t.run(stops[i])
            plain_diff += t.res64() != got[i]
    t_plain = time.perf_counter() - t0
    checked = mismatched = 0
    fx = _fixture("c4.json")
    if fx and fx.get("seed") == seed:
        byp = {c["p"]: c for c in fx["cases"]}
        for i, p in enumerate(ps):
            c = byp.get(p)
            if c and str(stops[i]) in c["checkpoints"]:
                checked += 1
                mismatched += got[i] != c["checkpoints"][str(stops[i])]["res64"]
    lns = [W.transformsize(p) for p in ps]
    batched_tps = len(ps) / t_batched
    plain_tps = len(idx) / t_plain
    log("batched: %d tests in %.3f s (%.1f tests/s); plain sample %d in %.3f s (%.2f tests/s)" % (
        len(ps), t_batched, batched_tps, len(idx), t_plain, plain_tps))
    return {"tests_per_s": round(batched_tps, 2), "plain_tests_per_s": round(plain_tps, 3),
            "vs_plain": round(batched_tps / plain_tps, 2), "batched_s": round(t_batched, 3),
            "plain_sample_s": round(t_plain, 3), "plain_sample": len(idx),
            "res64_checked_fixture": checked, "res64_mismatched_fixture": mismatched,
            "plain_res64_mismatched": plain_diff,
            "workload": "512 PRP tests, seed %d, log-uniform over n=2^8..2^16 (%s), min(p, %d) iterations each; "
                        "plain = one m2_test per test, timed on a seeded sample of %d" % (
                            seed, ",".join("2^%d:%d" % (ln, lns.count(ln)) for ln in sorted(set(lns))), cap,
                            len(idx))}


def whole_test_leg(ctx, seed: int):
    """Config 1: plain PRP tests at n = 512 run to completion, the seeded
    exponent from [10499, 20983] and p = 19937 (a Mersenne prime), each
    created, run for its p squarings and read (Res64 and verdict) alone;
    every final Res64 checked against the exact fixture."""
    import mersenne2gpu_amd as m2
    from mersenne2gpu_amd import workloads as W
    fx = _fixture("c1.json") or {"cases": []}
    want = {(c["p"], c["mode"]): c["checkpoints"][str(c["final"])]["res64"] for c in fx["cases"]}
    ps = [W.draw_uniform(seed, *W.C1_RANGE), 19937]
    cases = []
    for p in ps:
        with m2.Test(ctx, p, "prp") as w:  # warm: first launch of this length's kernels
            w.run(16)
            w.sync()
        t0 = time.perf_counter()
        with m2.Test(ctx, p, "prp") as t:
            t.run(p)
            r64 = t.res64()
            prime = t.is_prime()
        dt = time.perf_counter() - t0
        cases.append({"p": p, "iterations": p, "s": round(dt, 4), "us_per_iter": round(dt / p * 1e6, 3),
                      "res64": r64, "match": r64 == want.get((p, "prp")),
                      "verdict": "probable prime" if prime else "composite"})
    log("config 1: %s" % json.dumps(cases))
    return {"workload": "plain PRP tests to completion at n=512 (p from seed %d in [10499, 20983], and 19937); "
                        "wall time of create + p squarings + Res64/verdict read, one test at a time" % seed,
            "cases": cases, "ok": all(c["match"] for c in cases)}


def large_test_leg(ctx, seed: int, steps: int):
    """Config 3: the Lucas-Lehmer test at n = 2^24 (p from seed in
    [285212677, 553648103]); Res64 at iterations 100 and 1000 against the
    exact fixture, then `steps` more iterations timed with hipEvents and each
    kernel's mean duration, the dominant one priced against HBM peak."""
    import mersenne2gpu_amd as m2
    from mersenne2gpu_amd import workloads as W
    p = W.draw_uniform(seed, *W.C3_RANGE)
    ln = W.transformsize(p)
    h = 1 << (ln - 1)
    fx = _fixture("c3.json") or {"cases": []}
    want = {}
    for c in fx["cases"]:
        if c["p"] == p and c["mode"] == "ll":
            want = {int(k): v["res64"] for k, v in c["checkpoints"].items()}
    parity = {}
    with m2.Test(ctx, p, "ll") as t:
        done = 0
        for k in (100, 1000):
            t.run(k - done)
            done = k
            got = t.res64()
            parity[str(k)] = {"res64": got, "match": got == want.get(k)}
        ms = t.time(steps)
        kt = {k: v * 1000.0 for k, v in t.kernel_times(10).items()}
    us = ms * 1000.0 / steps
    dom = max(kt, key=kt.get)
    gbps = algorithmic_bytes(dom, h) / (kt[dom] * 1e-6) / 1e9
    out = {"workload": "plain single LL test, p=%d (seed %d), n=2^%d, %d iterations timed after 1000 checked" % (
               p, seed, ln, steps),
           "us_per_iter": round(us, 3), "hbm_gbps_algorithmic": round(48.0 * 2 * h / (us * 1e-6) / 1e9, 1),
           "per_kernel_us": {k: round(v, 3) for k, v in kt.items()},
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS,
                        "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4),
                        "bytes_per_launch": algorithmic_bytes(dom, h), "us_per_launch": round(kt[dom], 3)},
           "parity": parity, "ok": all(v["match"] for v in parity.values())}
    log("config 3: %s" % json.dumps(out))
    return out


def _pct(xs, q):
    xs = sorted(xs)
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def served_leg(ctx, seed: int, count: int, iters: int, budget_gb: float, slice_iters: int):
    """The first `count` requests of the serving stream, submitted at their
    arrival times to m2_server (memory budget, time slices), against the
    plain form measured the same way: the same requests served one at a
    time in arrival order (FIFO), each an m2_test created, run and read
    alone. Completion latency = completion - arrival."""
    import mersenne2gpu_amd as m2
    from mersenne2gpu_amd import workloads as W
    reqs = W.serve_stream(seed, 2000, 40.0, iters)[:count]
    srv = m2.Server(ctx, hbm_budget_bytes=int(budget_gb * 1e9), slice_iters=slice_iters)
    ids, done = {}, {}
    nxt = 0
    peak_active = 0
    t0 = time.perf_counter()
    while len(done) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt][0] <= now:
            _, p, mode, k = reqs[nxt]
            ids[srv.submit(p, mode, k)] = nxt
            nxt += 1
        busy = srv.step()
        tnow = time.perf_counter() - t0
        for rid, r64, _ in srv.poll():
            done[ids[rid]] = (tnow, r64)
        peak_active = max(peak_active, srv.stats()["active"])
        if not busy and nxt < len(reqs):
            wait = reqs[nxt][0] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.005))
    makespan = time.perf_counter() - t0
    srv.close()
    lat = [done[i][0] - reqs[i][0] for i in range(len(reqs))]
    # capacity: the same requests all submitted at once
    srv = m2.Server(ctx, hbm_budget_bytes=int(budget_gb * 1e9), slice_iters=slice_iters)
    t0 = time.perf_counter()
    cap_ids = {srv.submit(p, mode, k): i for i, (_, p, mode, k) in enumerate(reqs)}
    cap_r64 = {}
    while True:
        busy = srv.step()
        for rid, r64, _ in srv.poll():
            cap_r64[cap_ids[rid]] = r64
        if not busy and len(cap_r64) == len(reqs):
            break
        if time.perf_counter() - t0 > 600:
            raise RuntimeError("served capacity run did not finish")
    cap_makespan = time.perf_counter() - t0
    srv.close()
    # plain: FIFO, one m2_test at a time, in real time; each test's own
    # service time (create, run, read) is kept for the capacity comparison
    plain_lat, plain_r64, plain_service = [], [], []
    t0 = time.perf_counter()
    for arr, p, mode, k in reqs:
        wait = arr - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        s0 = time.perf_counter()
        with m2.Test(ctx, p, mode) as t:
            t.run(k)
            plain_r64.append(t.res64())
        plain_service.append(time.perf_counter() - s0)
        plain_lat.append(time.perf_counter() - t0 - arr)
    plain_makespan = time.perf_counter() - t0
    differ = sum(plain_r64[i] != done[i][1] or plain_r64[i] != cap_r64.get(i) for i in range(len(reqs)))
    checked = mismatched = 0
    fx = _fixture("c5.json")
    if fx and fx.get("seed") == seed:
        want = {(c["p"], c["mode"], int(k)): v["res64"] for c in fx["cases"] for k, v in c["checkpoints"].items()}
        for i, (_, p, mode, k) in enumerate(reqs):
            if (p, mode, k) in want:
                checked += 1
                mismatched += done[i][1] != want[(p, mode, k)]
    lns = [W.transformsize(p) for _, p, _, _ in reqs]
    out = {"tests_per_s": round(len(reqs) / makespan, 2), "p50_s": round(_pct(lat, 50), 4),
           "p99_s": round(_pct(lat, 99), 4), "makespan_s": round(makespan, 3), "peak_active": peak_active,
           "plain": {"tests_per_s": round(len(reqs) / plain_makespan, 2), "p50_s": round(_pct(plain_lat, 50), 4),
                     "p99_s": round(_pct(plain_lat, 99), 4), "makespan_s": round(plain_makespan, 3),
                     "how": "measured: FIFO in arrival order, one m2_test at a time"},
           "tests_per_s_vs_plain": round(plain_makespan / makespan, 2),
           "p99_vs_plain": round(_pct(plain_lat, 99) / _pct(lat, 99), 2),
           "capacity": {"tests_per_s": round(len(reqs) / cap_makespan, 2),
                        "plain_tests_per_s": round(len(reqs) / sum(plain_service), 2),
                        "vs_plain": round(sum(plain_service) / cap_makespan, 2),
                        "how": "all requests submitted at once; plain = the same tests back to back (sum of the "
                               "FIFO run's measured service times)"},
           "res64_same_alone_and_in_mix": differ == 0, "res64_differ": differ,
           "res64_checked_fixture": checked, "res64_mismatched_fixture": mismatched,
           "workload": "first %d requests of the seed-%d serving stream (Poisson 40/s over %.1f s, 70%% PRP; %s), "
                       "%d iterations each, %.0f GB budget, %d-iteration slices" % (
                           len(reqs), seed, reqs[-1][0],
                           ",".join("2^%d:%d" % (ln, lns.count(ln)) for ln in sorted(set(lns))), iters, budget_gb,
                           slice_iters)}
    log("served: %s" % json.dumps(out))
    return out


# the sources whose change can change a kernel's counters: the device code and
# the host code that plans launches and lays out tables
KERNEL_SOURCES = ("m2_kernels.hip", "m2_arith.h", "m2_digits.h", "m2_plan.h", "m2_internal.h", "m2_host.cpp",
                  "m2_api.cpp")


def kernel_sources_digest():
    """SHA-256 (first 16 hex digits) over the kernel and planner sources, as
    recorded in the committed counter summaries' `_sources_sha256`."""
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "mersenne2gpu_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def _profile_matches(name):
    """True when a committed counter summary was collected on the sources this
    run was built from, False when they differ, None when it records none."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            rec = json.load(f).get("_sources_sha256")
        return None if rec is None else rec == kernel_sources_digest()
    except (OSError, ValueError):
        return None


def _profile_source(name):
    """The `_source` note of a committed counter summary (which profile and
    build its figures come from: they are not measured by this run)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f).get("_source")
    except (OSError, ValueError):
        return None


def _fixture(name):
    path = os.path.join(ROOT, "tests", "golden", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


# Workloads: exponent range, mode, exact fixture and its checked iterations.
# The default (the bench line) is the plain PRP test at n = 2^20; the n = 2^24
# LL test is measured the same way on request.
WORKLOADS = {
    "prp-n2p20": {"range": "C2_RANGE", "mode": "prp", "fixture": "c2.json", "checks": (500, 1000, 2000),
                  "steps": 100000, "cpu_iters": (100, 9), "traffic_key": "c2"},
    "ll-n2p24": {"range": "C3_RANGE", "mode": "ll", "fixture": "c3.json", "checks": (100, 1000),
                 "steps": 5000, "cpu_iters": (8, 1), "traffic_key": "c3"},
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="prp-n2p20")
    ap.add_argument("--steps", type=int, default=None, help="timed iterations (default per workload)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--kernel-iters", type=int, default=200, help="iterations timed kernel by kernel")
    ap.add_argument("--cpu-iters", type=int, default=None,
                    help="oracle iterations on all host threads in the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-served", action="store_true", help="skip the served leg")
    ap.add_argument("--no-config1", action="store_true", help="skip the n = 512 whole-test leg")
    ap.add_argument("--no-config3", action="store_true", help="skip the n = 2^24 Lucas-Lehmer leg")
    ap.add_argument("--config3-steps", type=int, default=1000, help="timed iterations of the n = 2^24 leg")
    ap.add_argument("--batch-cap", type=int, default=2000, help="iterations per batched test")
    ap.add_argument("--batch-plain-sample", type=int, default=64, help="tests in the measured plain sample")
    ap.add_argument("--serve-count", type=int, default=400, help="requests of the serving stream")
    ap.add_argument("--serve-iters", type=int, default=2000, help="iterations per served request")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device 3 times (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    wl = WORKLOADS[args.workload]
    if args.steps is None:
        args.steps = wl["steps"]
    cpu_iters = (args.cpu_iters, wl["cpu_iters"][1]) if args.cpu_iters else wl["cpu_iters"]

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs one test on one GPU; it has no multi-GPU mode "
              "(independent tests shard across GPUs with no communication)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: 3 x 256 MiB device-to-device")

    import mersenne2gpu_amd as m2
    from mersenne2gpu_amd import workloads as W

    mode = wl["mode"]
    p = W.draw_uniform(args.seed, *getattr(W, wl["range"]))
    ln = W.transformsize(p)
    n, h = 1 << ln, 1 << (ln - 1)
    fixture = None
    fx_path = os.path.join(ROOT, "tests", "golden", wl["fixture"])
    if os.path.exists(fx_path):
        with open(fx_path) as f:
            for c in json.load(f)["cases"]:
                if c["p"] == p and c["mode"] == mode:
                    fixture = c

    ctx = m2.Context(0)
    log("device:", ctx.name, ctx.props())
    t = m2.Test(ctx, p, mode)
    log(t.plan().rstrip())

    # parity gate at the fixture's checkpoints
    parity = {}
    done = 0
    for k in wl["checks"]:
        t.run(k - done)
        done = k
        got = t.res64()
        want = fixture["checkpoints"][str(k)]["res64"] if fixture and str(k) in fixture["checkpoints"] else None
        parity[str(k)] = {"res64": got, "match": (got == want) if want else None}
        log("Res64@%d %s %s" % (k, got, "ok" if want == got else ("MISMATCH want " + want if want else "")))
    parity_ok = all(v["match"] is not False for v in parity.values()) and fixture is not None

    # warm-up, then the timed region (device time via hipEvents, wall clock beside it)
    t.run(args.warmup)
    t.sync()
    w0 = time.perf_counter()
    ms = t.time(args.steps)
    wall = time.perf_counter() - w0
    us_per_iter = ms * 1000.0 / args.steps
    log("timed %d iterations: %.3f ms device, %.3f s wall -> %.3f us/iter" % (args.steps, ms, wall, us_per_iter))

    # per-kernel device time (each kernel's dispatch start/stop timestamps)
    kt = t.kernel_times(args.kernel_iters)
    per_kernel_us = {k: v * 1000.0 for k, v in kt.items()}  # mean per launch
    log("per-kernel us:", json.dumps({k: round(v, 3) for k, v in per_kernel_us.items()}))
    dom = max(per_kernel_us, key=per_kernel_us.get)
    bytes_dom = algorithmic_bytes(dom, h)
    achieved = bytes_dom / (per_kernel_us[dom] * 1e-6) / 1e9
    traffic = load_traffic(dom, wl["traffic_key"])
    whole_gbps = 48.0 * n / (us_per_iter * 1e-6) / 1e9
    # secondary (integer) roofline: the transform's algorithmic base products,
    # 49.5 h per field per iteration (forward and inverse 7.25 h complex
    # products each at radix-4 counting, 3 base products apiece, plus the
    # centre's 6 h), against the measured time
    modmuls = 2 * 49.5 * h
    # each M61 product is four 32x32->64 multiply-adds (v_mad_u64_u32), each
    # M31 product one: the integer roofline prices them at the device's
    # measured v_mad_u64_u32 issue rate
    products = 49.5 * h * (4 + 1)
    mad_rate = ctx.product_rate()
    int_us = products / mad_rate * 1e6
    log("v_mad_u64_u32 rate %.2f T lane-op/s -> product floor %.2f us/iter" % (mad_rate / 1e12, int_us))
    t.close()

    def gpu_res64(k):
        with m2.Test(ctx, p, mode) as tk:
            tk.run(k)
            return tk.res64()

    # the other workloads and legs belong to the headline (n = 2^20) line
    legs = args.workload == "prp-n2p20"
    config1 = None if (args.no_config1 or not legs) else whole_test_leg(ctx, args.seed)
    config3 = None if (args.no_config3 or not legs) else large_test_leg(ctx, args.seed, args.config3_steps)
    batched = None if (args.no_batched or not legs) else batched_leg(ctx, args.seed, args.batch_cap,
                                                                        args.batch_plain_sample)
    served = None if (args.no_served or not legs) else served_leg(ctx, args.seed, args.serve_count, args.serve_iters,
                                                                   200.0, 1000)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, mode, cpu_iters[0], cpu_iters[1], gpu_res64)
        log("cpu baseline:", cpu)
    ctx.close()
    traffic_src = _profile_source("traffic.json") if traffic is not None else None

    out = {
        "metric": METRIC,
        "value": round(us_per_iter, 4),
        "unit": "µs/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(us_per_iter / 1000.0, 7),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (%s; the only inputs are p and the mode)" % (
            "PRP base 3 from the residue 3" if mode == "prp" else "Lucas-Lehmer from 4"),
        "config": {"workload": "plain single %s test, p=%d (seed %d), n=2^%d, GF(M61^2) x GF(M31^2), "
                               "%d iterations timed after %d checked + %d warm-up" % (
                                   mode.upper(), p, args.seed, ln, args.steps, wl["checks"][-1], args.warmup),
                   "p": p, "n": n, "mode": mode, "seed": args.seed, "parallelism": "1 GPU, 1 test"},
        "iters_per_s": round(1e6 / us_per_iter, 1),
        "hbm_gbps_algorithmic": round(whole_gbps, 1),
        "hbm_frac_algorithmic": round(whole_gbps / HBM_PEAK_GBPS, 4),
        "parity": {"ok": parity_ok, "checkpoints": parity},
        "per_kernel_us": {k: round(v, 3) for k, v in per_kernel_us.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "traffic_sources_match": _profile_matches("traffic.json") if traffic is not None else None,
                     "bytes_per_launch": bytes_dom, "us_per_launch": round(per_kernel_us[dom], 3)},
        "int_roofline": {"bound": "v_mad_u64_u32 issue", "unit": "T products/s",
                         "achieved": round(products / (us_per_iter * 1e-6) / 1e12, 3),
                         "peak": round(mad_rate / 1e12, 3), "peak_how": "measured: m2_debug_product_rate",
                         "frac": round(int_us / us_per_iter, 4),
                         "products_per_iter": int(products), "product_floor_us": round(int_us, 3),
                         "modmuls_per_iter": int(modmuls),
                         "achieved_gmodmul_s": round(modmuls / (us_per_iter * 1e-6) / 1e9, 1),
                         "kernel": dom, "valu_busy": load_valu(dom, wl["traffic_key"]),
                         "valu_busy_source": _profile_source("valu.json"),
                         "valu_busy_sources_match": _profile_matches("valu.json")},
        "config1": config1,
        "config3": config3,
        "batched": batched,
        "served": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ok = parity_ok
    for leg in (config1, config3):
        if leg:
            ok = ok and leg["ok"]
    if batched:
        ok = ok and batched["res64_mismatched_fixture"] == 0 and batched["plain_res64_mismatched"] == 0
    if served:
        ok = ok and served["res64_differ"] == 0 and served["res64_mismatched_fixture"] == 0
    if cpu:
        ok = ok and cpu["res64_match"]
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
