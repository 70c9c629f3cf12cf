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
                rocprofv3 counter summary, when present;
  int_roofline  the integer co-bound: algorithmic base products per iteration
                (49.5 h per field) per second, and the dominant kernel's VALU
                busy fraction per SIMD from the committed counter summary;
  cpu_baseline  the single-threaded CPU restatement of the reference's
                algorithm (oracle/) timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
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


def cpu_baseline(p: int, mode: str, fixture_case, iters: int):
    """Times the oracle (CPU restatement, 1 thread) for `iters` iterations and
    checks its Res64 against the fixture when the fixture has that checkpoint."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import Oracle
    o = Oracle(so)
    t = o.L.m2o_test_new(p, 1 if mode == "prp" else 0, 64)
    try:
        o.L.m2o_test_run(t, 1)  # first iteration outside the sample (cache warm-up)
        t0 = time.perf_counter()
        o.L.m2o_test_run(t, iters)
        dt = time.perf_counter() - t0
        res = "%016X" % o.L.m2o_test_res64(t)
    finally:
        o.L.m2o_test_free(t)
    ok = None
    if fixture_case and str(iters + 1) in fixture_case["checkpoints"]:
        ok = fixture_case["checkpoints"][str(iters + 1)]["res64"] == res
    return {"value": dt / iters * 1e6, "unit": "µs/iter", "cores": 1, "kind": "port",
            "sample": "oracle/m2_oracle.c (CPU restatement of the reference's radix-4 IBDWT, 1 thread), "
                      "iterations 2..%d of the same %s test on this host; Res64 at %d %s" % (
                          iters + 1, mode.upper(), iters + 1,
                          {True: "matches", False: "MISMATCH", None: "unchecked"}[ok])}


# Workloads: exponent range, mode, exact fixture and its checked iterations.
# The default (the bench line) is the plain PRP test at n = 2^20; the n = 2^24
# LL test is measured the same way on request.
WORKLOADS = {
    "prp-n2p20": {"range": "C2_RANGE", "mode": "prp", "fixture": "c2.json", "checks": (500, 1000, 2000),
                  "steps": 100000, "cpu_iters": 9, "traffic_key": "c2"},
    "ll-n2p24": {"range": "C3_RANGE", "mode": "ll", "fixture": "c3.json", "checks": (100, 1000),
                 "steps": 5000, "cpu_iters": 2, "traffic_key": "c3"},
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="prp-n2p20")
    ap.add_argument("--steps", type=int, default=None, help="timed iterations (default per workload)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--kernel-iters", type=int, default=200, help="iterations timed kernel by kernel")
    ap.add_argument("--cpu-iters", type=int, default=None, help="oracle iterations in the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device 3 times (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    wl = WORKLOADS[args.workload]
    if args.steps is None:
        args.steps = wl["steps"]
    if args.cpu_iters is None:
        args.cpu_iters = wl["cpu_iters"]

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
    t.close()
    ctx.close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, mode, fixture, args.cpu_iters)
        log("cpu baseline:", cpu)

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
                     "bytes_per_launch": bytes_dom, "us_per_launch": round(per_kernel_us[dom], 3)},
        "int_roofline": {"modmuls_per_iter": int(modmuls),
                         "achieved_gmodmul_s": round(modmuls / (us_per_iter * 1e-6) / 1e9, 1),
                         "kernel": dom, "valu_busy": load_valu(dom, wl["traffic_key"]),
                         "valu_busy_source": "profiles/valu.json: rocprofv3 SQ counters, VALU instructions "
                                             "per SIMD per 4-cycle issue slot"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0 if parity_ok else 1


if __name__ == "__main__":
    sys.exit(main())
