"""Batches and the server: each test's result must be the one it gets alone.

Expected values are the committed exact fixtures (tests/golden), which the
single-test parity suite already checks the engine against one exponent at a
time. Bit-exact Res64 and residue hashes, no tolerance.
"""
import hashlib
import re

import pytest

import mersenne2gpu_amd as m2
from conftest import load_fixture

pytestmark = pytest.mark.gpu


def _sweep_cases():
    return [c for c in load_fixture("sweep.json")["cases"] if c["p"] > 3]


def test_batch_sweep_mixed_lengths_checkpoints(gpu_ctx):
    """Every transform length from 2^4 to 2^17, LL and PRP, in one batch;
    advanced 1, 4, 45, 150 iterations at a time and checked at each stop."""
    cases = _sweep_cases()
    print("batch of %d sweep cases" % len(cases), flush=True)
    with m2.Batch(gpu_ctx, [c["p"] for c in cases], [c["mode"] for c in cases]) as b:
        done = 0
        for k in (1, 5, 50, 200):
            b.run(k - done)
            done = k
            for i, c in enumerate(cases):
                assert b.iteration(i) == k
                assert b.res64(i) == c["checkpoints"][str(k)]["res64"], (c["p"], c["mode"], k)


def test_batch_uneven_stops(gpu_ctx):
    """Tests stop at their own iteration counts inside one batch (the active
    set of a launch shrinks as tests finish)."""
    cases = _sweep_cases()
    stops = [(1, 5, 50, 200)[i % 4] for i in range(len(cases))]
    with m2.Batch(gpu_ctx, [c["p"] for c in cases], [c["mode"] for c in cases], iters=stops) as b:
        b.run(0)
        b.sync()
        for i, c in enumerate(cases):
            assert b.iteration(i) == stops[i]
            assert b.res64(i) == c["checkpoints"][str(stops[i])]["res64"], (c["p"], c["mode"], stops[i])


def test_batch_c4_512_tests(gpu_ctx):
    """The 512-test batch workload (log-uniform n = 2^8..2^16, PRP): Res64 and
    residue hash at iteration 2000 for all, and the full test (verdict and
    final residue) for every case with p <= 60000."""
    cases = load_fixture("c4.json")["cases"]
    stops = [c["final"] if c["p"] <= 60000 else 2000 for c in cases]
    print("c4 batch: %d tests, %d run to completion" % (len(cases), sum(s != 2000 for s in stops)), flush=True)
    with m2.Batch(gpu_ctx, [c["p"] for c in cases], [c["mode"] for c in cases], iters=stops) as b:
        b.run(2000)
        for i, c in enumerate(cases):
            want = c["checkpoints"]["2000"]
            assert b.res64(i) == want["res64"], (c["p"], 2000)
            assert hashlib.sha256(b.residue(i)).hexdigest() == want["sha256"], (c["p"], 2000)
        b.run(0)
        b.sync()
        for i, c in enumerate(cases):
            if stops[i] == 2000:
                continue
            want = c["checkpoints"][str(stops[i])]
            assert b.iteration(i) == stops[i]
            assert b.res64(i) == want["res64"], (c["p"], stops[i])
            assert hashlib.sha256(b.residue(i)).hexdigest() == want["sha256"], (c["p"], stops[i])
            # a composite's PRP residue is not 9 unless it is a base-3 probable prime
            assert b.is_prime(i) == (want["res64"] == "0000000000000009")


def test_batch_full_tests_verdicts(gpu_ctx):
    """Known primes and composites run to their verdict inside one batch."""
    mersenne_prime_exps = {3, 5, 7, 13, 17, 19, 31, 61, 89, 107, 127, 521, 607, 1279, 2203, 2281, 9941}
    cases = [c for c in load_fixture("kat.json")["cases"] if c["p"] > 3]
    with m2.Batch(gpu_ctx, [c["p"] for c in cases], [c["mode"] for c in cases]) as b:
        b.run(0)
        for i, c in enumerate(cases):
            assert b.iteration(i) == c["final"]
            assert b.res64(i) == c["checkpoints"][str(c["final"])]["res64"], (c["p"], c["mode"])
            assert b.is_prime(i) == (c["p"] in mersenne_prime_exps), (c["p"], c["mode"])


@pytest.mark.parametrize("wg", [0, 1024])
def test_batch_multipass_plan(gpu_ctx, wg):
    """Several tests sharing the multi-pass launches (grid row per test), with
    the 512- and the 1024-thread centre."""
    cases = [c for c in _sweep_cases() if 6 <= c["ln"] <= 17 and (c["p"] >> c["ln"]) >= 12]
    with m2.Batch(gpu_ctx, [c["p"] for c in cases], [c["mode"] for c in cases], local_stage_cap=16, wg=wg) as b:
        done = 0
        for k in (5, 200):
            b.run(k - done)
            done = k
            for i, c in enumerate(cases):
                assert b.res64(i) == c["checkpoints"][str(k)]["res64"], (c["p"], c["mode"], k)


def test_batch_n2p20_checkpoints(gpu_ctx):
    """Batched tests at n = 2^20 run the separate outer passes and the
    512-thread centre with that length's step geometry compiled in (a single
    test runs the fused pass and the 1024-thread centre instead): every test
    of the batch against the checkpoints of the same exponent."""
    c = load_fixture("c2.json")["cases"][0]
    ks = [k for k in sorted(int(k) for k in c["checkpoints"]) if k <= 1000]
    print("batch of 3, p=%d %s ks=%s" % (c["p"], c["mode"], ks), flush=True)
    with m2.Batch(gpu_ctx, [c["p"]] * 3, c["mode"]) as b:
        done = 0
        for k in ks:
            b.run(k - done)
            done = k
            for i in range(3):
                assert b.res64(i) == c["checkpoints"][str(k)]["res64"], (i, k)


def test_batch_rejects_bad_input(gpu_ctx):
    with pytest.raises(m2.M2Error):
        m2.Batch(gpu_ctx, [127, 2], ["prp", "prp"])
    with pytest.raises(ValueError):
        m2.Batch(gpu_ctx, [127], ["xyz"])
    with m2.Batch(gpu_ctx, [127], ["prp"]) as b:
        with pytest.raises(m2.M2Error):
            b.res64(1)


def test_server_results_match_alone(gpu_ctx):
    """Requests of mixed lengths through a small memory budget: admission
    waits for memory, every completion equals the test's result alone."""
    sweep = {(c["p"], c["mode"]): c for c in _sweep_cases()}
    reqs = []
    for (p, mode), c in sorted(sweep.items()):
        if c["ln"] >= 6:
            reqs.append((p, mode, 200 if c["ln"] % 2 else 50, c))
    c4 = load_fixture("c4.json")["cases"]
    for c in c4[:40]:
        k = c["final"] if c["p"] <= 60000 else 2000
        reqs.append((c["p"], c["mode"], k, c))
    big = max(m2.footprint(p) for p, _, _, _ in reqs)
    budget = 3 * big
    print("server: %d requests, budget %d bytes" % (len(reqs), budget), flush=True)
    with m2.Server(gpu_ctx, hbm_budget_bytes=budget, slice_iters=1000) as s:
        ids = {}
        for p, mode, k, c in reqs:
            ids[s.submit(p, mode, k)] = (p, mode, k, c)
        got = {}
        steps = 0
        while True:
            busy = s.step()
            steps += 1
            st = s.stats()
            assert st["bytes_in_use"] <= budget or st["active"] == 1
            for rid, r64, verdict in s.poll():
                got[rid] = (r64, verdict)
            if not busy:
                break
            assert steps < 10000
        for rid, r64, verdict in s.poll():
            got[rid] = (r64, verdict)
        assert set(got) == set(ids)
        for rid, (p, mode, k, c) in ids.items():
            assert got[rid][0] == c["checkpoints"][str(k)]["res64"], (p, mode, k)
        st = s.stats()
        assert st == {"queued": 0, "active": 0, "bytes_in_use": 0}


def test_server_one_request_larger_than_budget_runs_alone(gpu_ctx):
    c = [c for c in _sweep_cases() if c["ln"] == 12 and c["mode"] == "prp"][0]
    with m2.Server(gpu_ctx, hbm_budget_bytes=1, slice_iters=7) as s:
        a = s.submit(c["p"], "prp", 50)
        b = s.submit(c["p"], "prp", 200)
        out = {}
        while s.step():
            assert s.stats()["active"] <= 1
            out.update({rid: r64 for rid, r64, _ in s.poll()})
        out.update({rid: r64 for rid, r64, _ in s.poll()})
        assert out[a] == c["checkpoints"]["50"]["res64"]
        assert out[b] == c["checkpoints"]["200"]["res64"]


@pytest.mark.parametrize("variant", [dict(carry_pairs=16), dict(max_radix=4), dict(max_radix=4, carry_pairs=4),
                                     dict(wg=1024)])
def test_plan_variants_match(gpu_ctx, variant):
    """Carry-block width, radix and centre workgroup size (what the autotuner
    varies) leave every residue unchanged, on the single-workgroup and
    multi-pass plans."""
    for c in _sweep_cases():
        if not (5 <= c["ln"] <= 17 and c["mode"] == "prp"):
            continue
        for ls in (0, 64):
            if ls and (c["p"] >> c["ln"]) < 12:
                continue
            with m2.Test(gpu_ctx, c["p"], "prp", local_stage_cap=ls, **variant) as t:
                t.run(50)
                assert t.res64() == c["checkpoints"]["50"]["res64"], (c["p"], ls, variant)


@pytest.mark.parametrize("ls", [4, 8, 16, 256])
def test_wide_centre_on_small_blocks(gpu_ctx, ls):
    """The 1024-thread centre on centre blocks of 2^2..2^8 points, where most
    of its threads have no butterfly or squaring pair, gives the exact
    residues."""
    cl = max(2, ls.bit_length() - 1)  # centre blocks of 2^cl <= ls points; multi-pass from n = 2^(cl + 3)
    for c in _sweep_cases():
        if not (cl + 3 <= c["ln"] <= 14 and c["mode"] == "prp") or (c["p"] >> c["ln"]) < 12:
            continue
        with m2.Test(gpu_ctx, c["p"], "prp", local_stage_cap=ls, wg=1024) as t:
            assert "%d threads" % 1024 in t.plan(), t.plan()
            t.run(50)
            assert t.res64() == c["checkpoints"]["50"]["res64"], (c["p"], ls)


def test_autotune_and_tuned_batch(gpu_ctx):
    """The tuner reports every candidate and picks one; a batch built with
    autotune=1 gives the exact results."""
    c16 = [c for c in _sweep_cases() if c["ln"] == 16 and c["mode"] == "prp"][0]
    opts, report = m2.autotune(gpu_ctx, c16["p"], 4)
    print(report)
    assert "best:" in report and opts["max_radix"] in (0, 4, 8)
    cases = [c for c in _sweep_cases() if 8 <= c["ln"] <= 16]
    with m2.Batch(gpu_ctx, [c["p"] for c in cases], [c["mode"] for c in cases], autotune=1) as b:
        b.run(50)
        for i, c in enumerate(cases):
            assert b.res64(i) == c["checkpoints"]["50"]["res64"], (c["p"], c["mode"])


def test_fused_and_separate_outer_passes_agree(gpu_ctx):
    """Single tests fuse the outer passes of consecutive iterations where the
    plan allows; the separate launches give the same residues."""
    for c in [c for c in _sweep_cases() if 13 <= c["ln"] <= 17 and c["mode"] == "prp"][:4]:
        with m2.Test(gpu_ctx, c["p"], "prp", local_stage_cap=64) as a, \
                m2.Test(gpu_ctx, c["p"], "prp", local_stage_cap=64, no_fuse=1) as b:
            for k in (5, 50, 200):
                a.run(k - a.iteration)
                b.run(k - b.iteration)
                assert a.res64() == b.res64() == c["checkpoints"][str(k)]["res64"], (c["p"], k)
    # one-step outer passes (n = 2^13, 2^14 with the largest centre: 2 and 3
    # outer levels), whose fused pass stages through LDS instead of reading
    # and writing global memory directly at the pass edges
    one_step = {c["ln"]: c for c in _sweep_cases() if c["ln"] in (13, 14) and c["mode"] == "prp"}
    for ln, c in sorted(one_step.items()):
        with m2.Test(gpu_ctx, c["p"], "prp", local_stage_cap=2048) as a, \
                m2.Test(gpu_ctx, c["p"], "prp", local_stage_cap=2048, no_fuse=1) as b:
            assert re.search(r"levels 0\.\.%d radix \d\n" % (ln - 12), a.plan()), a.plan()
            a.run(50)
            b.run(50)
            assert a.res64() == b.res64() == c["checkpoints"]["50"]["res64"], (c["p"], 50)
    c2 = load_fixture("c2.json")["cases"][0]
    with m2.Test(gpu_ctx, c2["p"], "prp", no_fuse=1) as b:
        b.run(100)
        assert b.res64() == c2["checkpoints"]["100"]["res64"]


def test_fused_pass_reports_a_wait_time_out(gpu_ctx):
    """The fused outer pass's bounded wait: with the poll limit forced to 0
    from the host, every fused launch reports a time-out (without waiting)
    and the next sync fails with M2_E_STATE; with the limit restored, the
    same test set back to its start by a residue, and a fresh test, run
    exactly. Direct launches only (graph_iters=1), so no captured
    graph keeps the forced value."""
    import ctypes
    c2 = load_fixture("c2.json")["cases"][0]
    prev = ctypes.c_uint32()
    m2.lib().m2_debug_fused_spin_limit(0, ctypes.byref(prev))
    p = c2["p"]
    try:
        with m2.Test(gpu_ctx, p, "prp", graph_iters=1) as t:
            t.run(3)
            with pytest.raises(m2.M2Error) as ei:
                t.sync()
            assert ei.value.code == -6
            # a test whose state timed out is restored by setting a residue
            # (here the start, 3, at iteration 0), which clears the report
            m2.lib().m2_debug_fused_spin_limit(prev.value, None)
            t.set_residue((3).to_bytes((p + 7) // 8, "little"), 0)
            t.run(10)
            assert t.res64() == c2["checkpoints"]["10"]["res64"]
    finally:
        m2.lib().m2_debug_fused_spin_limit(prev.value, None)
    with m2.Test(gpu_ctx, p, "prp", graph_iters=1) as t:
        t.run(10)
        assert t.res64() == c2["checkpoints"]["10"]["res64"]


def _c5_requests():
    """The first 400 requests of the serving workload's stream (seed 1:
    Poisson arrivals, 70% PRP, n = 2^10..2^22), with the stream's 2000
    iterations per request except n >= 2^17, which run big_iters (100), and
    the exact Res64 of each (tests/golden/c5.json)."""
    from mersenne2gpu_amd import workloads as W
    fx = load_fixture("c5.json")
    if "big_iters" not in fx:
        pytest.skip("c5.json lacks the n >= 2^17 entries (gen_fixtures.py c5big)")
    want = {(c["p"], c["mode"], int(k)): v["res64"] for c in fx["cases"] for k, v in c["checkpoints"].items()}
    reqs = []
    for _, p, mode, k in W.serve_stream(fx["seed"])[:400]:
        if W.transformsize(p) >= fx["big_ln"]:
            k = fx["big_iters"]
        reqs.append((p, mode, k, want.get((p, mode, k))))
    return reqs


@pytest.mark.parametrize("budget_gb", [200.0, 1.0])
def test_server_c5_stream(gpu_ctx, budget_gb):
    """The serving workload's stream through m2_server: 200 GB (the workload's
    budget: everything admitted at once) and 1 GB (admission waits for
    memory; the n = 2^22 lanes take a quarter of it each). Every completion is
    checked against the exact value where the fixture holds it (all n <= 2^14
    requests, six at n = 2^16 and every n >= 2^17 one, including the n = 2^20
    and 2^22 lanes), and every one against the same request run alone."""
    from mersenne2gpu_amd import workloads as W
    reqs = _c5_requests()
    lns = [W.transformsize(p) for p, _, _, _ in reqs]
    assert 20 in lns and 22 in lns
    budget = int(budget_gb * 1e9)
    print("c5 stream: %d requests, %d with exact values, budget %.0f GB" % (
        len(reqs), sum(w is not None for _, _, _, w in reqs), budget_gb), flush=True)
    got = {}
    with m2.Server(gpu_ctx, hbm_budget_bytes=budget, slice_iters=1000) as s:
        ids = {s.submit(p, mode, k): i for i, (p, mode, k, _) in enumerate(reqs)}
        steps = waited = 0
        while True:
            busy = s.step()
            steps += 1
            st = s.stats()
            assert st["bytes_in_use"] <= budget or st["active"] == 1
            waited += st["queued"] > 0 and st["active"] > 0
            for rid, r64, _ in s.poll():
                got[ids[rid]] = r64
            if not busy:
                break
            assert steps < 200000
        for rid, r64, _ in s.poll():
            got[ids[rid]] = r64
    assert len(got) == len(reqs)
    if budget_gb < 10:
        assert waited > 0
    checked = 0
    for i, (p, mode, k, want) in enumerate(reqs):
        if want is not None:
            checked += 1
            assert got[i] == want, (i, p, mode, k)
    assert checked >= 300
    # alone: one request of each length, and every n >= 2^20 one
    seen = set()
    for i, (p, mode, k, _) in enumerate(reqs):
        ln = lns[i]
        if ln in seen and ln < 20:
            continue
        seen.add(ln)
        with m2.Test(gpu_ctx, p, mode) as t:
            t.run(k)
            assert t.res64() == got[i], (i, p, mode, k)


def test_batch_known_mersenne_primes_to_completion(gpu_ctx):
    """Whole PRP tests of the Mersenne primes 2^p - 1 with p = 216091,
    756839, 859433, 1257787 and 1398269 (n = 2^13, 2^15, 2^16), a whole LL
    test of 2^216091 - 1 and a whole PRP test of the composite 2^216103 - 1,
    all in one batch: the mathematics fixes the primes' final residues (PRP:
    3^(2^p) = 9 mod 2^p - 1; LL: 0), so every step of these long runs is
    checked at the end without a fixture."""
    ps = [216091, 756839, 859433, 1257787, 1398269, 216091, 216103]
    modes = ["prp"] * 5 + ["ll", "prp"]
    print("batch of %d whole tests" % len(ps), flush=True)
    with m2.Batch(gpu_ctx, ps, modes) as b:
        b.run(0)
        b.sync()
        for i, (p, mode) in enumerate(zip(ps, modes)):
            assert b.iteration(i) == (p if mode == "prp" else p - 2)
        for i in range(5):
            assert b.res64(i) == "0000000000000009", ps[i]
            assert b.is_prime(i)
        assert b.res64(5) == "0000000000000000" and b.is_prime(5)
        assert not b.is_prime(6)
