"""Parity of the HIP engine (libm2.so through its C ABI) with exact residues.

Expected values are the committed fixtures under tests/golden (exact GMP /
Python arithmetic, cross-checked against the oracle by test_oracle.py) and,
for randomised cases, the oracle itself. No tolerance: Res64 and residues must
be bit-identical.
"""
import hashlib
import os
import subprocess

import pytest

import mersenne2gpu_amd as m2
from conftest import load_fixture

pytestmark = pytest.mark.gpu


def _check(ctx, case, ls=0, max_k=None, graph_iters=0, **opts):
    ks = sorted(int(k) for k in case["checkpoints"])
    if max_k is not None:
        ks = [k for k in ks if k <= max_k]
    print("case p=%d %s ls=%d ks=%s %s" % (case["p"], case["mode"], ls, ks, opts), flush=True)
    t = m2.Test(ctx, case["p"], case["mode"], graph_iters=graph_iters, **opts) if not ls else _test_ls(ctx, case, ls)
    with t:
        done = 0
        for k in ks:
            t.run(k - done)
            done = k
            want = case["checkpoints"][str(k)]
            assert t.res64() == want["res64"], (case["p"], case["mode"], k, ls)
            if "sha256" in want:
                r = t.residue()
                assert hashlib.sha256(r).hexdigest() == want["sha256"], (case["p"], k)
            if "residue_hex" in want:
                assert t.residue()[::-1].hex().upper() == want["residue_hex"]


def _test_ls(ctx, case, ls):
    import ctypes
    opts = m2._Opts()
    opts.local_stage_cap = ls
    h = ctypes.c_void_p()
    rc = m2.lib().m2_test_create(ctx.h, case["p"], 1 if case["mode"] == "prp" else 0, ctypes.byref(opts),
                                 ctypes.byref(h))
    m2._chk(rc, "m2_test_create")
    t = m2.Test.__new__(m2.Test)
    t.p, t.mode, t.h = case["p"], case["mode"], h
    return t


def test_kat_full_tests(gpu_ctx):
    for c in load_fixture("kat.json")["cases"]:
        _check(gpu_ctx, c)


def test_sweep_every_transform_length(gpu_ctx):
    for c in load_fixture("sweep.json")["cases"]:
        _check(gpu_ctx, c)


@pytest.mark.parametrize("ls", [4, 16, 64, 256])
def test_sweep_multipass_plan_at_small_sizes(gpu_ctx, ls):
    """-LS forces the multi-pass kernels (outer/centre, and the middle pass
    when the outer levels exceed 9) onto small transforms."""
    for c in load_fixture("sweep.json")["cases"]:
        if 4 <= c["ln"] <= 17 and (c["p"] >> c["ln"]) >= 12:
            _check(gpu_ctx, c, ls=ls)


def test_c1_full_prp_tests(gpu_ctx):
    for c in load_fixture("c1.json")["cases"]:
        _check(gpu_ctx, c)


def test_c1_on_multipass_plan(gpu_ctx):
    for c in load_fixture("c1.json")["cases"][:2]:
        _check(gpu_ctx, c, ls=64, max_k=2000)


def test_c2_checkpoints_n2p20(gpu_ctx):
    for c in load_fixture("c2.json")["cases"]:
        _check(gpu_ctx, c)


def test_lengths_2p18_to_2p23_checkpoints(gpu_ctx):
    """One exponent per length between the n = 2^20 and n = 2^24 workloads
    (tests/golden/mid.json, exact GMP arithmetic): their plans mix outer and
    middle passes that take the step geometry as arguments with the centres
    that have it compiled in."""
    for c in load_fixture("mid.json")["cases"]:
        _check(gpu_ctx, c)


def test_c2_centre_workgroup_sizes(gpu_ctx):
    """The 512-thread (radix-8) and 1024-thread (radix-4) centre give the
    same residues; the plan text names the size in use, and a single test
    at this length takes the 1024-thread centre by default."""
    c = load_fixture("c2.json")["cases"][0]
    for wg in (512, 1024):
        _check(gpu_ctx, c, max_k=1000, wg=wg)
        with m2.Test(gpu_ctx, c["p"], c["mode"], wg=wg) as t:
            assert "2 blocks x 1024, %d threads" % wg in t.plan(), t.plan()
    with m2.Test(gpu_ctx, c["p"], c["mode"]) as t:
        assert "2 blocks x 1024, 1024 threads" in t.plan(), t.plan()


def test_c2_graph_and_direct_launches_agree(gpu_ctx):
    c = load_fixture("c2.json")["cases"][0]
    _check(gpu_ctx, c, max_k=100, graph_iters=1)
    _check(gpu_ctx, c, max_k=100, graph_iters=7)


def test_c3_ll_checkpoints_n2p24(gpu_ctx):
    for c in load_fixture("c3.json")["cases"]:
        _check(gpu_ctx, c)


@pytest.mark.parametrize("p,mode,k1,k2", [(19937, "prp", 700, 1300), (27899687, "prp", 3, 10),
                                          (132049, "ll", 50, 120)])
def test_resume_from_residue(gpu_ctx, p, mode, k1, k2):
    with m2.Test(gpu_ctx, p, mode) as a:
        a.run(k1)
        r = a.residue()
        a.run(k2 - k1)
        want = a.residue()
    with m2.Test(gpu_ctx, p, mode) as b:
        b.set_residue(r, k1)
        assert b.iteration == k1
        b.run(k2 - k1)
        assert b.residue() == want


@pytest.mark.parametrize("p,mode,prime", [(9941, "prp", True), (9941, "ll", True), (11213, "prp", True),
                                          (10501, "prp", False), (21701, "ll", True), (21011, "ll", False)])
def test_verdicts(gpu_ctx, p, mode, prime):
    with m2.Test(gpu_ctx, p, mode) as t:
        t.run(t.total_iters)
        assert t.is_prime() == prime


def test_cli_full_prp_test_and_residue_lines(tmp_path):
    dump = tmp_path / "r.bin"
    r = subprocess.run([m2.CLI_PATH, "19937", "prp", "--res64-at", "100,2000", "--dump-residue", str(dump)],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "p=19937, ln=9, transform=256 (n=512)"
    assert lines[1].startswith("device=0 (")
    assert lines[2].startswith("carry_blocks=")
    assert "Res64@100: EB67D29C64046DB3" in lines
    assert "Res64@2000: B5DBC55987E2FD16" in lines
    assert any(l.startswith("2^19937 - 1 is a probable prime, elapsed ") for l in lines)
    assert lines[-1] == "Res64: 0000000000000009"
    assert dump.read_bytes() == (9).to_bytes((19937 + 7) // 8, "little")


def test_cli_ll_default_mode_and_benchmark_iters():
    r = subprocess.run([m2.CLI_PATH, "20983", "-N", "1000"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "benchmark_iters=1000" in r.stdout
    assert "iter 1000/1000 (100.0%)" in r.stdout
    assert "2^20983 - 1 is composite" in r.stdout


def _profile_blocks(out):
    """Kernel profile summaries printed by --profile-kernels: title -> {kernel: (ms, launches)}."""
    blocks, cur = {}, None
    for line in out.splitlines():
        if line.startswith("Kernel profile summary"):
            cur = blocks.setdefault(line.rstrip(":"), {})
        elif cur is not None and line.startswith("  ") and " ms total (" in line:
            name, rest = line.strip().split(": ", 1)
            ms = float(rest.split(" ms total")[0])
            launches = int(rest.split("launches=")[1].rstrip(")"))
            cur[name] = (ms, launches)
        else:
            cur = None
    return blocks


@pytest.mark.parametrize("p, iters, every, kernels", [
    (27899687, 500, 250, {"k_centre": 500}),  # n = 2^20: fused outer pass + centre per iteration
    (19937, 2000, 1000, {"k_small": 2000}),    # n = 512: single-workgroup plan, one launch per iteration
])
def test_cli_profile_kernels(p, iters, every, kernels):
    r = subprocess.run([m2.CLI_PATH, str(p), "prp", "-N", str(iters), "--profile-kernels", "--profile-every",
                        str(every), "--res64-at", str(iters)],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    blocks = _profile_blocks(r.stdout)
    want = ["Kernel profile summary at iter %d" % k for k in range(every, iters + 1, every)]
    assert list(blocks) == want + ["Kernel profile summary (final)"]
    final = blocks["Kernel profile summary (final)"]
    for prefix, launches in kernels.items():
        got = [v for k, v in final.items() if k.startswith(prefix)]
        assert got and sum(n for _, n in got) == launches, final
    assert all(ms > 0 for ms, _ in final.values())
    # totals only grow from one summary to the next
    for a, b in zip(want, want[1:] + ["Kernel profile summary (final)"]):
        assert all(blocks[b][k][1] >= n for k, (_, n) in blocks[a].items())
    # profiled launches compute the same residue as graphed ones
    if p == 27899687:
        c = [c for c in load_fixture("c2.json")["cases"] if c["p"] == p][0]
        assert "Res64@%d: %s" % (iters, c["checkpoints"][str(iters)]["res64"]) in r.stdout
    else:
        assert "Res64@2000: B5DBC55987E2FD16" in r.stdout  # as in the graphed run above
