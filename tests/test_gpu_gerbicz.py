"""PRP runs under the Gerbicz check (m2_test_run_checked): the same residues as
plain runs, no failed check on a clean run, and a planted error (one flipped
residue bit between two checks, set from the host) caught and rolled back."""
import subprocess

import pytest

import mersenne2gpu_amd as m2
from conftest import load_fixture

pytestmark = pytest.mark.gpu


def test_checked_full_prp_test(gpu_ctx):
    with m2.Test(gpu_ctx, 19937, "prp") as t:
        errors = t.run_checked(19937, 1000, 4)
        assert errors == 0
        assert t.iteration == 19937
        assert t.res64() == "0000000000000009" and t.is_prime()


@pytest.mark.parametrize("fault_block", [1, 3, 4, 7])
def test_checked_run_catches_a_planted_error(gpu_ctx, fault_block):
    c = load_fixture("c1.json")["cases"][0]
    with m2.Test(gpu_ctx, c["p"], "prp") as t:
        t.run(400)
        m2.debug_gerbicz_fault(fault_block)
        errors = t.run_checked(1600, 200, 4)  # checks after blocks 4 and 8
        assert errors == 1
        assert t.iteration == 2000
        assert t.res64() == c["checkpoints"]["2000"]["res64"]


def test_checked_tail_and_partial_check(gpu_ctx):
    """iters not a multiple of the block: the tail is run twice and compared."""
    c = load_fixture("c1.json")["cases"][0]
    with m2.Test(gpu_ctx, c["p"], "prp") as t:
        assert t.run_checked(2000 - 77, 300, 5) == 0
        t.run(77)
        assert t.res64() == c["checkpoints"]["2000"]["res64"]


def test_checked_run_at_n2p20(gpu_ctx):
    c = load_fixture("c2.json")["cases"][0]
    with m2.Test(gpu_ctx, c["p"], c["mode"]) as t:
        assert t.run_checked(1000, 250, 2) == 0
        assert t.res64() == c["checkpoints"]["1000"]["res64"]


def test_checked_run_refuses_ll_and_bad_blocks(gpu_ctx):
    with m2.Test(gpu_ctx, 19937, "ll") as t:
        with pytest.raises(m2.M2Error):
            t.run_checked(100, 10, 2)
    with m2.Test(gpu_ctx, 19937, "prp") as t:
        with pytest.raises(m2.M2Error):
            t.run_checked(100, 0, 2)


def test_cli_gerbicz_option():
    r = subprocess.run([m2.CLI_PATH, "19937", "prp", "--gerbicz", "500"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert "gerbicz: block 500, 0 failed checks" in lines
    assert any(l.startswith("2^19937 - 1 is a probable prime, elapsed ") for l in lines)
    assert lines[-1] == "Res64: 0000000000000009"
