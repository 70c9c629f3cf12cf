"""GPU checks at the edges of the 64-bit carry forms (digit widths 16..44).

Tests whose digits are q or q + 1 bits wide with 16 <= q <= 31 normalise and
ripple their carries in two 64-bit words (norm_digit32, ripple_digit32 in
m2_digits.h), those with 32 <= q <= 43 in one (norm_digit64). q = 16 (only at
n >= 2^25) and q = 31 are the first form's ends (w = 16 and w = 32, where the
digit mask is all 32 bits and the carry's shift by 32 - w is 0), q = 32 the
second form's first width; q = 17 occurs only at n = 8 (p = 137) and at the
bottom of the ranges of n = 2^21..2^24, which no golden fixture reaches. Each
exponent here runs:

  * from the initial state, against the oracle (or, at p = 137, the whole
    test against Python integers; n = 2^25 only from the residues below);
  * from near-maximal residues, whose results the mathematics fixes:
    PRP from M_p - 3 gives 9, 81, 6561; LL from M_p - 3 gives 7, 47, 2207 and
    from M_p - 1 stays at M_p - 1.
"""
import pytest

import mersenne2gpu_amd as m2

pytestmark = pytest.mark.gpu

# (p, log2 n, q = p >> log2 n)
EDGES = [(137, 3, 17), (262139, 13, 31), (33554393, 20, 31), (33554467, 20, 32), (36700201, 21, 17),
         (553648171, 25, 16)]


def _r64(x):
    return "%016X" % ((x - 1 if x < 0 else x) & ((1 << 64) - 1))


@pytest.mark.parametrize("p,ln,q", EDGES)
def test_widths_at_the_edges(gpu_ctx, p, ln, q):
    assert m2.transform_log2(p) == ln and p >> ln == q
    for mode, d, c in (("prp", 3, 0), ("ll", 3, 2), ("ll", 1, 2)):
        print("case p=%d q=%d %s M_p-%d" % (p, q, mode, d), flush=True)
        with m2.Test(gpu_ctx, p, mode) as t:
            t.set_residue(((1 << p) - 1 - d).to_bytes((p + 7) // 8, "little"), 0)
            v = d * d - c
            for it in range(3):
                t.run(1)
                assert t.res64() == _r64(v), (p, mode, d, it)
                v = v * v - c if d != 1 else v


@pytest.mark.parametrize("p,ln,q", [e for e in EDGES if 1000 < e[0] < 10 ** 8])
def test_widths_against_the_oracle(gpu_ctx, oracle, p, ln, q):
    ks = [40, 60]
    for mode in ("prp", "ll"):
        print("case p=%d q=%d %s from the start, checkpoints %s" % (p, q, mode, ks), flush=True)
        want, _, _ = oracle.run(p, mode, ks)
        got = m2.run_checkpoints(gpu_ctx, p, mode, ks)
        assert got == want, (p, mode, got, want)


def test_whole_tests_at_q17_n8(gpu_ctx):
    p = 137
    M = (1 << p) - 1
    prp = pow(3, 1 << p, M)
    s = 4
    for _ in range(p - 2):
        s = (s * s - 2) % M
    for mode, iters, want in (("prp", p, prp), ("ll", p - 2, s)):
        with m2.Test(gpu_ctx, p, mode) as t:
            t.run(iters)
            assert t.res64() == "%016X" % (want & ((1 << 64) - 1)), (mode, t.res64())
            assert t.residue() == want.to_bytes((p + 7) // 8, "little")
