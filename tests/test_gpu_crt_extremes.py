"""Closed-form GPU checks at the largest convolution terms the engine meets.

At the top of a transform length's exponent range the squaring's convolution
terms reach almost M61*M31 (about 2^92): with ln + 2q + 2 = 91 (odd ln, q =
p >> ln) a term of a residue whose digits are all near their maximum lies
above 3 2^90, and the CRT must still read it as the positive value it is.
Residues of that kind are the states next to the end of a test of a Mersenne
prime: M_p - 3 = 3^(2^(p-1)) one PRP squaring before the verdict, and
+-2^((p+1)/2) one Lucas-Lehmer step before it. Each case sets such a residue,
runs a few iterations and compares Res64 (and the residue) with what the
mathematics fixes:

  PRP  from M_p - 3: 9, 81, 6561;      from M_p - 1: 1 (every iteration)
  LL   from M_p - 1: M_p - 1 (a fixed point of s -> s^2 - 2)
       from M_p - 3: 7, 47, 2207;      from M_p - 2^((p+1)/2): 0

The exponents are the reference table's largest prime of every odd length
2^3..2^23 (mersenne2_gpu.cpp:12-42 lists them), and of two even lengths.
Lengths up to 2^12 run the single-workgroup kernel, longer ones the
multi-pass kernels (one iteration: separate inverse pass; more: the fused
outer pass), and -LS forces the multi-pass kernels onto the short lengths.
"""
import pytest

import mersenne2gpu_amd as m2

pytestmark = pytest.mark.gpu

# largest prime exponent of each length (ln: p)
TOPS = {3: 349, 5: 1373, 7: 5351, 9: 20983, 11: 81919, 13: 319483, 15: 1245169, 17: 4849651, 19: 18874367,
        20: 36700159, 21: 73400311, 23: 285212659, 24: 553648103}


def _res(p, x):
    """Canonical residue bytes of a small x (or of M_p + x for a small
    negative x); never a long division by M_p, which is slow at these p."""
    if x < 0:
        return ((1 << p) - 1 + x).to_bytes((p + 7) // 8, "little")
    return x.to_bytes((p + 7) // 8, "little")


def _r64(x):
    """Res64 of x mod M_p for a small x: M_p + x = 2^p - 1 + x for x < 0."""
    return "%016X" % ((x - 1 if x < 0 else x) & ((1 << 64) - 1))


def _run_case(ctx, p, mode, d, iters, ls=0, small=False, **opts):
    """Sets R = M_p - d (d small, or a power of two 2^k with p < 2k < 2p),
    runs `iters` iterations one at a time and checks each Res64 against the
    closed form: the first squaring gives d^2 (2^(2k - p) for d = 2^k), every
    later value stays small."""
    label = "M_p-%d" % d if d < 1 << 20 else "M_p-2^%d" % (d.bit_length() - 1)
    print("case p=%d %s start=%s iters=%d ls=%d" % (p, mode, label, iters, ls), flush=True)
    c = 2 if mode == "ll" else 0
    k = d.bit_length() - 1
    v = (1 << (2 * k - p)) if d == 1 << k and 2 * k > p else d * d
    with m2.Test(ctx, p, mode, local_stage_cap=ls, **opts) as t:
        assert ("k_small" in t.plan()) == small or not (small or opts), t.plan()
        t.set_residue(((1 << p) - 1 - d).to_bytes((p + 7) // 8, "little"), 0)
        for it in range(1, iters + 1):
            t.run(1)
            v = v - c if it == 1 else v * v - c
            assert abs(v) < 1 << 62
            assert t.res64() == _r64(v), (p, mode, label, it, ls)
        if p < 200000:
            assert t.residue() == _res(p, v), (p, mode, label, ls)


@pytest.mark.parametrize("ln", sorted(TOPS))
def test_near_maximal_residues_at_range_top(gpu_ctx, ln):
    p = TOPS[ln]
    iters = 3 if ln <= 21 else 2
    _run_case(gpu_ctx, p, "prp", 3, iters)
    _run_case(gpu_ctx, p, "prp", 1, iters)
    _run_case(gpu_ctx, p, "ll", 1, iters)
    _run_case(gpu_ctx, p, "ll", 3, iters)
    _run_case(gpu_ctx, p, "ll", 1 << ((p + 1) // 2), 1)


@pytest.mark.parametrize("ln", [5, 7, 9, 11])
def test_near_maximal_residues_on_multipass_kernels(gpu_ctx, ln):
    p = TOPS[ln]
    for ls in (16, 256):
        _run_case(gpu_ctx, p, "prp", 3, 3, ls=ls)
        _run_case(gpu_ctx, p, "ll", 1, 3, ls=ls)


def test_lucas_lehmer_end_of_a_prime_test(gpu_ctx):
    """The last two LL steps of the Mersenne prime 2^21701 - 1 (n = 2^10,
    odd-length top window): s_(p-4) -> s_(p-3) = +-2^((p+1)/2) -> 0. The state
    s_(p-4) comes from a whole run, so this is the engine's own path to a
    near-maximal residue."""
    p = 21701
    with m2.Test(gpu_ctx, p, "ll") as t:
        t.run(p - 4)
        M = (1 << p) - 1
        s = int.from_bytes(t.residue(), "little")
        t.run(1)
        s3 = int.from_bytes(t.residue(), "little")
        assert s3 == (s * s - 2) % M
        assert s3 in ((1 << ((p + 1) // 2)) % M, M - (1 << ((p + 1) // 2)))
        t.run(1)
        assert t.res64() == "0000000000000000" and t.is_prime()


@pytest.mark.parametrize("radix", [8, 4])
def test_near_maximal_residues_on_lds_resident_n2p13(gpu_ctx, radix):
    """n = 2^13 on the single-workgroup plan (h = 2^12: two carry segments per
    thread, the last step's and the squaring twiddles read from global
    memory), which a plain single test of that length does not use: an
    explicit radix selects the default plan instead of the single-test
    profile."""
    for p in (TOPS[13], 159763):
        _run_case(gpu_ctx, p, "prp", 3, 3, small=True, max_radix=radix)
        _run_case(gpu_ctx, p, "ll", 1, 3, small=True, max_radix=radix)
        _run_case(gpu_ctx, p, "ll", 1 << ((p + 1) // 2), 1, small=True, max_radix=radix)
