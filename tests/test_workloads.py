"""Seeded exponent draws used by the benchmark and the fixtures."""
from mersenne2gpu_amd import workloads as W


def test_mt19937_64_reference_value():
    # 10000th output for the default seed 5489 (C++ standard, [rand.predef])
    g = W.MT19937_64(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042


def test_draws_are_prime_and_in_range():
    for seed in (1, 2, 3):
        for lo, hi in (W.C1_RANGE, W.C2_RANGE, W.C3_RANGE):
            p = W.draw_uniform(seed, lo, hi)
            assert lo <= p <= hi and W.is_prime(p)
    ps = W.batch_exponents(1)
    assert len(ps) == 512 and all(W.C4_RANGE[0] <= p <= W.C4_RANGE[1] and W.is_prime(p) for p in ps)
    assert {W.transformsize(p) for p in ps} == set(range(8, 17))


def test_exponent_ranges_match_table():
    assert W.exponent_range(9) == (10499, 20983)
    assert W.exponent_range(20) == (18874379, 36700159)
    assert W.exponent_range(24) == (285212677, 553648103)
