"""The size of the squaring's convolution terms, which the CRT must read.

The engine reads a CRT value x in [0, M), M = (2^61 - 1)(2^31 - 1), as the
term itself up to 2^92 - 2^64 and as x - M in a narrow band below M
(crt61_31 in m2_arith.h). That is exact only if every term of the weighted
cyclic self-convolution of the digits lies in [0, 2^(ln+2q+3) - 2^(2q+2)].
These checks compute the terms exactly with Python integers:

* for residues with every digit near its maximum (M_p - 3 and M_p - 1) at the
  top of odd-length exponent ranges, where ln + 2q + 2 = 91: the largest
  terms are above 3 2^90 (so a reading that took that band as negative
  fails) and within the bound;
* for the digit vector with every digit at 2^width (one more than a
  canonical digit can be, as the engine's segment tops may be): the bound
  holds at every exponent tried, and at the top of odd-length ranges the
  largest term comes within a factor 1 - 2/n of it, so it cannot be lowered.
"""
import random

import pytest

P61 = (1 << 61) - 1
P31 = (1 << 31) - 1
M = P61 * P31


def transformsize(p):
    ln = 3
    while ln + 2 * ((p >> ln) + 1) >= 92:
        ln += 1
    return ln


def bitpos(p, n):
    return [-(-p * j // n) for j in range(n + 1)]


def digits_of(p, n, x):
    B = bitpos(p, n)
    return [(x >> B[j]) & ((1 << (B[j + 1] - B[j])) - 1) for j in range(n)]


def terms(p, n, d):
    """c_k = sum over i + j = k (mod n) of d_i d_j 2^(B_i + B_j - B_k - p [i + j >= n])."""
    B = bitpos(p, n)
    c = [0] * n
    for i in range(n):
        di, bi = d[i], B[i]
        if not di:
            continue
        for j in range(n):
            k = i + j
            if k >= n:
                k -= n
                c[k] += (di * d[j]) << (bi + B[j] - B[k] - p)
            else:
                c[k] += (di * d[j]) << (bi + B[j] - B[k])
    return c


def bound(p, ln):
    q = p >> ln
    return (1 << (ln + 2 * q + 3)) - (1 << (2 * q + 2))


def test_weighted_convolution_matches_integer_square():
    """terms() is the IBDWT's convolution: sum c_k 2^B_k = x^2 (mod 2^p - 1)."""
    rnd = random.Random(5)
    for p in (89, 349, 1373):
        ln = transformsize(p)
        n = 1 << ln
        x = rnd.randrange(1 << p)
        c = terms(p, n, digits_of(p, n, x))
        B = bitpos(p, n)
        assert sum(ck << B[k] for k, ck in enumerate(c)) % ((1 << p) - 1) == x * x % ((1 << p) - 1)


@pytest.mark.parametrize("p", [1373, 5351, 20983])
def test_near_maximal_residues_reach_above_3_2_90(p):
    ln = transformsize(p)
    n = 1 << ln
    q = p >> ln
    assert ln % 2 == 1 and ln + 2 * q + 2 == 91
    for r in ((1 << p) - 4, (1 << p) - 2):  # M_p - 3, M_p - 1
        c = terms(p, n, digits_of(p, n, r))
        top = max(c)
        assert min(c) >= 0
        assert top <= bound(p, ln) <= (1 << 92) - (1 << 64) < M
        # the terms the replaced reading (negative above 3 2^90) got wrong
        assert sum(1 for x in c if x >= 3 << 90) > 0, (p, top.bit_length())


@pytest.mark.parametrize("seed", range(6))
def test_bound_for_digits_at_two_to_the_width(seed):
    rnd = random.Random(seed)
    # one exponent per length: the range top (odd ln: the tight case) or a random one
    ln = rnd.choice([4, 5, 6, 7, 8, 9])
    q_max = (89 - ln) // 2 if ln % 2 else (88 - ln) // 2
    p = rnd.choice([(q_max + 1) * (1 << ln) - 1, rnd.randrange(q_max << ln, (q_max + 1) << ln) | 1])
    assert transformsize(p) == ln
    n = 1 << ln
    B = bitpos(p, n)
    d = [1 << (B[j + 1] - B[j]) for j in range(n)]
    c = terms(p, n, d)
    assert max(c) <= bound(p, ln)
    if ln + 2 * (p >> ln) + 2 == 91 and p == (q_max + 1) * (1 << ln) - 1:
        # at the range top the bound is met but for one digit of width q
        assert max(c) * n >= bound(p, ln) * (n - 2)


def test_two_pairs_at_digit_zero_are_unweighted():
    """For i = 0 or j = 0 the weight exponent B_i + B_j - B_k is 0, which is
    what keeps the bound below 2^(ln+2q+3)."""
    for p in (349, 1373, 5351, 81919, 27899687):
        ln = transformsize(p)
        n = 1 << ln
        B = bitpos(p, n)
        for k in range(0, n, max(1, n // 64)):
            assert B[0] + B[k] - B[k] == 0
        # every other pair's exponent is 0 or 1
        rnd = random.Random(p)
        for _ in range(2000):
            i, j = rnd.randrange(n), rnd.randrange(n)
            k = (i + j) % n
            e = B[i] + B[j] - B[k] - (p if i + j >= n else 0)
            assert e in (0, 1)
