"""Seeded exponent draws for the benchmark and test workloads.

An exponent is drawn from a seed as: mt19937-64(seed) -> value in [lo, hi]
(uniform, or uniform in ln p for log-uniform draws) -> the next prime <= hi,
else the previous prime. The generator is written out here (not taken from
numpy) so that C++ and Python agree on the draws.
"""
from __future__ import annotations

import math
from typing import List, Tuple

_MASK = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto & Nishimura, 2004)."""

    def __init__(self, seed: int):
        self.mt = [0] * 312
        self.mt[0] = seed & _MASK
        for i in range(1, 312):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & _MASK
        self.idx = 312

    def _twist(self) -> None:
        mt = self.mt
        for i in range(312):
            x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def next(self) -> int:
        if self.idx >= 312:
            self._twist()
        x = self.mt[self.idx]
        self.idx += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & _MASK

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for q in small:
        if n % q == 0:
            return n == q
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def prime_near(v: int, lo: int, hi: int) -> int:
    """Next prime >= v that is <= hi, else the previous prime >= lo."""
    x = max(v, lo)
    while x <= hi:
        if is_prime(x):
            return x
        x += 1
    x = min(v, hi)
    while x >= lo:
        if is_prime(x):
            return x
        x -= 1
    raise ValueError("no prime in [%d, %d]" % (lo, hi))


def draw_uniform(seed: int, lo: int, hi: int) -> int:
    g = MT19937_64(seed)
    return prime_near(lo + g.next() % (hi - lo + 1), lo, hi)


def draw_log_uniform(g: MT19937_64, lo: int, hi: int) -> int:
    v = int(math.exp(math.log(lo) + g.uniform() * (math.log(hi) - math.log(lo))))
    return prime_near(v, lo, hi)


def transformsize(p: int) -> int:
    """log2 of the transform length for exponent p (mersenne2_gpu.cpp:286-293)."""
    ln = 2
    while True:
        ln += 1
        if ln + 2 * ((p >> ln) + 1) < 92:
            return ln


def exponent_range(ln: int) -> Tuple[int, int]:
    """Smallest and largest prime exponent whose transform length is 2^ln."""
    lo = 3
    while transformsize(lo) < ln:
        lo = lo * 2
    # binary search the first p with transformsize >= ln
    a, b = 2, max(lo, 4)
    while transformsize(b) < ln:
        b *= 2
    while a < b:
        mid = (a + b) // 2
        if transformsize(mid) >= ln:
            b = mid
        else:
            a = mid + 1
    first = a
    a, b = first, first * 4
    while transformsize(b) <= ln:
        b *= 2
    while a < b:
        mid = (a + b) // 2
        if transformsize(mid) > ln:
            b = mid
        else:
            a = mid + 1
    last = a - 1
    p0 = first
    while not is_prime(p0):
        p0 += 1
    p1 = last
    while not is_prime(p1):
        p1 -= 1
    return p0, p1


# Workload definitions (exponent ranges of the transform-length table,
# mersenne2_gpu.cpp:12-42).
C1_RANGE = (10499, 20983)         # n = 512
C2_RANGE = (18874379, 36700159)   # n = 2^20
C3_RANGE = (285212677, 553648103)  # n = 2^24
C4_RANGE = (5381, 2424827)        # n = 2^8 .. 2^16


def batch_exponents(seed: int, count: int = 512, lo: int = C4_RANGE[0], hi: int = C4_RANGE[1]) -> List[int]:
    g = MT19937_64(seed)
    return [draw_log_uniform(g, lo, hi) for _ in range(count)]


def log_uniform_over_lengths(g: MT19937_64, ln_lo: int, ln_hi: int) -> int:
    """Exponent log-uniform over the exponent ranges of n = 2^ln_lo .. 2^ln_hi."""
    lo = exponent_range(ln_lo)[0]
    hi = exponent_range(ln_hi)[1]
    return draw_log_uniform(g, lo, hi)


def serve_stream(seed: int, count: int = 2000, rate: float = 40.0, iters: int = 2000,
                 prp_frac: float = 0.7) -> List[Tuple[float, int, str, int]]:
    """Poisson arrival stream of LL/PRP requests: (arrival_s, p, mode, iterations).

    Inter-arrival gaps are exponential with mean 1/rate; 70% PRP; exponents
    drawn so that 80% fall in n = 2^10..2^14, 18% in 2^16..2^20 and 2% at
    n = 2^22 (log-uniform within each band). Each request runs
    min(iters, full test) iterations (iters = 0: the full test)."""
    g = MT19937_64(seed)
    out = []
    t = 0.0
    for _ in range(count):
        t += -math.log(1.0 - g.uniform()) / rate
        mode = "prp" if g.uniform() < prp_frac else "ll"
        u = g.uniform()
        if u < 0.80:
            p = log_uniform_over_lengths(g, 10, 14)
        elif u < 0.98:
            p = log_uniform_over_lengths(g, 16, 20)
        else:
            p = log_uniform_over_lengths(g, 22, 22)
        full = p if mode == "prp" else p - 2
        out.append((t, p, mode, min(full, iters) if iters else full))
    return out
