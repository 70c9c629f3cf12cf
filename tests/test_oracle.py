"""Pins the oracle (oracle/m2_oracle.c) before results are verified against it.

Each oracle function is checked against something that does not restate its
own formula: Python integers, closed forms, brute-force transforms, the
convolution theorem, and exact residues from GMP / Python (tests/golden).
"""
import ctypes
import random

import pytest

from conftest import load_fixture
from oracle_ctypes import GF

P61 = (1 << 61) - 1
P31 = (1 << 31) - 1
EDGE61 = [0, 1, 2, P61 - 1, P61 - 2, 1 << 60, (1 << 32) - 1, 1 << 32]
EDGE31 = [0, 1, 2, P31 - 1, P31 - 2, 1 << 30, (1 << 16) + 3]


def test_scalar_fields_match_python_integers(oracle):
    rnd = random.Random(7)
    xs = EDGE61 + [rnd.randrange(P61) for _ in range(300)]
    for a in xs[:60]:
        for b in xs[::7]:
            assert oracle.add61(a, b) == (a + b) % P61
            assert oracle.sub61(a, b) == (a - b) % P61
            assert oracle.mul61(a, b) == (a * b) % P61
    for a in xs[:40]:
        for s in range(0, 61):
            assert oracle.shl61(a, s) == (a << s) % P61
    ys = EDGE31 + [rnd.randrange(P31) for _ in range(300)]
    for a in ys[:60]:
        for b in ys[::7]:
            assert oracle.add31(a, b) == (a + b) % P31
            assert oracle.sub31(a, b) == (a - b) % P31
            assert oracle.mul31(a, b) == (a * b) % P31
    for a in ys[:40]:
        for s in range(0, 31):
            assert oracle.shl31(a, s) == (a << s) % P31


def test_gf_product_matches_complex_modular_product(oracle):
    rnd = random.Random(3)
    for _ in range(500):
        a = GF(rnd.randrange(P61), rnd.randrange(P61), rnd.randrange(P31), rnd.randrange(P31))
        b = GF(rnd.randrange(P61), rnd.randrange(P61), rnd.randrange(P31), rnd.randrange(P31))
        c = GF()
        oracle.gf_mul(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        assert c.s0 == (a.s0 * b.s0 - a.s1 * b.s1) % P61
        assert c.s1 == (a.s0 * b.s1 + a.s1 * b.s0) % P61
        assert c.t0 == (a.t0 * b.t0 - a.t1 * b.t1) % P31
        assert c.t1 == (a.t0 * b.t1 + a.t1 * b.t0) % P31


# Exponent -> transform length table of the reference (mersenne2_gpu.cpp:12-42).
TABLE = [(353, 683, 4), (691, 1373, 5), (1381, 2687, 6), (2689, 5351, 7), (5381, 10487, 8), (10499, 20983, 9),
         (21001, 40949, 10), (40961, 81919, 11), (81929, 159739, 12), (159763, 319483, 13), (319489, 622577, 14),
         (622603, 1245169, 15), (1245187, 2424827, 16), (2424833, 4849651, 17), (4849687, 9437179, 18),
         (9437189, 18874367, 19), (18874379, 36700159, 20), (36700201, 73400311, 21), (73400329, 142606333, 22),
         (142606357, 285212659, 23), (285212677, 553648103, 24), (553648171, 1107296251, 25),
         (1107296257, 2147483647, 26), (2147483659, 4294967291, 27)]


def test_transformsize_matches_reference_table(oracle):
    for lo, hi, ln in TABLE:
        assert oracle.transformsize(lo) == ln
        assert oracle.transformsize(hi) == ln
    # exactness bound: n * (2^(q+1))^2 < 2^92 > M61*M31 margin
    for p in [3, 100, 349, 19937, 27899687, 449846671]:
        ln = oracle.transformsize(p)
        assert ln + 2 * ((p >> ln) + 1) < 92


@pytest.mark.parametrize("p", [3, 7, 127, 353, 19937, 21001, 132049, 2424827])
def test_digit_tables_closed_forms(oracle, p):
    ln = oracle.transformsize(p)
    n = 1 << ln
    rc, wd, w61, w31 = oracle.tables(p, ln)
    assert rc == 0
    q = p >> ln
    assert sum(wd) == p
    assert set(wd) <= {q, q + 1}
    # weight a_j = 2^(ceil(pj/n) - pj/n) means a_j^n = 2^(n ceil(pj/n) - pj):
    # in M61, 2 has order 61, so w61_j * n = n ceil(pj/n) - pj (mod 61).
    for j in range(n):
        c = -(-p * j // n)
        assert (w61[j] * n - (n * c - p * j)) % 61 == 0
        assert (w31[j] * n - (n * c - p * j)) % 31 == 0


def _gpow(a, e, P):
    r = (1, 0)
    while e:
        if e & 1:
            r = ((r[0] * a[0] - r[1] * a[1]) % P, (r[0] * a[1] + r[1] * a[0]) % P)
        a = ((a[0] * a[0] - a[1] * a[1]) % P, (2 * a[0] * a[1]) % P)
        e >>= 1
    return r


@pytest.mark.parametrize("k", [1, 2, 3, 9, 20, 27])
def test_roots_have_exact_order(oracle, k):
    r = GF()
    oracle.root(k, ctypes.byref(r))
    for P, v in ((P61, (r.s0, r.s1)), (P31, (r.t0, r.t1))):
        assert _gpow(v, 1 << k, P) == (1, 0)
        assert _gpow(v, 1 << (k - 1), P) == (P - 1, 0)


def _bitrev(i, bits):
    return int(format(i, "0%db" % bits)[::-1], 2) if bits else 0


def _rand_gf(rnd, h):
    z = (GF * h)()
    for i in range(h):
        z[i] = GF(rnd.randrange(P61), rnd.randrange(P61), rnd.randrange(P31), rnd.randrange(P31))
    return z


@pytest.mark.parametrize("ln", [3, 4, 5, 6, 7, 8])
def test_forward_is_bitreversed_dft(oracle, ln):
    """Brute-force O(h^2) DFT with w_h = root of order h: position k holds X[bitrev(k)]."""
    h = 1 << (ln - 1)
    rnd = random.Random(ln)
    z = _rand_gf(rnd, h)
    x = [(z[i].s0, z[i].s1, z[i].t0, z[i].t1) for i in range(h)]
    w = oracle.twiddles(ln)
    oracle.forward(ln, z, w)
    r = GF()
    oracle.root(ln - 1, ctypes.byref(r))
    bits = ln - 1
    for k in range(h):
        f = _bitrev(k, bits)
        for P, idx, rv in ((P61, (0, 1), (r.s0, r.s1)), (P31, (2, 3), (r.t0, r.t1))):
            acc = (0, 0)
            wf = _gpow(rv, f, P)
            t = (1, 0)
            for i in range(h):
                xi = (x[i][idx[0]], x[i][idx[1]])
                acc = ((acc[0] + xi[0] * t[0] - xi[1] * t[1]) % P, (acc[1] + xi[0] * t[1] + xi[1] * t[0]) % P)
                t = ((t[0] * wf[0] - t[1] * wf[1]) % P, (t[0] * wf[1] + t[1] * wf[0]) % P)
            got = (z[k].s0, z[k].s1) if P == P61 else (z[k].t0, z[k].t1)
            assert got == acc, (ln, k)


def test_planted_twiddle_mistake_is_caught(oracle):
    """A wrong twiddle entry must break the brute-force DFT identity."""
    ln, h = 6, 32
    rnd = random.Random(11)
    z = _rand_gf(rnd, h)
    z2 = (GF * h)(*z)
    w = oracle.twiddles(ln)
    oracle.forward(ln, z, w)
    bad = oracle.twiddles(ln)
    bad[5], bad[6] = bad[6], bad[5]
    oracle.forward(ln, z2, bad)
    assert any((z[i].s0, z[i].s1) != (z2[i].s0, z2[i].s1) for i in range(h))


@pytest.mark.parametrize("ln", [3, 4, 7, 10, 13])
def test_inverse_after_forward_is_h_times_identity(oracle, ln):
    h = 1 << (ln - 1)
    rnd = random.Random(100 + ln)
    z = _rand_gf(rnd, h)
    x = [(z[i].s0, z[i].s1, z[i].t0, z[i].t1) for i in range(h)]
    w = oracle.twiddles(ln)
    oracle.forward(ln, z, w)
    oracle.inverse(ln, z, w)
    for i in range(h):
        assert (z[i].s0, z[i].s1, z[i].t0, z[i].t1) == (
            x[i][0] * h % P61, x[i][1] * h % P61, x[i][2] * h % P31, x[i][3] * h % P31)


@pytest.mark.parametrize("ln", [3, 4, 5, 8, 10])
def test_square_half_is_cyclic_self_convolution(oracle, ln):
    """forward -> square_half -> inverse on a real length-n sequence packed as
    n/2 complex values gives 4h times its cyclic self-convolution."""
    n, h = 1 << ln, 1 << (ln - 1)
    rnd = random.Random(ln * 31)
    a = [rnd.randrange(1 << 20) for _ in range(n)]
    z = (GF * h)()
    for k in range(h):
        z[k] = GF(a[2 * k], a[2 * k + 1], a[2 * k] % P31, a[2 * k + 1] % P31)
    w = oracle.twiddles(ln)
    oracle.forward(ln, z, w)
    oracle.square_half(ln, z, w)
    oracle.inverse(ln, z, w)
    conv = [0] * n
    for i in range(n):
        for j in range(n):
            conv[(i + j) % n] += a[i] * a[j]
    for k in range(h):
        for d, (s, t) in enumerate(((z[k].s0, z[k].t0), (z[k].s1, z[k].t1))):
            c = conv[2 * k + d] * 4 * h
            assert s == c % P61 and t == c % P31


def test_garner_round_trip(oracle):
    rnd = random.Random(5)
    M = P61 * P31
    out = (ctypes.c_uint64 * 2)()
    for x in [0, 1, P31, P61, M - 1, (1 << 91) + 12345] + [rnd.randrange(M) for _ in range(2000)]:
        oracle.garner(x % P61, x % P31, out)
        assert out[0] | (out[1] << 64) == x


def test_digit_residue_reconstruction(oracle):
    """Residue of a digit vector equals sum d_j 2^ceil(pj/n) mod M_p, with M_p -> 0."""
    rnd = random.Random(9)
    for p in [127, 607, 19937]:
        ln = oracle.transformsize(p)
        n = 1 << ln
        _, wd, _, _ = oracle.tables(p, ln)
        for trial in range(4):
            if trial == 0:
                d = [(1 << w) - 1 for w in wd]  # all ones = M_p = 0
            else:
                d = [rnd.randrange(1 << (w + 1)) for w in wd]  # not normalised
            arr = (ctypes.c_uint64 * n)(*d)
            nb = (p + 7) // 8
            buf = ctypes.create_string_buffer(nb)
            assert oracle.digits_residue(p, ln, arr, buf, nb) == 0
            val, pos = 0, 0
            for j in range(n):
                val += d[j] << pos
                pos += wd[j]
            assert int.from_bytes(buf.raw, "little") == val % ((1 << p) - 1)


def _check_cases(oracle, cases, carry_pairs=64, max_iters=None):
    for c in cases:
        ks = [int(k) for k in c["checkpoints"]]
        if max_iters is not None:
            ks = [k for k in ks if k <= max_iters]
        if not ks:
            continue
        got, verdict, residue = oracle.run(c["p"], c["mode"], ks, carry_pairs)
        for k in ks:
            assert got[k] == c["checkpoints"][str(k)]["res64"], (c["p"], c["mode"], k)
        last = c["checkpoints"][str(max(ks))]
        if "residue_hex" in last:
            assert residue[::-1].hex().upper() == last["residue_hex"]


def test_oracle_matches_kat_fixtures(oracle):
    fx = load_fixture("kat.json")
    _check_cases(oracle, fx["cases"])
    # verdicts of the reference's own self-tests (fused.cpp:536-539): PRP for
    # these p, composite for 11 and 23
    for c in fx["cases"]:
        if c["mode"] == "prp" and c["p"] in (3, 5, 7, 13, 17, 19, 31, 61, 89, 107, 127, 11, 23):
            _, verdict, _ = oracle.run(c["p"], "prp", [c["final"]])
            assert verdict == (0 if c["p"] in (11, 23) else 1)


def test_oracle_matches_sweep_fixtures(oracle):
    fx = load_fixture("sweep.json")
    cases = [c for c in fx["cases"] if c["ln"] <= 14]
    _check_cases(oracle, cases, max_iters=50)


@pytest.mark.parametrize("carry_pairs", [1, 16, 256])
def test_oracle_block_size_does_not_change_results(oracle, carry_pairs):
    fx = load_fixture("c1.json")
    _check_cases(oracle, fx["cases"][:1], carry_pairs=carry_pairs, max_iters=500)


def test_oracle_matches_c1_checkpoints(oracle):
    fx = load_fixture("c1.json")
    _check_cases(oracle, fx["cases"], max_iters=2000)
