// CPU checks of the transform arithmetic in m2_arith.h (the same inline
// functions the gfx950 kernels compile): split-twiddle products, the 64-bit
// M31 reduction, the lazily reduced M61 butterflies and the centre squaring
// formula, against plain modular arithmetic on random and extreme inputs; and
// the 64-bit normalisation of a carry segment (norm_digit32) against the
// 128-bit Garner reading, mask and shift.
#include <stdio.h>

#include <random>

#include "m2_arith.h"
#include "m2_digits.h"

using namespace m2;
typedef unsigned __int128 u128;

static int fails = 0;
#define CHECK(c, ...)          \
  do {                         \
    if (!(c)) {                \
      if (fails < 20) {        \
        printf("FAIL %s: ", #c); \
        printf(__VA_ARGS__);   \
        printf("\n");          \
      }                        \
      ++fails;                 \
    }                          \
  } while (0)

static uint64_t m61(u128 x) { return (uint64_t)(x % P61); }
static uint64_t canon61(uint64_t x) { return x % P61; }  // P61 (and above) as residues
static uint32_t canon31(uint32_t x) { return x % P31; }
static bool eq61(e61 x, e61 y) { return canon61(x.a) == canon61(y.a) && canon61(x.b) == canon61(y.b); }
static bool eq31(e31 x, e31 y) { return canon31(x.a) == canon31(y.a) && canon31(x.b) == canon31(y.b); }

static std::mt19937_64 rng(12345);
// values in [0, bound] with the ends over-represented
static uint64_t pick(uint64_t bound) {
  switch (rng() % 8) {
    case 0: return 0;
    case 1: return bound;
    case 2: return bound - (rng() % 8 < bound ? rng() % 8 : 0);
    case 3: return P61 < bound ? P61 : bound;
    default: return bound == ~0ull ? rng() : rng() % (bound + 1);
  }
}
static uint32_t pick31() {
  switch (rng() % 6) {
    case 0: return 0;
    case 1: return P31;
    case 2: return P31 - 1;
    default: return (uint32_t)(rng() % (uint64_t(P31) + 1));
  }
}
static e61 rnd61(uint64_t bound) { return {pick(bound), pick(bound)}; }
static e31 rnd31() { return {pick31(), pick31()}; }
static e61 reduced(e61 x) { return {canon61(x.a), canon61(x.b)}; }

int main() {
  const int N = 200000;
  // dot61s: a, b < 2^62, c, d <= P61 (split)
  for (int i = 0; i < N; ++i) {
    const uint64_t a = pick((1ull << 62) - 1), b = pick((1ull << 62) - 1), c = pick(P61), d = pick(P61);
    const uint64_t r = dot61s(a, split31(c), b, split31(d));
    CHECK(r <= P61, "dot61s range %llx", (unsigned long long)r);
    CHECK(canon61(r) == m61((u128)a * c + (u128)b * d), "dot61s a=%llx b=%llx c=%llx d=%llx",
          (unsigned long long)a, (unsigned long long)b, (unsigned long long)c, (unsigned long long)d);
    CHECK(unsplit31(split31(c)) == c, "split round trip");
  }
  // the limb product the kernels use: x, y < 2^62, fully reduced result
  for (int i = 0; i < N; ++i) {
    const uint64_t x = pick((1ull << 62) - 1), y = pick((1ull << 62) - 1);
    const uint64_t r = mul61_limbs(x, y);
    CHECK(r <= P61, "mul61_limbs range %llx", (unsigned long long)r);
    CHECK(canon61(r) == m61((u128)x * y), "mul61_limbs x=%llx y=%llx", (unsigned long long)x,
          (unsigned long long)y);
  }
  // squares with unreduced factors against the reduced-factor square
  for (int i = 0; i < N; ++i) {
    const e61 x = rnd61(P61), q = sqr_lazy(x);
    CHECK(q.a <= P61 && q.b <= P61, "sqr_lazy range");
    CHECK(eq61(q, sqr(x)) && eq61(q, mul(x, x)), "sqr_lazy");
  }
  // twiddle products against the plain GF(p^2) products
  for (int i = 0; i < N; ++i) {
    const e61 x = rnd61(P61x2), w = rnd61(P61);
    const e61 xr = reduced(x);
    const e61 p = mul_tw(x, split31(w)), q = mulconj_tw(x, split31(w));
    CHECK(p.a <= P61 && p.b <= P61 && q.a <= P61 && q.b <= P61, "mul_tw range");
    CHECK(eq61(p, mul(xr, w)), "mul_tw 61");
    CHECK(eq61(q, mulconj(xr, w)), "mulconj_tw 61");
    const e31 y = rnd31(), v = rnd31();
    const e31 s = mul_tw(y, v), t = mulconj_tw(y, v);
    CHECK(s.a <= P31 && s.b <= P31 && t.a <= P31 && t.b <= P31, "mul_tw 31 range");
    CHECK(eq31(s, mul(y, v)), "mul_tw 31");
    CHECK(eq31(t, mulconj(y, v)), "mulconj_tw 31");
  }
  for (int i = 0; i < N; ++i) {
    const uint64_t x = pick((1ull << 63) - 1);
    CHECK(red31(x) <= P31 && canon31(red31(x)) == x % P31, "red31 %llx", (unsigned long long)x);
    const uint64_t P2 = 2ull * P31 * P31, z = i < 64 ? P2 - (uint64_t)i : pick(P2);
    CHECK(red31_dot(z) <= P31 && canon31(red31_dot(z)) == z % P31, "red31_dot %llx", (unsigned long long)z);
    const uint64_t y = pick(~0ull);
    CHECK(rot30_lazy(y) < (1ull << 61) + (1ull << 33) && canon61(rot30_lazy(y)) == m61((u128)y << 30),
          "rot30_lazy");
  }
  // rotations: by the full width they are the identity (the CRT relies on it)
  for (int i = 0; i < N; ++i) {
    const uint64_t x = pick(P61);
    const uint32_t y = pick31();
    for (uint32_t sh = 0; sh <= 61; ++sh)
      CHECK(canon61(rot61(x, sh)) == m61((u128)x << sh) && rot61(x, sh) <= P61, "rot61 %u", sh);
    CHECK(rot61(x, 61) == x && rot31(y, 31) == y, "full rotation");
    for (uint32_t sh = 0; sh <= 31; ++sh)
      CHECK(canon31(rot31(y, sh)) == (uint32_t)(((uint64_t)y << sh) % P31) && rot31(y, sh) <= P31, "rot31 %u", sh);
    if (i > N / 16) break;
  }
  // lazy butterflies against the reduced generic ones: forward inputs have
  // element 0 <= P61 + 7, the rest <= P61; inverse inputs <= P61
  for (int i = 0; i < N / 4; ++i) {
    e61 v[8], ref[8];
    for (int e = 0; e < 8; ++e) v[e] = rnd61(e == 0 ? P61 + 7 : P61);
    for (int e = 0; e < 8; ++e) ref[e] = reduced(v[e]);
    e61 v4[4], r4[4], v2[2], r2[2];
    for (int e = 0; e < 4; ++e) { v4[e] = v[e]; r4[e] = ref[e]; }
    for (int e = 0; e < 2; ++e) { v2[e] = v[e]; r2[e] = ref[e]; }
    dft8_fwd_lazy(v);
    dft8_fwd(ref);
    dft4_fwd_lazy(v4);
    dft4_fwd(r4);
    dft2_fwd_lazy(v2[0], v2[1]);
    dft2(r2[0], r2[1]);
    for (int e = 0; e < 8; ++e) {
      CHECK(v[e].a <= P61 + 7 && v[e].b <= P61 + 7, "fwd8 range");
      CHECK(eq61(v[e], ref[e]), "dft8_fwd_lazy e=%d", e);
    }
    for (int e = 0; e < 4; ++e) CHECK(v4[e].a <= P61 + 7 && eq61(v4[e], r4[e]), "dft4_fwd_lazy e=%d", e);
    for (int e = 0; e < 2; ++e) CHECK(v2[e].a <= P61 + 7 && eq61(v2[e], r2[e]), "dft2_fwd_lazy e=%d", e);

    for (int e = 0; e < 8; ++e) v[e] = ref[e] = rnd61(P61);
    for (int e = 0; e < 4; ++e) v4[e] = r4[e] = v[e];
    for (int e = 0; e < 2; ++e) v2[e] = r2[e] = v[e];
    dft8_inv_lazy(v);
    dft8_inv(ref);
    dft4_inv_lazy(v4);
    dft4_inv(r4);
    dft2_inv_lazy(v2[0], v2[1]);
    dft2(r2[0], r2[1]);
    CHECK(v[0].a <= P61 && v[0].b <= P61 && v4[0].a <= P61 && v4[0].b <= P61, "inv output 0 reduced");
    CHECK(v2[0].a <= P61 && v2[1].a <= P61 && v2[0].b <= P61 && v2[1].b <= P61, "inv2 reduced");
    for (int e = 0; e < 8; ++e) {
      CHECK(v[e].a <= P61 + 7 && v[e].b <= P61 + 7, "inv8 range");
      CHECK(fold61(v[e].a) <= P61 && fold61(v[e].b) <= P61, "inv8 one fold reduces");
      CHECK(eq61(v[e], ref[e]), "dft8_inv_lazy e=%d", e);
    }
    for (int e = 0; e < 4; ++e) CHECK(eq61(v4[e], r4[e]) && v4[e].a <= P61 + 7 && v4[e].b <= P61 + 7, "dft4_inv_lazy e=%d", e);
    for (int e = 0; e < 2; ++e) CHECK(eq61(v2[e], r2[e]) && v2[e].a <= P61 + 7 && v2[e].b <= P61 + 7, "dft2_inv_lazy e=%d", e);
  }
  // centre squaring: 4 z_k^2 - C, 4 z_mk^2 - conj(C) equals the defining
  // formula v0 +/- v1 with u0 = z_k + conj(z_mk), u1 = z_k - conj(z_mk)
  for (int i = 0; i < N / 4; ++i) {
    const e61 zk = rnd61(P61 + 7), zmk = rnd61(P61 + 7), tw = rnd61(P61);
    e61 a = zk, b = zmk;
    square_pair(a, b, split31(add(e61{1, 0}, tw)));
    const e61 k = reduced(zk), c = conj(reduced(zmk));
    const e61 u0 = add(k, c), u1 = sub(k, c);
    const e61 v0 = sub(sqr(u0), mul(sqr(u1), tw)), v1 = mul(u0, add(u1, u1));
    CHECK(a.a <= P61 && a.b <= P61 && b.a <= P61 && b.b <= P61, "square range");
    CHECK(eq61(a, add(v0, v1)) && eq61(b, conj(sub(v0, v1))), "square_pair 61");
    const e31 yk = rnd31(), ymk = rnd31(), t3 = rnd31();
    e31 p = yk, q = ymk;
    square_pair(p, q, add(e31{1, 0}, t3));
    const e31 c3 = conj(ymk), w0 = add(yk, c3), w1 = sub(yk, c3);
    const e31 x0 = sub(sqr(w0), mul(sqr(w1), t3)), x1 = mul(w0, add(w1, w1));
    CHECK(eq31(p, add(x0, x1)) && eq31(q, conj(sub(x0, x1))), "square_pair 31");
  }
  // 92-bit CRT: every x in [-2^60, 2^92 - 2^64] comes back from
  // (x mod P61, x mod P31), with either residue in its second representation
  // of 0 where it is 0; the edges of the old 3 2^90 rule and of the band
  // below M are taken explicitly.
  {
    typedef __int128 i128;
    const i128 M = (i128)P61 * P31;
    const i128 POS_MAX = ((i128)1 << 92) - ((i128)1 << 64);
    const i128 NEG_MIN = -((i128)1 << 60);
    auto res61 = [](i128 x) { i128 r = x % (i128)P61; return (uint64_t)(r < 0 ? r + P61 : r); };
    auto res31 = [](i128 x) { i128 r = x % (i128)P31; return (uint32_t)(r < 0 ? r + P31 : r); };
    auto check = [&](i128 x, uint64_t add) {
      uint64_t s = res61(x);
      uint32_t t = res31(x);
      for (int alt = 0; alt < 4; ++alt) {
        const uint64_t sa = (alt & 1) && s == 0 ? P61 : s;
        const uint32_t ta = (alt & 2) && t == 0 ? P31 : t;
        const i128 got = crt61_31(sa, ta, add);
        CHECK(got == x + (i128)add, "crt61_31 x=%016llx%016llx alt=%d", (unsigned long long)(uint64_t)(x >> 64),
              (unsigned long long)(uint64_t)x, alt);
      }
    };
    const i128 edges[] = {0, 1, 2, P31 - 1, P31, (i128)P61, (i128)3 << 90, ((i128)3 << 90) - 1, ((i128)3 << 90) + 1,
                          ((i128)3 << 90) + ((i128)3 << 59), ((i128)1 << 91), ((i128)1 << 92) - ((i128)1 << 66),
                          POS_MAX, POS_MAX - 1, -1, -2, -(i128)P31, -((i128)1 << 40), NEG_MIN, NEG_MIN + 1};
    for (const i128 x : edges) check(x, 0);
    for (int i = 0; i < N; ++i) {
      const int band = (int)(rng() % 4);
      i128 x;
      if (band == 0) x = (i128)((((u128)rng() << 64) | rng()) % (u128)(POS_MAX + 1));  // any positive term
      else if (band == 1) x = ((i128)3 << 90) + (i128)((((u128)rng() << 64) | rng()) % (u128)(POS_MAX - ((i128)3 << 90) + 1));  // the old rule's band
      else if (band == 2) x = POS_MAX - (i128)(rng() % (1ull << 40));                  // just below the bound
      else x = -(i128)(1 + rng() % (uint64_t)(-(NEG_MIN)));                            // negative terms
      check(x, i % 5 == 0 ? (rng() % (1ull << 40)) : 0);
    }
    // the reading the band replaced (negative from u >= 3 2^59) misreads the
    // terms this check covers; the check above must see the difference
    auto old_rule = [](uint64_t s, uint32_t t) {
      uint64_t u = sub61(s, t);
      u = add61(u, rot61(u, 31));
      const int64_t v = (int64_t)u - ((u >= (3ull << 59)) ? (int64_t)P61 : 0);
      return ((i128)v << 31) - (i128)v + (i128)t;
    };
    const i128 hi_term = ((i128)3 << 90) + 12345;
    CHECK(old_rule(res61(hi_term), res31(hi_term)) == hi_term - M, "old rule reads 3 2^90 + 12345 as negative");
    CHECK(crt61_31(res61(hi_term), res31(hi_term)) == hi_term, "crt61_31 reads 3 2^90 + 12345 as positive");
    CHECK(crt61_31(res61(M - 1), res31(M - 1)) == -1, "M - 1 is -1");
  }
  // segment normalisation with 64-bit words (16 <= q <= 31) against the
  // 128-bit reading: 8 terms in [0, 2^92 - 2^64] with random unweighting
  // rotations, residues in either representation of 0, PRP (add 0) and LL
  // (add 2^w - 1, carry-in -2 at the first digit), and positive carry-ins
  {
    typedef __int128 i128;
    const i128 POS_MAX = ((i128)1 << 92) - ((i128)1 << 64);
    auto res61 = [](i128 x) { return (uint64_t)(x % (i128)P61); };
    auto res31 = [](i128 x) { return (uint32_t)(x % (i128)P31); };
    for (int trial = 0; trial < N / 4; ++trial) {
      const uint32_t q = 16u + (uint32_t)(rng() % 16);
      CHECK(norm32_applies(q), "norm32_applies %u", q);
      const int mode = (int)(rng() % 3);  // 0 PRP, 1 LL, 2 PRP with a positive carry-in
      const i128 cin = mode == 1 ? (rng() % 2 ? -2 : 0) : (mode == 2 ? (i128)(rng() % (1ull << 32)) : 0);
      i128 c = cin;
      uint64_t cA = (uint64_t)(int64_t)cin, cB = 0;
      for (int i = 0; i < 8; ++i) {
        const uint32_t a61 = (uint32_t)(rng() % 61), a31 = (uint32_t)(rng() % 31), wide = (uint32_t)(rng() % 2);
        const uint32_t meta = (a61 << 16) | (a31 << 24) | (wide << 31);
        const uint32_t w = q + wide;
        const int band = (int)(rng() % 4);
        i128 x;
        if (band == 0) x = (i128)((((u128)rng() << 64) | rng()) % (u128)(POS_MAX + 1));
        else if (band == 1) x = POS_MAX - (i128)(rng() % (1ull << 40));
        else if (band == 2) x = (i128)(rng() % 4) * (i128)P61 * (rng() % 2 ? 1 : (i128)P31);  // residues 0
        else x = (i128)(rng() % (1ull << 40));
        uint64_t s1 = res61(x);
        uint32_t t1 = res31(x);
        if (s1 == 0 && rng() % 2) s1 = P61;
        if (t1 == 0 && rng() % 2) t1 = P31;
        // the stored residues are the weighted ones: rotating them by the
        // unweighting counts gives s1, t1
        const uint64_t s = rot61(s1, 61u - meta_u61(meta));
        const uint32_t t = rot31(t1, 31u - meta_u31(meta));
        const uint64_t add = mode == 1 ? (1ull << w) - 1ull : 0ull;
        const i128 acc = crt61_31(rot61(s, meta_u61(meta)), rot31(t, meta_u31(meta)), add) + c;
        const uint64_t want = (uint64_t)acc & ((1ull << w) - 1ull);
        c = acc >> w;
        const uint32_t got = norm_digit32(s, t, meta, q, add, cA, cB);
        CHECK(got == want, "norm_digit32 q=%u digit %d: %x against %llx", q, i, got, (unsigned long long)want);
      }
      const i128 cout = ((i128)cB << 32) + (i128)cA;
      CHECK(cout == c, "norm_digit32 carry-out q=%u", q);
    }
    CHECK(norm32_applies(16) && !norm32_applies(15) && !norm32_applies(32), "norm32_applies range");
    CHECK(norm64_applies(32) && norm64_applies(43) && !norm64_applies(31) && !norm64_applies(44),
          "norm64_applies range");
    // widths 32..44 (q = 32..43): the one-word carry against the 128-bit reading
    for (int trial = 0; trial < N / 4; ++trial) {
      const uint32_t q = 32u + (uint32_t)(rng() % 12);
      const int mode = (int)(rng() % 3);
      const i128 cin = mode == 1 ? (rng() % 2 ? -2 : 0) : (mode == 2 ? (i128)(rng() % (1ull << 60)) : 0);
      i128 c = cin;
      uint64_t cw = (uint64_t)(int64_t)cin;
      for (int i = 0; i < 8; ++i) {
        const uint32_t a61 = (uint32_t)(rng() % 61), a31 = (uint32_t)(rng() % 31), wide = (uint32_t)(rng() % 2);
        const uint32_t meta = (a61 << 16) | (a31 << 24) | (wide << 31);
        const uint32_t w = q + wide;
        const int band = (int)(rng() % 4);
        i128 x;
        if (band == 0) x = (i128)((((u128)rng() << 64) | rng()) % (u128)(POS_MAX + 1));
        else if (band == 1) x = POS_MAX - (i128)(rng() % (1ull << 40));
        else if (band == 2) x = (i128)(rng() % 4) * (i128)P61 * (rng() % 2 ? 1 : (i128)P31);
        else x = (i128)(rng() % (1ull << 40));
        uint64_t s1 = res61(x);
        uint32_t t1 = res31(x);
        if (s1 == 0 && rng() % 2) s1 = P61;
        if (t1 == 0 && rng() % 2) t1 = P31;
        const uint64_t s = rot61(s1, 61u - meta_u61(meta));
        const uint32_t t = rot31(t1, 31u - meta_u31(meta));
        const uint64_t add = mode == 1 ? (1ull << w) - 1ull : 0ull;
        const i128 acc = crt61_31(rot61(s, meta_u61(meta)), rot31(t, meta_u31(meta)), add) + c;
        const uint64_t want = (uint64_t)acc & ((1ull << w) - 1ull);
        c = acc >> w;
        const uint64_t got = norm_digit64(s, t, meta, q, add, cw);
        CHECK(got == want, "norm_digit64 q=%u digit %d: %llx against %llx", q, i, (unsigned long long)got,
              (unsigned long long)want);
      }
      CHECK((i128)cw == c, "norm_digit64 carry-out q=%u", q);
      // ripple of the carry-out through 7 digits below 2^w, the top digit taking the rest
      uint64_t d[8], got[8], want[8];
      for (int i = 0; i < 8; ++i) d[i] = rng() % 3 == 0 ? (1ull << q) - 1ull : rng() % (1ull << q);
      i128 ref = c;
      uint64_t cr = cw;
      for (int i = 0; i < 7; ++i) {
        const i128 a = (i128)d[i] + ref;
        want[i] = (uint64_t)a & ((1ull << q) - 1ull);
        ref = a >> q;
        got[i] = cr ? ripple_digit64(d[i], 0u, q, cr) : d[i];
      }
      want[7] = d[7] + (uint64_t)ref;
      got[7] = d[7] + cr;
      for (int i = 0; i < 8; ++i) CHECK(got[i] == want[i], "ripple_digit64 q=%u digit %d", q, i);
    }
    // the ripple of a carry in [0, 2^77) through 7 digits below 2^w, the top
    // digit taking the rest, against 128-bit addition and shifts
    for (int trial = 0; trial < N / 4; ++trial) {
      const uint32_t q = 16u + (uint32_t)(rng() % 16);
      uint64_t d[8];
      uint32_t meta[8];
      for (int i = 0; i < 8; ++i) {
        meta[i] = (uint32_t)(rng() % 2) << 31;
        const uint32_t w = q + (meta[i] >> 31);
        d[i] = rng() % 3 == 0 ? (1ull << w) - 1ull : rng() % (1ull << w);
      }
      const int kind = (int)(rng() % 3);
      i128 c = kind == 0 ? (i128)(rng() % (1ull << 40)) : (kind == 1 ? (i128)((((u128)rng() << 64) | rng()) % ((u128)1 << 77)) : 0);
      uint64_t cA = (uint64_t)c & 0xFFFFFFFFull, cB = (uint64_t)(c >> 32);
      i128 ref = c;
      uint64_t got[8], want[8];
      for (int i = 0; i < 7; ++i) {
        const uint32_t w = q + (meta[i] >> 31);
        const i128 acc = (i128)d[i] + ref;
        want[i] = (uint64_t)acc & ((1ull << w) - 1ull);
        ref = acc >> w;
        got[i] = (cA | cB) ? ripple_digit32(d[i], meta[i], q, cA, cB) : d[i];
      }
      want[7] = d[7] + (uint64_t)ref;
      got[7] = d[7] + cA + (cB << 32);
      for (int i = 0; i < 8; ++i) CHECK(got[i] == want[i], "ripple_digit32 q=%u digit %d", q, i);
    }
  }
  printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
