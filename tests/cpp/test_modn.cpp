// CPU checks of the host residue arithmetic of the Gerbicz check
// (m2_modn.h): addition, negation and division by 4 mod 2^p - 1, against
// 128-bit integers for p <= 127 and against identities for large p.
#include <stdio.h>

#include <random>

#include "m2_modn.h"

using namespace m2::modn;
typedef unsigned __int128 u128;

static int fails = 0;
#define CHECK(c, ...)            \
  do {                           \
    if (!(c)) {                  \
      if (fails < 20) {          \
        printf("FAIL %s: ", #c); \
        printf(__VA_ARGS__);     \
        printf("\n");            \
      }                          \
      ++fails;                   \
    }                            \
  } while (0)

static Num from128(u128 v, const ModN& m) {
  Num a(m.nw, 0);
  a[0] = (uint64_t)v;
  if (m.nw > 1) a[1] = (uint64_t)(v >> 64);
  return a;
}
static u128 to128(const Num& a) { return (u128)a[0] | (a.size() > 1 ? (u128)a[1] << 64 : 0); }

int main() {
  std::mt19937_64 rng(7);
  for (uint32_t p : {3u, 5u, 7u, 13u, 31u, 61u, 63u, 64u + 1, 89u, 107u, 127u}) {
    const ModN m(p);
    const u128 N = ((u128)1 << p) - 1;
    auto rnd = [&]() {
      const u128 v = ((u128)rng() << 64 | rng()) % N;
      switch (rng() % 8) {
        case 0: return (u128)0;
        case 1: return N - 1;
        case 2: return (u128)1;
        default: return v;
      }
    };
    for (int it = 0; it < 20000; ++it) {
      const u128 a = rnd(), b = rnd();
      const Num A = from128(a, m), B = from128(b, m);
      CHECK(to128(add(A, B, m)) == (a + b) % N, "p=%u add", p);
      CHECK(to128(neg(A, m)) == (N - a) % N, "p=%u neg", p);
      // a / 4: the x with 4 x = a mod N
      const u128 q = to128(quarter(A, m));
      const u128 q2 = (q + q) % N;  // doubling twice stays below 2^128 for p <= 127
      CHECK(q < N && (q2 + q2) % N == a, "p=%u quarter", p);
    }
  }
  // large p: identities (a - a = 0, 4 (a / 4) = a, commutativity, round trip through bytes)
  for (uint32_t p : {19937u, 86243u, 1398269u}) {
    const ModN m(p);
    for (int it = 0; it < 20; ++it) {
      Num a(m.nw), b(m.nw);
      for (size_t i = 0; i < m.nw; ++i) {
        a[i] = rng();
        b[i] = rng();
      }
      a[m.nw - 1] &= m.top;
      b[m.nw - 1] &= m.top;
      if (it == 0) std::fill(a.begin(), a.end(), 0);
      if (it == 1) {  // N - 1
        std::fill(a.begin(), a.end(), ~0ull);
        a[m.nw - 1] = m.top;
        a[0] ^= 1;
      }
      const Num zero(m.nw, 0);
      CHECK(add(a, neg(a, m), m) == zero, "p=%u a - a", p);
      CHECK(add(a, b, m) == add(b, a, m), "p=%u commutes", p);
      const Num q = quarter(a, m);
      const Num q2 = add(q, q, m);
      CHECK(add(q2, q2, m) == a, "p=%u 4 (a / 4)", p);
      CHECK(from_bytes(to_bytes(a, m), m) == a, "p=%u bytes", p);
      CHECK(add(add(a, b, m), neg(b, m), m) == a, "p=%u (a + b) - b", p);
    }
  }
  printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
