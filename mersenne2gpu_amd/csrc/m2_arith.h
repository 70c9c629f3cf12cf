// Exact arithmetic in GF(M61^2) and GF(M31^2), i^2 = -1, for gfx950 HIP kernels
// and for the host table builder (the same inline functions compile for both).
//
// Elements lie in [0, M] between operations, M = 2^k - 1 standing for 0 as
// well, so that every reduction is a Mersenne fold x = lo + hi (mod 2^k - 1)
// or an unsigned min, without compares. A product is 32x32->64 multiplies
// (v_mad_u64_u32) whose pieces are folded together. Multiplication by a power
// of two is a rotation of the k-bit word, which is how the IBDWT weights are
// applied.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define M2_HD __host__ __device__ __forceinline__
#else
#define M2_HD static inline
#endif

namespace m2 {

constexpr uint64_t P61 = (uint64_t(1) << 61) - 1;
constexpr uint32_t P31 = (uint32_t(1) << 31) - 1;

struct alignas(16) e61 { uint64_t a, b; };  // a + b i in GF(M61^2)
struct alignas(8) e31 { uint32_t a, b; };   // a + b i in GF(M31^2)

// ---- M61 ----
// Values are kept in [0, P61]: P61 is a second representation of 0, which
// lets sums and products be reduced by folding (x mod 2^61 + x >> 61) alone.
M2_HD uint64_t fold61(uint64_t x) { return (x & P61) + (x >> 61); }
M2_HD uint64_t add61(uint64_t x, uint64_t y) { return fold61(x + y); }
M2_HD uint64_t neg61(uint64_t x) { return x ^ P61; }  // P61 - x
M2_HD uint64_t sub61(uint64_t x, uint64_t y) { return fold61(x + (y ^ P61)); }
// Unreduced pieces of a 61x61-bit product: with 32-bit limbs x = x1 2^32 + x0,
// x y = p11 2^64 + p01 2^32 + p00, and 2^64 = 2^3, 2^61 = 1 (mod P61).
M2_HD uint64_t fold_product61(uint64_t p00, uint64_t p01, uint64_t p11) {
  const uint64_t r = (p11 << 3) + (p00 & P61) + (p00 >> 61) + (p01 >> 29) + ((p01 & 0x1FFFFFFFull) << 32);
  return fold61(fold61(r));
}
M2_HD uint64_t mul61(uint64_t x, uint64_t y) {
#if defined(__HIP_DEVICE_COMPILE__)
  // one v_mad_u64_u32 per partial product
  const uint32_t x0 = uint32_t(x), x1 = uint32_t(x >> 32);
  const uint32_t y0 = uint32_t(y), y1 = uint32_t(y >> 32);
  return fold_product61(uint64_t(x0) * y0, uint64_t(x0) * y1 + uint64_t(x1) * y0, uint64_t(x1) * y1);
#else
  unsigned __int128 t = (unsigned __int128)x * y;
  return fold61(fold61((uint64_t(t) & P61) + uint64_t(t >> 61)));
#endif
}
// a b + c d (mod P61) with the two products' pieces summed before folding.
M2_HD uint64_t dot61(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t a0 = uint32_t(a), a1 = uint32_t(a >> 32), b0 = uint32_t(b), b1 = uint32_t(b >> 32);
  const uint32_t c0 = uint32_t(c), c1 = uint32_t(c >> 32), d0 = uint32_t(d), d1 = uint32_t(d >> 32);
  const uint64_t lo = fold61(uint64_t(a0) * b0) + fold61(uint64_t(c0) * d0);       // < 2^62 + 16
  const uint64_t mid = uint64_t(a0) * b1 + uint64_t(a1) * b0;                          // < 2^62
  const uint64_t mid2 = uint64_t(c0) * d1 + uint64_t(c1) * d0;                         // < 2^62
  const uint64_t hi = uint64_t(a1) * b1 + uint64_t(c1) * d1;                           // < 2^59
  const uint64_t r = lo + (hi << 3) + (mid >> 29) + ((mid & 0x1FFFFFFFull) << 32) + (mid2 >> 29) +
                     ((mid2 & 0x1FFFFFFFull) << 32);                                    // < 2^64
  return fold61(fold61(r));
#else
  return add61(mul61(a, b), mul61(c, d));
#endif
}
// x * 2^s for 0 <= s < 61: rotate the 61-bit word left by s.
M2_HD uint64_t rot61(uint64_t x, uint32_t s) { return ((x << s) & P61) | (x >> (61u - s)); }

// ---- M31 ----
// Also in [0, P31] with P31 as a second 0; the unsigned min picks the
// reduced one of t and t -/+ P31.
M2_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }
M2_HD uint32_t add31(uint32_t x, uint32_t y) { const uint32_t t = x + y; return umin32(t, t - P31); }
M2_HD uint32_t sub31(uint32_t x, uint32_t y) { const uint32_t t = x - y; return umin32(t, t + P31); }
M2_HD uint32_t neg31(uint32_t x) { return P31 - x; }
M2_HD uint32_t mul31(uint32_t x, uint32_t y) {
  const uint64_t t = uint64_t(x) * y;
  const uint32_t r = (uint32_t(t) & P31) + uint32_t(t >> 31);
  return umin32(r, r - P31);
}
M2_HD uint32_t rot31(uint32_t x, uint32_t s) {
  const uint64_t t = uint64_t(x) << s;
  return (uint32_t(t) & P31) | uint32_t(t >> 31);
}

// ---- GF(M61^2) ----
M2_HD e61 add(e61 x, e61 y) { return {add61(x.a, y.a), add61(x.b, y.b)}; }
M2_HD e61 sub(e61 x, e61 y) { return {sub61(x.a, y.a), sub61(x.b, y.b)}; }
M2_HD e61 conj(e61 x) { return {x.a, neg61(x.b)}; }
M2_HD e61 muli(e61 x) { return {neg61(x.b), x.a}; }   // x * i
M2_HD e61 mulmi(e61 x) { return {x.b, neg61(x.a)}; }  // x * -i
// (a + bi)(c + di) = (ac + (P-b) d) + (ad + bc) i: each component is one dot
// product, whose four 64-bit partial products share two folds.
M2_HD e61 mul(e61 x, e61 y) { return {dot61(x.a, y.a, neg61(x.b), y.b), dot61(x.a, y.b, x.b, y.a)}; }
// x * conj(y) = (ac + bd) + (bc - ad) i
M2_HD e61 mulconj(e61 x, e61 y) { return {dot61(x.a, y.a, x.b, y.b), dot61(x.b, y.a, neg61(x.a), y.b)}; }
M2_HD e61 sqr(e61 x) {
  const uint64_t t = mul61(x.a, x.b);
  return {mul61(add61(x.a, x.b), sub61(x.a, x.b)), add61(t, t)};
}
// The primitive 8th root of unity in this field is -(1+i) 2^30; its powers are
// rotations plus add/sub, no multiplication.
M2_HD e61 mulw8(e61 x) {  // * w8
  return {rot61(sub61(x.b, x.a), 30), rot61(neg61(add61(x.a, x.b)), 30)};
}
M2_HD e61 mulw8c(e61 x) {  // * conj(w8) = -(1-i) 2^30
  return {rot61(neg61(add61(x.a, x.b)), 30), rot61(sub61(x.a, x.b), 30)};
}
M2_HD e61 rot(e61 x, uint32_t s0, uint32_t s1) { return {rot61(x.a, s0), rot61(x.b, s1)}; }

// ---- GF(M31^2) ----
M2_HD e31 add(e31 x, e31 y) { return {add31(x.a, y.a), add31(x.b, y.b)}; }
M2_HD e31 sub(e31 x, e31 y) { return {sub31(x.a, y.a), sub31(x.b, y.b)}; }
M2_HD e31 conj(e31 x) { return {x.a, neg31(x.b)}; }
M2_HD e31 muli(e31 x) { return {neg31(x.b), x.a}; }
M2_HD e31 mulmi(e31 x) { return {x.b, neg31(x.a)}; }
M2_HD e31 mul(e31 x, e31 y) {
  const uint32_t ac = mul31(x.a, y.a), bd = mul31(x.b, y.b);
  const uint32_t k = mul31(add31(x.a, x.b), add31(y.a, y.b));
  return {sub31(ac, bd), sub31(sub31(k, ac), bd)};
}
M2_HD e31 mulconj(e31 x, e31 y) {
  const uint32_t ac = mul31(x.a, y.a), bd = mul31(x.b, y.b);
  const uint32_t k = mul31(add31(x.a, x.b), sub31(y.a, y.b));
  return {add31(ac, bd), add31(sub31(k, ac), bd)};
}
M2_HD e31 sqr(e31 x) {
  const uint32_t t = mul31(x.a, x.b);
  return {mul31(add31(x.a, x.b), sub31(x.a, x.b)), add31(t, t)};
}
// The primitive 8th root of unity here is (1+i) 2^15.
M2_HD e31 mulw8(e31 x) { return {rot31(sub31(x.a, x.b), 15), rot31(add31(x.a, x.b), 15)}; }
M2_HD e31 mulw8c(e31 x) { return {rot31(add31(x.a, x.b), 15), rot31(sub31(x.b, x.a), 15)}; }
M2_HD e31 rot(e31 x, uint32_t s0, uint32_t s1) { return {rot31(x.a, s0), rot31(x.b, s1)}; }

// ---- generic butterflies over either field ----
// Forward step of radix r over pre-twisted inputs: the r-point DFT with
// bit-reversed output (evaluation of the block polynomial at tau * w_r^k).
template <class E> M2_HD void dft2(E& x0, E& x1) { E t = x0; x0 = add(t, x1); x1 = sub(t, x1); }
template <class E> M2_HD void dft4_fwd(E* v) {
  E y0 = add(v[0], v[2]), y2 = sub(v[0], v[2]);
  E y1 = add(v[1], v[3]), y3 = muli(sub(v[1], v[3]));
  v[0] = add(y0, y1); v[1] = sub(y0, y1); v[2] = add(y2, y3); v[3] = sub(y2, y3);
}
template <class E> M2_HD void dft4_inv(E* v) {
  E z0 = add(v[0], v[1]), z1 = sub(v[0], v[1]);
  E z2 = add(v[2], v[3]), z3 = mulmi(sub(v[2], v[3]));
  v[0] = add(z0, z2); v[2] = sub(z0, z2); v[1] = add(z1, z3); v[3] = sub(z1, z3);
}
template <class E> M2_HD void dft8_fwd(E* v) {
  E y0 = add(v[0], v[4]), y4 = sub(v[0], v[4]);
  E y1 = add(v[1], v[5]), y5 = sub(v[1], v[5]);
  E y2 = add(v[2], v[6]), y6 = muli(sub(v[2], v[6]));
  E y3 = add(v[3], v[7]), y7 = muli(sub(v[3], v[7]));
  E z0 = add(y0, y2), z2 = sub(y0, y2);
  E z1 = add(y1, y3), z3 = muli(sub(y1, y3));
  E z4 = add(y4, y6), z6 = sub(y4, y6);
  E z5 = mulw8(add(y5, y7)), z7 = mulw8(muli(sub(y5, y7)));
  v[0] = add(z0, z1); v[1] = sub(z0, z1);
  v[2] = add(z2, z3); v[3] = sub(z2, z3);
  v[4] = add(z4, z5); v[5] = sub(z4, z5);
  v[6] = add(z6, z7); v[7] = sub(z6, z7);
}
template <class E> M2_HD void dft8_inv(E* v) {
  E z0 = add(v[0], v[1]), z1 = sub(v[0], v[1]);
  E z2 = add(v[2], v[3]), z3 = mulmi(sub(v[2], v[3]));
  E z4 = add(v[4], v[5]), z5 = mulw8c(sub(v[4], v[5]));
  E z6 = add(v[6], v[7]), z7 = mulw8c(mulmi(sub(v[6], v[7])));
  E y0 = add(z0, z2), y2 = sub(z0, z2);
  E y1 = add(z1, z3), y3 = sub(z1, z3);
  E y4 = add(z4, z6), y6 = mulmi(sub(z4, z6));
  E y5 = add(z5, z7), y7 = mulmi(sub(z5, z7));
  v[0] = add(y0, y4); v[4] = sub(y0, y4);
  v[1] = add(y1, y5); v[5] = sub(y1, y5);
  v[2] = add(y2, y6); v[6] = sub(y2, y6);
  v[3] = add(y3, y7); v[7] = sub(y3, y7);
}

}  // namespace m2
