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
// x y (mod P61) in [0, P61] from 32-bit limbs, one v_mad_u64_u32 per partial
// product, for x, y < 2^62: x1, y1 < 2^30, so p11 < 2^60 and p01 < 2^63, and
// the folded sum stays below 2^63 + 2^62 + 2^35.
M2_HD uint64_t mul61_limbs(uint64_t x, uint64_t y) {
  const uint32_t x0 = uint32_t(x), x1 = uint32_t(x >> 32);
  const uint32_t y0 = uint32_t(y), y1 = uint32_t(y >> 32);
  return fold_product61(uint64_t(x0) * y0, uint64_t(x0) * y1 + uint64_t(x1) * y0, uint64_t(x1) * y1);
}
M2_HD uint64_t mul61(uint64_t x, uint64_t y) {
#if defined(__HIP_DEVICE_COMPILE__)
  return mul61_limbs(x, y);
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
// x * 2^s for 0 <= s <= 61: rotate the 61-bit word left by s (s = 61 gives x).
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
// x * 2^s for x <= P31, 0 <= s <= 31 (s = 31 gives x).
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
// The same square for components <= P61, with the factors a + b, a - b
// (as a + (P61 - b)) and 2b left unreduced: each is <= 2 P61 < 2^62, which
// mul61 takes as it is.
M2_HD e61 sqr_lazy(e61 x) { return {mul61(x.a + x.b, x.a + neg61(x.b)), mul61(x.a, x.b + x.b)}; }
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

// ---- twiddle products and lazily reduced butterflies (transform steps) ----
//
// Twiddle tables hold M61 values in split form (c1 << 32) | c0 with
// c = c1 2^31 + c0, c0 < 2^31: then a c + b d for a, b < 2^62 sums its
// low, middle and high partial products in 64 bits without overflow and
// needs one fold of the low part plus two final folds.
M2_HD uint64_t split31(uint64_t c) { return ((c >> 31) << 32) | (c & 0x7FFFFFFFull); }
M2_HD uint64_t unsplit31(uint64_t w) { return ((w >> 32) << 31) | (w & 0x7FFFFFFFull); }
M2_HD e61 split31(e61 w) { return {split31(w.a), split31(w.b)}; }

constexpr uint64_t P61x2 = 2 * P61, P61x4 = 4 * P61, P61x8 = 8 * P61;

// Values hidden from the instruction combiner (no instruction on the device):
// 2 x1 of a limb x1 = x >> 32 is then one shift, not the funnel shift and mask
// that (x >> 31) & ~1 becomes, and a product by the constant 2^31 stays one
// v_mad_u64_u32 that adds into the sum instead of a 64-bit shift and two masks.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ uint32_t opaque_v(uint32_t x) { asm("" : "+v"(x)); return x; }
__device__ __forceinline__ uint32_t opaque_s(uint32_t x) { asm("" : "+s"(x)); return x; }
#else
M2_HD uint32_t opaque_v(uint32_t x) { return x; }
M2_HD uint32_t opaque_s(uint32_t x) { return x; }
#endif

// a c + b d (mod P61) in [0, P61]: a, b < 2^62; c, d in split form.
//   a c = a0 c0 + (a0 c1 + 2 a1 c0) 2^31 + a1 c1 2^63, with 2^63 = 4 and
//   2^61 = 1 (mod P61); the middle sum M contributes (M >> 30) + (M mod 2^30) 2^31.
M2_HD uint64_t dot61s(uint64_t a, uint64_t c, uint64_t b, uint64_t d) {
  const uint32_t a0 = uint32_t(a), a1 = opaque_v(uint32_t(a >> 32)), b0 = uint32_t(b), b1 = opaque_v(uint32_t(b >> 32));
  const uint32_t c0 = uint32_t(c), c1 = uint32_t(c >> 32), d0 = uint32_t(d), d1 = uint32_t(d >> 32);
  const uint64_t lo = uint64_t(a0) * c0 + uint64_t(b0) * d0;  // each < 2^63
  const uint64_t mid = uint64_t(a0) * c1 + uint64_t(2u * a1) * c0 + uint64_t(b0) * d1 + uint64_t(2u * b1) * d0;
  const uint64_t hi = uint64_t(a1) * c1 + uint64_t(b1) * d1;  // < 2^61
  const uint64_t r = fold61(lo) + (hi << 2) + (mid >> 30) +
                     uint64_t(uint32_t(mid) & 0x3FFFFFFFu) * opaque_s(0x80000000u);  // < 2^64
  return fold61(fold61(r));
}
// P61 - c in split form for c <= P61 in split form: both fields complemented
// within their widths (2^30 - 1 - c1, 2^31 - 1 - c0), i.e. two XORs.
M2_HD uint64_t negsplit31(uint64_t w) { return w ^ 0x3FFFFFFF7FFFFFFFull; }
// x w and x conj(w) for x with components < 2^62, w a split twiddle (the
// negated product is that with the negated twiddle component); results in
// [0, P61].
M2_HD e61 mul_tw(e61 x, e61 w) { return {dot61s(x.a, w.a, x.b, negsplit31(w.b)), dot61s(x.a, w.b, x.b, w.a)}; }
M2_HD e61 mulconj_tw(e61 x, e61 w) { return {dot61s(x.a, w.a, x.b, w.b), dot61s(x.b, w.a, x.a, negsplit31(w.b))}; }

// x mod P31 in [0, P31] for x < 2^63 (2^32 = 2 mod P31).
M2_HD uint32_t red31(uint64_t x) {
  const uint64_t t = (x >> 32) * 2u + (x & 0xFFFFFFFFull);  // < 2^33
  const uint32_t r = uint32_t(t & P31) + uint32_t(t >> 31);
  return umin32(r, r - P31);
}
// x mod P31 in [0, P31] for x <= 2 P31^2 (a dot product of two reduced
// pairs): with x = a 2^62 + b 2^31 + c (a <= 1; b, c < 2^31), x = a + b + c
// (mod P31), and a + b + c <= 2 P31 up to that bound (a = 1 leaves b <=
// 2^31 - 4), so one three-input add and an unsigned min reduce it, all in
// 32-bit words.
M2_HD uint32_t red31_dot(uint64_t x) {
  const uint32_t y = uint32_t(x >> 31);  // bits 31..62
  const uint32_t r = (y & P31) + (y >> 31) + (uint32_t(x) & P31);
  return umin32(r, r - P31);
}
// GF(M31^2) twiddle products as two 2-term dot products each.
M2_HD e31 mul_tw(e31 x, e31 w) {
  const uint32_t nd = P31 - w.b;
  return {red31_dot(uint64_t(x.a) * w.a + uint64_t(x.b) * nd), red31_dot(uint64_t(x.a) * w.b + uint64_t(x.b) * w.a)};
}
M2_HD e31 mulconj_tw(e31 x, e31 w) {
  const uint32_t nd = P31 - w.b;
  return {red31_dot(uint64_t(x.a) * w.a + uint64_t(x.b) * w.b), red31_dot(uint64_t(x.b) * w.a + uint64_t(x.a) * nd)};
}

// x 2^30 (mod P61) for any x < 2^64, result < 2^61 + 2^33.
M2_HD uint64_t rot30_lazy(uint64_t x) { return (x >> 31) + ((x & 0x7FFFFFFFull) << 30); }
M2_HD uint64_t red61(uint64_t x) { return fold61(fold61(x)); }  // [0, 2^64) -> [0, P61]

// Forward butterflies over GF(M61^2) with sums kept unreduced between levels
// (bounds in the comments, P = P61). Inputs: element 0 <= P + 7 (the output
// of a previous forward step), the others <= P (twiddle products). Outputs
// are folded once: <= P + 7.
M2_HD void dft2_fwd_lazy(e61& x0, e61& x1) {
  const e61 t = x0;
  x0 = {fold61(t.a + x1.a), fold61(t.b + x1.b)};
  x1 = {fold61(t.a + neg61(x1.a)), fold61(t.b + neg61(x1.b))};
}
M2_HD void dft4_fwd_lazy(e61* v) {
  const uint64_t y0a = v[0].a + v[2].a, y0b = v[0].b + v[2].b;                    // <= 2P + 7
  const uint64_t y2a = v[0].a + neg61(v[2].a), y2b = v[0].b + neg61(v[2].b);
  const uint64_t y1a = v[1].a + v[3].a, y1b = v[1].b + v[3].b;                    // <= 2P
  const uint64_t y3a = v[3].b + neg61(v[1].b), y3b = v[1].a + neg61(v[3].a);    // i (v1 - v3)
  v[0] = {fold61(y0a + y1a), fold61(y0b + y1b)};
  v[1] = {fold61(y0a + (P61x2 - y1a)), fold61(y0b + (P61x2 - y1b))};
  v[2] = {fold61(y2a + y3a), fold61(y2b + y3b)};
  v[3] = {fold61(y2a + (P61x2 - y3a)), fold61(y2b + (P61x2 - y3b))};
}
M2_HD void dft8_fwd_lazy(e61* v) {
  // level 1
  const uint64_t y0a = v[0].a + v[4].a, y0b = v[0].b + v[4].b;                    // <= 2P + 7
  const uint64_t y4a = v[0].a + neg61(v[4].a), y4b = v[0].b + neg61(v[4].b);
  const uint64_t y1a = v[1].a + v[5].a, y1b = v[1].b + v[5].b;                    // <= 2P
  const uint64_t y5a = v[1].a + neg61(v[5].a), y5b = v[1].b + neg61(v[5].b);
  const uint64_t y2a = v[2].a + v[6].a, y2b = v[2].b + v[6].b;
  const uint64_t y6a = v[6].b + neg61(v[2].b), y6b = v[2].a + neg61(v[6].a);    // i (v2 - v6)
  const uint64_t y3a = v[3].a + v[7].a, y3b = v[3].b + v[7].b;
  const uint64_t y7a = v[7].b + neg61(v[3].b), y7b = v[3].a + neg61(v[7].a);    // i (v3 - v7)
  // level 2
  const uint64_t z0a = y0a + y2a, z0b = y0b + y2b;                                // <= 4P + 7
  const uint64_t z2a = y0a + (P61x2 - y2a), z2b = y0b + (P61x2 - y2b);
  const uint64_t z1a = y1a + y3a, z1b = y1b + y3b;                                // <= 4P
  const uint64_t z3a = y3b + (P61x2 - y1b), z3b = y1a + (P61x2 - y3a);            // i (y1 - y3)
  const uint64_t z4a = y4a + y6a, z4b = y4b + y6b;
  const uint64_t z6a = y4a + (P61x2 - y6a), z6b = y4b + (P61x2 - y6b);
  const uint64_t sa = y5a + y7a, sb = y5b + y7b;                                  // <= 4P
  const uint64_t z5a = rot30_lazy(sb + (P61x4 - sa)), z5b = rot30_lazy(P61x8 - sa - sb);  // w8 s
  const uint64_t ta = y7b + (P61x2 - y5b), tb = y5a + (P61x2 - y7a);              // i (y5 - y7)
  const uint64_t z7a = rot30_lazy(tb + (P61x4 - ta)), z7b = rot30_lazy(P61x8 - ta - tb);  // w8 t
  // level 3 (sums < 2^64)
  v[0] = {fold61(z0a + z1a), fold61(z0b + z1b)};
  v[1] = {fold61(z0a + (P61x4 - z1a)), fold61(z0b + (P61x4 - z1b))};
  v[2] = {fold61(z2a + z3a), fold61(z2b + z3b)};
  v[3] = {fold61(z2a + (P61x4 - z3a)), fold61(z2b + (P61x4 - z3b))};
  v[4] = {fold61(z4a + z5a), fold61(z4b + z5b)};
  v[5] = {fold61(z4a + (P61x2 - z5a)), fold61(z4b + (P61x2 - z5b))};
  v[6] = {fold61(z6a + z7a), fold61(z6b + z7b)};
  v[7] = {fold61(z6a + (P61x2 - z7a)), fold61(z6b + (P61x2 - z7b))};
}
// Inverse butterflies: inputs <= P; output 0 fully reduced (<= P), the others
// folded once (<= P + 7, ready for mulconj_tw; one more fold reduces them).
M2_HD void dft2_inv_lazy(e61& x0, e61& x1) {
  const e61 t = x0;
  x0 = {fold61(t.a + x1.a), fold61(t.b + x1.b)};
  x1 = {fold61(t.a + neg61(x1.a)), fold61(t.b + neg61(x1.b))};
}
M2_HD void dft4_inv_lazy(e61* v) {
  const uint64_t z0a = v[0].a + v[1].a, z0b = v[0].b + v[1].b;                    // <= 2P
  const uint64_t z1a = v[0].a + neg61(v[1].a), z1b = v[0].b + neg61(v[1].b);
  const uint64_t z2a = v[2].a + v[3].a, z2b = v[2].b + v[3].b;
  const uint64_t z3a = v[2].b + neg61(v[3].b), z3b = v[3].a + neg61(v[2].a);    // -i (v2 - v3)
  v[0] = {red61(z0a + z2a), red61(z0b + z2b)};
  v[2] = {fold61(z0a + (P61x2 - z2a)), fold61(z0b + (P61x2 - z2b))};
  v[1] = {fold61(z1a + z3a), fold61(z1b + z3b)};
  v[3] = {fold61(z1a + (P61x2 - z3a)), fold61(z1b + (P61x2 - z3b))};
}
M2_HD void dft8_inv_lazy(e61* v) {
  // level 1
  const uint64_t z0a = v[0].a + v[1].a, z0b = v[0].b + v[1].b;                    // <= 2P
  const uint64_t z1a = v[0].a + neg61(v[1].a), z1b = v[0].b + neg61(v[1].b);
  const uint64_t z2a = v[2].a + v[3].a, z2b = v[2].b + v[3].b;
  const uint64_t z3a = v[2].b + neg61(v[3].b), z3b = v[3].a + neg61(v[2].a);    // -i (v2 - v3)
  const uint64_t z4a = v[4].a + v[5].a, z4b = v[4].b + v[5].b;
  const uint64_t da = v[4].a + neg61(v[5].a), db = v[4].b + neg61(v[5].b);
  const uint64_t z5a = rot30_lazy(P61x4 - da - db), z5b = rot30_lazy(da + (P61x2 - db));  // conj(w8) d
  const uint64_t z6a = v[6].a + v[7].a, z6b = v[6].b + v[7].b;
  const uint64_t ma = v[6].b + neg61(v[7].b), mb = v[7].a + neg61(v[6].a);      // -i (v6 - v7)
  const uint64_t z7a = rot30_lazy(P61x4 - ma - mb), z7b = rot30_lazy(ma + (P61x2 - mb));
  // level 2 (z5, z7 < 2^61 + 2^33)
  const uint64_t y0a = z0a + z2a, y0b = z0b + z2b;                                // <= 4P
  const uint64_t y2a = z0a + (P61x2 - z2a), y2b = z0b + (P61x2 - z2b);
  const uint64_t y1a = z1a + z3a, y1b = z1b + z3b;
  const uint64_t y3a = z1a + (P61x2 - z3a), y3b = z1b + (P61x2 - z3b);
  const uint64_t y4a = z4a + z6a, y4b = z4b + z6b;
  const uint64_t y6a = z4b + (P61x2 - z6b), y6b = z6a + (P61x2 - z4a);            // -i (z4 - z6)
  const uint64_t y5a = z5a + z7a, y5b = z5b + z7b;                                // < 2P
  const uint64_t y7a = z5b + (P61x2 - z7b), y7b = z7a + (P61x2 - z5a);            // -i (z5 - z7)
  // level 3 (sums <= 8P)
  v[0] = {red61(y0a + y4a), red61(y0b + y4b)};
  v[4] = {fold61(y0a + (P61x4 - y4a)), fold61(y0b + (P61x4 - y4b))};
  v[1] = {fold61(y1a + y5a), fold61(y1b + y5b)};
  v[5] = {fold61(y1a + (P61x4 - y5a)), fold61(y1b + (P61x4 - y5b))};
  v[2] = {fold61(y2a + y6a), fold61(y2b + y6b)};
  v[6] = {fold61(y2a + (P61x4 - y6a)), fold61(y2b + (P61x4 - y6b))};
  v[3] = {fold61(y3a + y7a), fold61(y3b + y7b)};
  v[7] = {fold61(y3a + (P61x4 - y7a)), fold61(y3b + (P61x4 - y7b))};
}

// ---- 92-bit CRT (Garner) ----
//
// x from its residues s = x mod P61 and t = x mod P31 (either may be in its
// second representation of 0, P61 or P31): with P31^-1 = 2^31 + 1 (mod P61),
// u = (s - t)(2^31 + 1) mod P61 and x = u P31 + t in [0, M), M = P61 P31
// (the reference's unsigned reading, mersenne2_gpu.cpp:2018-2030).
//
// The squaring's convolution terms are non-negative as long as every digit is
// (PRP always; LL with the -2 applied as described at the carry), and with
// digits d_j in [0, 2^width_j] each term is at most 2^(ln+2q+3) - 2^(2q+2)
// <= 2^92 - 2^64 for ln <= 27 (q = p >> ln, ln + 2q + 2 <= 91): every pair
// d_i d_j 2^e of the weighted cyclic convolution has width_i, width_j <= q + 1
// and e in {0, 1}, and the two pairs with i = 0 or j = 0 have e = 0. Such a
// term has u = floor(x / P31) <= P61 - 7 2^30 + 1. Values with
// u >= P61 - 2^30 are read as x - M = (u - P61) P31 + t: negative terms down
// to -2^60, whichever representation of 0 a residue has (only the smallest
// exponents, whose first digit is 1 bit wide, make any). `add` (< 2^63) is added to the result: the carry step
// adds each LL digit's all-ones value there.
constexpr uint64_t CRT_NEG_U = P61 - (uint64_t(1) << 30);
M2_HD __int128 crt61_31(uint64_t s, uint32_t t, uint64_t add = 0) {
  uint64_t u = sub61(s, t);
  u = add61(u, rot61(u, 31));  // * (2^31 + 1)
  const int64_t v = (int64_t)u - ((u >= CRT_NEG_U) ? (int64_t)P61 : 0);
  // v 2^31 - v + t (+ add) as a 128-bit two's-complement integer
  return ((__int128)v << 31) - (__int128)v + (__int128)((uint64_t)t + add);
}

// Real-data pairing and squaring of one pair of spectrum values (the
// reference's square_half, mersenne2_gpu.cpp:856-870): with
// u0 = z_k + conj(z_mk), u1 = z_k - conj(z_mk), v0 = u0^2 - u1^2 tw,
// v1 = 2 u0 u1, it sets z_k <- v0 + v1, z_mk <- conj(v0 - v1). Since
// u0 + u1 = 2 z_k and u0 - u1 = 2 conj(z_mk), the same values are
//   z_k <- 4 z_k^2 - C,  z_mk <- 4 z_mk^2 - conj(C),  C = u1^2 (1 + tw),
// three squarings and one twiddle product; tw1 = 1 + tw (split form for M61).
M2_HD e61 red(e61 x) { return {fold61(x.a), fold61(x.b)}; }  // <= P61 + 7 -> <= P61
M2_HD e31 red(e31 x) { return x; }
M2_HD e61 times4(e61 x) { return {rot61(x.a, 2), rot61(x.b, 2)}; }
M2_HD e31 times4(e31 x) { return {rot31(x.a, 2), rot31(x.b, 2)}; }
template <class E>
M2_HD void square_pair(E& zk, E& zmk, E tw1) {
  const E a = red(zk), b = red(zmk);
  const E u1 = sub(a, conj(b));
  const E C = mul_tw(sqr(u1), tw1);
  zk = sub(times4(sqr(a)), C);
  zmk = sub(times4(sqr(b)), conj(C));
}
M2_HD void square_pair(e61& zk, e61& zmk, e61 tw1) {
  const e61 a = red(zk), b = red(zmk);
  const e61 u1 = sub(a, conj(b));
  const e61 C = mul_tw(sqr_lazy(u1), tw1);
  zk = sub(times4(sqr_lazy(a)), C);
  zmk = sub(times4(sqr_lazy(b)), conj(C));
}


}  // namespace m2
