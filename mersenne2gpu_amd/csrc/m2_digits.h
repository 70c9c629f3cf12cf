// Digit metadata walked on the fly (see DigitGen in m2_plan.h); shared by the
// kernels and the host-side self-check in build_plan.
#pragma once
#include "m2_arith.h"
#include "m2_plan.h"

namespace m2 {

struct DWalk {
  uint32_t r, v61, v31;
};

M2_HD DWalk dwalk_at(const DigitGen& D, uint32_t j) {
  DWalk w;
  w.r = (D.p * j) & D.nmask;  // low 32 bits of p j suffice: n <= 2^27
  const uint32_t m = D.nmask + 1u - w.r;
  w.v61 = (D.lr61 * (m % 61u)) % 61u;
  w.v31 = (D.lr31 * (m % 31u)) % 31u;
  return w;
}
M2_HD uint32_t dwalk_width(const DigitGen& D, const DWalk& w) {
  return D.q + ((w.r + D.s > D.nmask + 1u) ? 1u : 0u) + (w.r == 0 ? 1u : 0u);
}
M2_HD uint32_t dwalk_w61(const DWalk& w) { return w.r ? w.v61 : 0u; }
M2_HD uint32_t dwalk_w31(const DWalk& w) { return w.r ? w.v31 : 0u; }
M2_HD void dwalk_next(const DigitGen& D, DWalk& w) {
  const uint32_t t = w.r + D.s;
  const uint32_t wrap = (t > D.nmask) ? 1u : 0u;
  w.r = t & D.nmask;
  // a in [1, 122], b in [1, 62]: two unsigned-min steps (a - 61 wraps to a
  // large value when a < 61) leave them in [0, 61) and [0, 31)
  uint32_t a = w.v61 + 61u - D.d61 + wrap;
  a = umin32(a, a - 61u);
  a = umin32(a, a - 61u);
  w.v61 = a;
  uint32_t b = w.v31 + 31u - D.d31 + wrap;
  b = umin32(b, b - 31u);
  b = umin32(b, b - 31u);
  w.v31 = b;
}

// Digit metadata of one carry segment (the 8 digits from j0), walked once and
// kept in registers for the segment's normalise, ripple and weight steps. Per
// digit, one byte each: its weight shifts w61 (bits 0-7) and w31 (8-15), and
// a61 = (w61 + ln + 1) mod 61 (16-23) and a31 = (w31 + ln + 1) mod 31 (24-30),
// the right rotations that unweight it and also remove 1/h and the centre's
// factor 4; bit 31: its width is q + 1 (else q). A count may also hold its
// modulus for 0 (a rotation by 61 or 31 is the identity).
//
// From digit j to j + 1 (r = p j mod n, s = p mod n) every count moves by the
// same uniform step, -(lr s) mod 61 or 31, plus 1 when r + s wraps past n, so
// the four counts advance together as the bytes of one word: one add, then
// one range correction for all of them (subtract the modulus from every byte
// that exceeds it). The wrap is also the width bit, except at the two digits
// where r + s meets n exactly or r = 0, i.e. j = n - 1 and j = 0 (p is odd),
// which are set apart.
struct SegMeta {
  uint32_t m[2 * SEG_ELEMS];
};
M2_HD SegMeta seg_meta(const DigitGen& D, uint32_t j0) {
  constexpr uint32_t ONES = 0x01010101u, HIGH = 0x80808080u;
  constexpr uint32_t MOD = 61u | (31u << 8) | (61u << 16) | (31u << 24);
  // per byte 128 - (modulus + 1): bit 7 of a byte of S + OVER is set iff the
  // count in it exceeds its modulus
  constexpr uint32_t OVER = (128u - 62u) | ((128u - 32u) << 8) | ((128u - 62u) << 16) | ((128u - 32u) << 24);
  const uint32_t c61 = D.d61 ? 61u - D.d61 : 0u, c31 = D.d31 ? 31u - D.d31 : 0u;
  const uint32_t C = c61 | (c31 << 8) | (c61 << 16) | (c31 << 24);
  const DWalk w = dwalk_at(D, j0);
  const uint32_t a61 = umin32(w.v61 + D.k61, w.v61 + D.k61 - 61u);
  const uint32_t a31 = umin32(w.v31 + D.k31, w.v31 + D.k31 - 31u);
  uint32_t S = w.v61 | (w.v31 << 8) | (a61 << 16) | (a31 << 24);  // bytes <= 60, 30, 60, 30
  uint32_t r = w.r;
  SegMeta M;
#pragma unroll
  for (int i = 0; i < 2 * SEG_ELEMS; ++i) {
    const uint32_t t = r + D.s;
    const uint32_t f = t > D.nmask ? ONES : 0u;
    M.m[i] = S | (f << 31);
    if (i + 1 < 2 * SEG_ELEMS) {
      r = t & D.nmask;
      S += C + f;  // bytes <= 122, 62, 122, 62: no carry between them
      const uint32_t g = (S + OVER) & HIGH;
      S -= ((g << 1) - (g >> 7)) & MOD;  // (g << 1) - (g >> 7): 0xFF in every flagged byte
    }
  }
  // digit 0: weight 1 and width q + 1; digit n - 1: width q
  if (j0 == 0u) M.m[0] = (D.k61 << 16) | (D.k31 << 24) | 0x80000000u;
  if (j0 + 2u * SEG_ELEMS == D.nmask + 1u) M.m[2 * SEG_ELEMS - 1] &= 0x7FFFFFFFu;
  return M;
}
M2_HD uint32_t meta_w61(uint32_t m) { return m & 0xFFu; }
M2_HD uint32_t meta_w31(uint32_t m) { return (m >> 8) & 0xFFu; }
// the unweighting rotations as left rotations (61 - a61, 31 - a31)
M2_HD uint32_t meta_u61(uint32_t m) { return 61u - ((m >> 16) & 0xFFu); }
M2_HD uint32_t meta_u31(uint32_t m) { return 31u - ((m >> 24) & 0x7Fu); }
M2_HD uint32_t meta_wide(uint32_t m) { return m >> 31; }

// Normalisation with 64-bit words, for digit widths w = q or q + 1 with
// 16 <= q <= 31 (so w <= 32). Such tests have no negative convolution term
// (every digit is non-negative and the first one at least 17 bits wide), so
// Garner's u (crt61_31) is read unsigned: with the M31 residue first made
// canonical (P31, its second 0, gives u = P61 - 1 for the term 0 otherwise),
// u is in [0, P61) or is P61, the second representation of 0, read as 0.
// The term x = u P31 + t + add is u0 P31 + (t + add) plus u1 P31 2^32
// (u = u1 2^32 + u0): two v_mad_u64_u32. The running carry is
// kept as cA + cB 2^32; with A = u0 P31 + t + add + cA and B = u1 P31 + cB the
// sum is A + B 2^32, its low w bits are A's (w <= 32), and the carry out is
//   (A >> w) + (B mod 2^w) 2^(32 - w)  +  (B >> w) 2^32.
// Bounds: cA < 2^(64-w) + 2^32 <= 2^48 + 2^32 and cB < 2^47, so A < 2^63 +
// 2^49 and B < 2^60 + 2^47 (u0 < 2^32, u1 < 2^29, t < 2^31, add < 2^32); a
// carry-in in [-2, 0) is cA = 2^64 - 2 or 2^64 - 1, and A, computed modulo
// 2^64, is the true non-negative sum as long as t + add + carry-in >= 0.
M2_HD bool norm32_applies(uint32_t q) { return q >= 16u && q <= 31u; }
M2_HD uint32_t norm_digit32(uint64_t s, uint32_t t, uint32_t meta, uint32_t q, uint64_t add, uint64_t& cA,
                            uint64_t& cB) {
  uint32_t tr = rot31(t, meta_u31(meta));
  tr = umin32(tr, tr - P31);  // P31 -> 0: u is then the unique one in [0, P61)
  // u = (s - t)(2^31 + 1) = P31^-1 (s - t) (mod P61): with v = s + (P61 - t)
  // <= 2 P61 (v1 < 2^30), v (2^31 + 1) = v + v0 2^31 + 4 v1 (2^63 = 4), a sum
  // below 2^62 + 2^63 + 2^32 that two folds reduce
  const uint64_t v = rot61(s, meta_u61(meta)) + neg61(uint64_t(tr));
  uint64_t u = uint64_t(uint32_t(v)) * opaque_s(0x80000000u) + v;
  u = uint64_t(uint32_t(v >> 32)) * opaque_s(4u) + u;
  u = fold61(fold61(u));
  u = (u == P61) ? 0ull : u;
  const uint32_t w = q + meta_wide(meta);
  const uint32_t mask = 0xFFFFFFFFu >> (32u - w);
  const uint64_t A = uint64_t(uint32_t(u)) * P31 + (uint64_t(tr) + add + cA);
  const uint64_t B = uint64_t(uint32_t(u >> 32)) * P31 + cB;
  // (B mod 2^w) 2^(32 - w) < 2^32
  cA = uint64_t((uint32_t(B) & mask) << (32u - w)) * opaque_s(1u) + (A >> w);
  cB = B >> w;
  return uint32_t(A) & mask;
}
// The same for digit widths w = q or q + 1 with 32 <= q <= 43 (w = 32..44),
// with the carry in one word: with A = u0 P31 + t + add + c (< 2^63 + 2^61 +
// 2^45) and C = (A >> 32) + u1 P31 (< 2^61) the sum is (A mod 2^32) + C 2^32,
// its low w bits are A's low word and C's low w - 32 bits above it, and the
// carry out is C >> (w - 32) (< 2^61). LL's addend 2^w - 1 is below 2^44 here.
M2_HD bool norm64_applies(uint32_t q) { return q >= 32u && q <= 43u; }
M2_HD uint64_t norm_digit64(uint64_t s, uint32_t t, uint32_t meta, uint32_t q, uint64_t add, uint64_t& c) {
  uint32_t tr = rot31(t, meta_u31(meta));
  tr = umin32(tr, tr - P31);
  const uint64_t v = rot61(s, meta_u61(meta)) + neg61(uint64_t(tr));
  uint64_t u = uint64_t(uint32_t(v)) * opaque_s(0x80000000u) + v;
  u = uint64_t(uint32_t(v >> 32)) * opaque_s(4u) + u;
  u = fold61(fold61(u));
  u = (u == P61) ? 0ull : u;
  const uint32_t ws = q + meta_wide(meta) - 32u;  // 0..12
  const uint64_t A = uint64_t(uint32_t(u)) * P31 + (uint64_t(tr) + add + c);
  const uint64_t C = uint64_t(uint32_t(u >> 32)) * P31 + (A >> 32);
  c = C >> ws;
  return (uint64_t(uint32_t(C) & ~(0xFFFFFFFFu << ws)) << 32) | uint32_t(A);
}
// One digit of a carry ripple for widths 32..44: digit d < 2^w, carry c <
// 2^61 in one word.
M2_HD uint64_t ripple_digit64(uint64_t d, uint32_t meta, uint32_t q, uint64_t& c) {
  const uint32_t w = q + meta_wide(meta);
  const uint64_t A = d + c;
  c = A >> w;
  return A & ((uint64_t(1) << w) - 1u);
}
// One digit of a carry ripple in the same form: the digit d (< 2^32 here)
// plus the carry cA + cB 2^32 (cA < 2^33); returns the new digit.
M2_HD uint32_t ripple_digit32(uint64_t d, uint32_t meta, uint32_t q, uint64_t& cA, uint64_t& cB) {
  const uint32_t w = q + meta_wide(meta);
  const uint32_t mask = 0xFFFFFFFFu >> (32u - w);
  const uint64_t A = d + cA;
  cA = uint64_t((uint32_t(cB) & mask) << (32u - w)) * opaque_s(1u) + (A >> w);
  cB >>= w;
  return uint32_t(A) & mask;
}

}  // namespace m2
