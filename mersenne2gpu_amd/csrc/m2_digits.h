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

}  // namespace m2
