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

}  // namespace m2
