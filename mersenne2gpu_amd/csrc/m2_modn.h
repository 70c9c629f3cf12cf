// Canonical residues mod N = 2^p - 1 on the host, for the Gerbicz check
// (m2_gerbicz.cpp): little-endian 64-bit words, values in [0, N) (all ones,
// which is N, stands for 0 and is mapped to it). Only what the check needs:
// addition, negation and division by 4.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace m2 {
namespace modn {

struct ModN {
  uint32_t p = 0;
  size_t nw = 0;         // words
  uint64_t top = 0;      // mask of the bits of the top word below p
  explicit ModN(uint32_t p_) : p(p_), nw((p_ + 63) / 64), top(p_ % 64 ? (1ull << (p_ % 64)) - 1 : ~0ull) {}
};
using Num = std::vector<uint64_t>;

inline Num from_bytes(const std::vector<uint8_t>& b, const ModN& m) {
  Num a(m.nw, 0);
  for (size_t i = 0; i < b.size() && i / 8 < m.nw; ++i) a[i / 8] |= (uint64_t)b[i] << (8 * (i % 8));
  return a;
}
inline std::vector<uint8_t> to_bytes(const Num& a, const ModN& m) {
  std::vector<uint8_t> b((m.p + 7) / 8, 0);
  for (size_t i = 0; i < b.size(); ++i) b[i] = (uint8_t)(a[i / 8] >> (8 * (i % 8)));
  return b;
}
inline bool all_ones(const Num& a, const ModN& m) {
  for (size_t i = 0; i + 1 < m.nw; ++i)
    if (a[i] != ~0ull) return false;
  return a[m.nw - 1] == m.top;
}
// a + b for canonical a, b: the bit carried out of position p re-enters at
// bit 0 (2^p = 1 mod N); a sum of all ones is N itself, i.e. 0.
inline Num add(const Num& a, const Num& b, const ModN& m) {
  Num s(m.nw);
  unsigned __int128 c = 0;
  for (size_t i = 0; i < m.nw; ++i) {
    c += (unsigned __int128)a[i] + b[i];
    s[i] = (uint64_t)c;
    c >>= 64;
  }
  // bits at and above p: at most one (a, b < 2^p)
  const unsigned sh = m.p % 64;
  uint64_t over = sh ? (s[m.nw - 1] >> sh) : (uint64_t)c;
  if (sh) s[m.nw - 1] &= m.top;
  for (size_t i = 0; over && i < m.nw; ++i) {
    const uint64_t v = s[i] + over;
    over = v < s[i] ? 1 : 0;
    s[i] = v;
  }
  if (sh && (s[m.nw - 1] >> sh)) {  // the re-entered bit carried out again (only if an input was N)
    s[m.nw - 1] &= m.top;
    for (size_t i = 0; i < m.nw; ++i)
      if (++s[i] != 0) break;
  }
  if (all_ones(s, m)) std::fill(s.begin(), s.end(), 0);
  return s;
}
// N - a (the complement of the p bits); N - 0 = N stands for 0.
inline Num neg(const Num& a, const ModN& m) {
  Num r(m.nw);
  for (size_t i = 0; i < m.nw; ++i) r[i] = ~a[i];
  r[m.nw - 1] &= m.top;
  if (all_ones(r, m)) std::fill(r.begin(), r.end(), 0);
  return r;
}
// a / 4 = a 2^(p-2): a rotation of the p bits right by 2 (bits 0 and 1 move
// to p - 2 and p - 1).
inline Num quarter(const Num& a, const ModN& m) {
  Num r(m.nw, 0);
  for (size_t i = 0; i < m.nw; ++i) r[i] = (a[i] >> 2) | (i + 1 < m.nw ? a[i + 1] << 62 : 0);
  const uint64_t low = a[0] & 3u;
  for (uint32_t j = 0; j < 2; ++j) {
    const uint64_t pos = m.p - 2 + j;
    r[pos / 64] |= ((low >> j) & 1u) << (pos % 64);
  }
  return r;
}

}  // namespace modn
}  // namespace m2
