// CPU checks of the engine's host code (no GPU calls): plan construction and
// its digit-walk self-check, digit/carry layouts, and exact residue packing.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <random>
#include <set>
#include <vector>

#include "m2_internal.h"

using namespace m2;

static int fails = 0;
#define CHECK(c, ...)                  \
  do {                                 \
    if (!(c)) {                        \
      printf("FAIL %s: ", #c);         \
      printf(__VA_ARGS__);             \
      printf("\n");                    \
      ++fails;                         \
    }                                  \
  } while (0)

int main() {
  std::mt19937_64 rng(42);
  const uint32_t ps[] = {3, 5, 7, 127, 349, 353, 607, 1279, 19937, 20983, 21001, 132049, 2424827, 27899687};
  const uint32_t caps[] = {0, 4, 16, 64, 256};
  // (carry_pairs, max_radix, centre workgroup size)
  const uint32_t variants[][3] = {{0, 8, 0}, {16, 8, 0}, {0, 4, 0}, {4, 4, 0}, {0, 8, 1024}};
  for (uint32_t p : ps) {
    for (uint32_t cap : caps) for (const auto& var : variants) {
      DevPlan P;
      ClassTables CT;
      DigitGen D;
      std::string err;
      PlanOpts po;
      po.ls_cap = cap;
      po.carry_pairs = var[0];
      po.max_radix = var[1];
      po.threads = var[2];
      const bool okc = build_class_plan(transform_log2(p), po, P, CT, err);
      CHECK(okc, "p=%u cap=%u class plan: %s", p, cap, err.c_str());
      if (!okc) continue;
      CHECK(CT.sq61.size() == P.h / 2 && CT.sq31.size() == P.h / 2, "p=%u centre table size", p);
      const bool ok = make_digit_gen(p, P, D, err);
      if (!ok) {
        // only the multi-pass carry bound (q >= 12) may refuse a plan
        CHECK(err.find("12 bits") != std::string::npos, "p=%u cap=%u: %s", p, cap, err.c_str());
        continue;
      }
      const uint64_t h = P.h, n = 2 * h;
      // layout is a permutation
      std::vector<char> seen(h, 0);
      for (uint64_t e = 0; e < h; ++e) {
        const uint64_t s = digit_slot(P, e);
        CHECK(s < h && !seen[s], "p=%u cap=%u slot %llu", p, cap, (unsigned long long)s);
        if (s < h) seen[s] = 1;
      }
      if (!P.small) {
        std::set<uint64_t> tgt;
        for (uint64_t g = 0; g < P.nseg; ++g) tgt.insert(carry_target_digit(P, g));
        CHECK(tgt.size() == P.nseg, "p=%u carry targets not distinct", p);
        CHECK(tgt.count(0) == 1, "p=%u no carry wraps to digit 0", p);
      }
      // residue round trip on random digits, including negatives and carries
      // every pass's steps stay within the kernels' step table, and cover the pass's levels
      for (const Pass* ps : {&P.outer, &P.mid, &P.centre}) {
        if (ps->A == 0) continue;
        CHECK(ps->nsteps >= 1 && ps->nsteps <= MAX_STEPS, "p=%u steps %d", p, ps->nsteps);
        int lv = 0;
        for (int k = 0; k < ps->nsteps; ++k) {
          CHECK(ps->steps[k].radix <= (int)(var[1]), "p=%u radix %d", p, ps->steps[k].radix);
          lv += ps->steps[k].radix == 8 ? 3 : (ps->steps[k].radix == 4 ? 2 : 1);
        }
        CHECK(lv == ps->A, "p=%u levels %d != %d", p, lv, ps->A);
      }
      if (!P.small) CHECK((1 << P.logW) >= SEG_ELEMS && (P.outer.A + P.logW) <= 11, "p=%u outer tile", p);
      if (!P.small && var[2]) {
        for (int k = 0; k < P.centre.nsteps; ++k) CHECK(P.centre.steps[k].radix == 4, "p=%u wide centre", p);
        CHECK(P.centre_nt == MP_THREADS_WIDE, "p=%u centre size", p);
      }
      if (cap == 0 && var[0] == 0 && var[1] == 8 && var[2] == 0 && n <= (1u << 16)) {
        auto width = [&](uint64_t j) { return digit_bitpos(p, P.ln, j + 1) - digit_bitpos(p, P.ln, j); };
        std::vector<int64_t> d(n);
        for (uint64_t j = 0; j < n; ++j) d[j] = (int64_t)(rng() % (1ull << width(j)));
        std::vector<uint8_t> r = residue_from_state(p, P.ln, d.data(), n, nullptr, nullptr, 0);
        std::vector<int64_t> d2(n);
        digits_from_residue(p, P.ln, r.data(), r.size(), d2.data());
        std::vector<uint8_t> r2 = residue_from_state(p, P.ln, d2.data(), n, nullptr, nullptr, 0);
        CHECK(r == r2, "p=%u residue round trip", p);
        // moving one unit between digit j and the next is value-neutral:
        // d_j += 2^w_j, d_{j+1} -= 1  (here on j = 1, so digit 2 may go negative)
        std::vector<int64_t> d3 = d2;
        d3[1] += (int64_t)1 << width(1);
        d3[2] -= 1;
        CHECK(residue_from_state(p, P.ln, d3.data(), n, nullptr, nullptr, 0) == r, "p=%u carry move", p);
        // a pending carry at digit 0 of 1 equals adding 1 to digit 0
        Carry128 c{1, 0};
        uint64_t at = 0;
        std::vector<int64_t> d4 = d2;
        d4[0] += 1;
        CHECK(residue_from_state(p, P.ln, d2.data(), n, &c, &at, 1) ==
                  residue_from_state(p, P.ln, d4.data(), n, nullptr, nullptr, 0),
              "p=%u pending carry", p);
        // a digit far above its width (summed term by term) equals the same
        // amount as a pending carry at that digit (digits streamed)
        std::vector<int64_t> d5 = d2;
        d5[3] += (int64_t)1 << 50;
        Carry128 c50{uint64_t(1) << 50, 0};
        uint64_t at3 = 3;
        CHECK(residue_from_state(p, P.ln, d5.data(), n, nullptr, nullptr, 0) ==
                  residue_from_state(p, P.ln, d2.data(), n, &c50, &at3, 1),
              "p=%u large digit", p);
      }
    }
  }
  // Strided passes hold at most 2^PASS_MAX_LOG_R rows and PASS_MAX_ELEMS
  // elements per workgroup (their LDS arrays are sized for that); a local
  // stage cap that would need taller passes is refused, whatever carry_pairs.
  for (uint32_t cp : {0u, 16u}) {
    const struct { int ln; uint32_t cap; bool ok; } cases[] = {{22, 4, false}, {23, 8, false}, {22, 8, true}};
    for (const auto& c : cases) {
      DevPlan P;
      ClassTables CT;
      std::string err;
      PlanOpts po;
      po.ls_cap = c.cap;
      po.carry_pairs = cp;
      const bool ok = build_class_plan(c.ln, po, P, CT, err);
      CHECK(ok == c.ok, "ln=%d cap=%u pairs=%u: %s", c.ln, c.cap, cp, err.c_str());
      if (!ok) {
        CHECK(err.find("too tall") != std::string::npos, "ln=%d cap=%u: %s", c.ln, c.cap, err.c_str());
        continue;
      }
      CHECK(P.outer.A <= PASS_MAX_LOG_R && (1 << (P.outer.A + P.logW)) <= PASS_MAX_ELEMS, "ln=%d outer tile", c.ln);
      if (P.has_mid)
        CHECK(P.mid.A <= PASS_MAX_LOG_R && (1 << (P.mid.A + P.mid_logW)) <= PASS_MAX_ELEMS, "ln=%d mid tile", c.ln);
    }
  }
  // 1024-thread centre: radix-4 centre steps, radix-8 strided passes, and
  // per-workgroup register records holding the same squaring twiddles as the
  // 512-thread layout (each workgroup squares the same pairs; only the thread
  // that takes a pair differs)
  {
    DevPlan Pn, Pw;
    ClassTables Cn, Cw;
    std::string err;
    PlanOpts pn, pw;
    pw.threads = MP_THREADS_WIDE;
    const bool okn = build_class_plan(20, pn, Pn, Cn, err);
    const bool okw = build_class_plan(20, pw, Pw, Cw, err);
    CHECK(okn && okw, "ln=20 plans: %s", err.c_str());
    if (okn && okw) {
      CHECK(Pn.centre_nt == MP_THREADS_NARROW && Pw.centre_nt == MP_THREADS_WIDE, "centre sizes");
      for (int k = 0; k < Pw.centre.nsteps; ++k) CHECK(Pw.centre.steps[k].radix == 4, "wide centre radix");
      for (int k = 0; k < Pw.outer.nsteps; ++k)
        CHECK(Pw.outer.steps[k].radix == Pn.outer.steps[k].radix, "outer steps unchanged");
      const CentreSlots cn = centre_slots(MP_THREADS_NARROW), cw = centre_slots(MP_THREADS_WIDE);
      auto early = [](const DevPlan& P) {
        uint32_t nE = 0;
        for (int si = 0; si + 1 < P.centre.nsteps; ++si)
          nE += 2u * (1u << (P.centre.steps[si].ell - P.centre.ell0)) * (uint32_t)(P.centre.steps[si].radix - 1);
        return nE;
      };
      const uint32_t nEn = early(Pn), nEw = early(Pw);
      CHECK(Pw.crec == nEw + (uint32_t)((cw.step_slots + cw.sq_slots) * cw.role_threads), "wide record size");
      const uint32_t ncentre = (Pw.h >> Pw.centre.A) / 2;
      CHECK(Cw.ctw61.size() == (size_t)ncentre * Pw.crec && Cw.ctw31.size() == Cw.ctw61.size(), "wide records");
      for (uint32_t g : {0u, 1u, 2u, 77u, ncentre - 1}) {
        std::vector<std::pair<uint64_t, uint64_t>> a, b;
        std::vector<uint64_t> a31, b31;
        const size_t sn = (size_t)g * Pn.crec + nEn + (size_t)cn.step_slots * cn.role_threads;
        const size_t sw = (size_t)g * Pw.crec + nEw + (size_t)cw.step_slots * cw.role_threads;
        for (int i = 0; i < cn.sq_slots * cn.role_threads; ++i) {
          a.push_back({Cn.ctw61[sn + i].a, Cn.ctw61[sn + i].b});
          a31.push_back((uint64_t)Cn.ctw31[sn + i].a << 32 | Cn.ctw31[sn + i].b);
        }
        for (int i = 0; i < cw.sq_slots * cw.role_threads; ++i) {
          b.push_back({Cw.ctw61[sw + i].a, Cw.ctw61[sw + i].b});
          b31.push_back((uint64_t)Cw.ctw31[sw + i].a << 32 | Cw.ctw31[sw + i].b);
        }
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        std::sort(a31.begin(), a31.end());
        std::sort(b31.begin(), b31.end());
        CHECK(a == b && a31 == b31, "g=%u squaring twiddles differ between centre sizes", g);
      }
    }
    PlanOpts bad;
    bad.threads = 256;
    CHECK(!build_class_plan(20, bad, Pw, Cw, err) && err.find("workgroup size") != std::string::npos,
          "workgroup size 256 refused: %s", err.c_str());
  }
  // Digit walk and segment metadata against the closed forms (make_digit_gen's
  // self-check: every digit up to n = 2^16, windows above) for random odd
  // exponents at every transform length 2^3..2^27, plus the extremes of
  // each bit length; the plan flag only relaxes the multi-pass width bound.
  {
    int lens = 0;
    std::set<int> seen_ln;
    for (int bits = 2; bits <= 32; ++bits) {
      std::vector<uint32_t> cand = {(uint32_t)((1ull << (bits - 1)) | 1u), (uint32_t)((1ull << bits) - 1)};
      for (int k = 0; k < 6; ++k) cand.push_back((uint32_t)((1ull << (bits - 1)) | (rng() & ((1ull << (bits - 1)) - 1)) | 1u));
      // p mod n a multiple of 61, of 31 and of both: the walk's step is 0
      // in that field and only the wraps move its counts
      if (bits >= 8) {
        const uint32_t p0 = cand.back();
        const uint64_t n = 1ull << transform_log2(p0);
        for (uint64_t f : {61ull, 31ull, 61ull * 31ull})
          for (uint64_t k = 1; f * k < n; k += 2 + 2 * (rng() % 8)) {
            cand.push_back((uint32_t)((p0 & ~(n - 1)) | (f * k)));
            if (k > 16) break;
          }
      }
      for (uint32_t p : cand) {
        if (p < 3 || (p & 1u) == 0) continue;
        const int ln = transform_log2(p);
        if (ln > 27) continue;
        DevPlan P{};
        P.ln = ln;
        P.small = true;
        DigitGen D;
        std::string e;
        CHECK(make_digit_gen(p, P, D, e), "p=%u ln=%d digit walk / segment metadata: %s", p, ln, e.c_str());
        seen_ln.insert(ln);
        ++lens;
      }
    }
    CHECK(seen_ln.size() == 25 && *seen_ln.begin() == 3 && *seen_ln.rbegin() == 27,
          "transform lengths covered: %zu", seen_ln.size());
    (void)lens;
  }
  printf("%s (%d failures)\n", fails ? "FAILED" : "OK", fails);
  return fails ? 1 : 0;
}
