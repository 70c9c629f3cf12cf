// Host side of the engine: transform size, digit tables, twiddle tables, the
// pass plan, and exact residue arithmetic on the host.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <sstream>

#include "m2_digits.h"
#include "m2_internal.h"

namespace m2 {

// Smallest ln >= 3 with ln + 2((p >> ln) + 1) < 92, so that n (2^(q+1))^2 stays
// below M61*M31 (the reference's transformsize, mersenne2_gpu.cpp:286-293).
int transform_log2(uint64_t p) {
  int ln = 3;
  while ((uint64_t)ln + 2u * ((p >> ln) + 1u) >= 92u) ++ln;
  return ln;
}

namespace {

// 2-Sylow generators: order 2^62 in GF(M61^2), 2^32 in GF(M31^2). They are the
// reference's constants (mersenne2_gpu.cpp:174, :209), kept so that spectra
// match the oracle element by element.
const e61 G61 = {264036120304204ull, 4677669021635377ull};
const e31 G31 = {7735u, 748621u};

template <class E>
E gpow(E x, uint64_t e) {
  E r;
  r.a = 1;
  r.b = 0;
  while (e) {
    if (e & 1) r = mul(r, x);
    x = mul(x, x);
    e >>= 1;
  }
  return r;
}
e61 root61(int k) { return gpow(G61, uint64_t(1) << (62 - k)); }  // order 2^k
e31 root31(int k) { return gpow(G31, uint64_t(1) << (32 - k)); }

uint32_t bitrev(uint32_t j, int bits) {
  uint32_t r = 0;
  for (int b = 0; b < bits; ++b) r |= ((j >> b) & 1u) << (bits - 1 - b);
  return r;
}

int log2i(uint32_t r) { return r == 8 ? 3 : (r == 4 ? 2 : 1); }

// Levels [ell0, ell0 + A). max_radix 8: one radix-2 or radix-4 step for
// A mod 3 levels, then radix-8 steps; max_radix 4: one radix-2 step if A is
// odd, then radix-4 steps. Twiddle blocks are appended to tw61/tw31.
Pass make_pass(int ell0, int A, std::vector<e61>& tw61, std::vector<e31>& tw31, int max_radix = 8) {
  Pass ps{};
  ps.ell0 = ell0;
  ps.A = A;
  int ell = ell0;
  auto add_step = [&](int radix) {
    Step& s = ps.steps[ps.nsteps++];
    s.ell = ell;
    s.radix = radix;
    s.twoff = (uint32_t)tw61.size();
    const int lr = log2i(radix);
    // tau_j = zeta of block 2^lr * j at level ell + lr = w_{2^(ell+lr)}^bitrev(j)
    const uint32_t nb = 1u << ell;
    const e61 w61 = root61(ell + lr);
    const e31 w31 = root31(ell + lr);
    std::vector<e61> p61(nb);
    std::vector<e31> p31(nb);
    e61 c61{1, 0};
    e31 c31{1, 0};
    for (uint32_t k = 0; k < nb; ++k) {
      p61[k] = c61;
      p31[k] = c31;
      c61 = mul(c61, w61);
      c31 = mul(c31, w31);
    }
    for (uint32_t j = 0; j < nb; ++j) {
      const uint32_t br = bitrev(j, ell);
      const e61 t61 = p61[br];
      const e31 t31 = p31[br];
      e61 a61 = t61;
      e31 a31 = t31;
      for (int e = 1; e < radix; ++e) {
        tw61.push_back(a61);
        tw31.push_back(a31);
        a61 = mul(a61, t61);
        a31 = mul(a31, t31);
      }
    }
    ell += lr;
  };
  if (max_radix == 4) {
    if (A % 2) add_step(2);
    while (ell < ell0 + A) add_step(4);
    return ps;
  }
  const int rem = A % 3;
  if (rem == 1) add_step(2);
  if (rem == 2) add_step(4);
  while (ell < ell0 + A) add_step(8);
  return ps;
}

// Register twiddles of centre thread t (either field), as k_centre keeps
// them with cs = centre_slots(nt): slot k (r - 1) + e - 1 holds tau^e of the
// butterfly the thread takes in iteration k of the last step (lds_step's
// thread -> butterfly map with one column and two blocks), slot
// cs.step_slots + k the squaring twiddle of the pair it takes in iteration k
// (square_self's map for workgroup 0, square_cross's otherwise). Unused slots
// stay 0.
template <class E>
void centre_register_twiddles(const Pass& cp, const CentreSlots& cs, uint32_t g, const uint32_t bl[2],
                              const std::vector<E>& tw, const std::vector<E>& sq, std::vector<E>& out) {
  const int c = cp.A;
  const Step& st = cp.steps[cp.nsteps - 1];
  const int r = st.radix, lr = log2i(r), lam = st.ell - cp.ell0;
  const int logd = c - lam - lr, logU = c - lr;
  const int total = 2 << logU;
  const int Cb = 1 << c, half = Cb >> 1;
  const int rt = cs.role_threads;
  const size_t base = out.size();
  out.resize(base + (size_t)(cs.step_slots + cs.sq_slots) * rt, E{});
  auto at = [&](int slot, int t) -> E& { return out[base + (size_t)slot * rt + t]; };
  for (int t = 0; t < rt; ++t) {
    for (int k = 0; t + k * rt < total; ++k) {
      const int bt = t + k * rt;
      const int b = bt >> logU, u = bt & ((1 << logU) - 1), blk = u >> logd;
      const uint32_t g0 = st.twoff + ((bl[b] << lam) + (uint32_t)blk) * (uint32_t)(r - 1);
      for (int e = 1; e < r; ++e) at(k * (r - 1) + e - 1, t) = tw[g0 + e - 1];
    }
    for (int k = 0; k < cs.sq_slots; ++k) {
      const int o = t + k * rt;
      if (o >= Cb) break;  // both maps cover Cb pairs
      const int ti = g == 0 ? o : ((o & 1) == 0 ? (o >> 1) : half + ((Cb - 1 - o) >> 1));
      const int b = ti >= half ? 1 : 0;
      at(cs.step_slots + k, t) = sq[((size_t)bl[b] << (c - 1)) + (size_t)(ti - b * half)];
    }
  }
}

}  // namespace

bool build_class_plan(int ln, const PlanOpts& opts, DevPlan& P, ClassTables& T, std::string& err) {
  if (ln < 3 || ln > 27) {
    err = "transform length 2^" + std::to_string(ln) + " outside 2^3..2^27";
    return false;
  }
  if (opts.max_radix != 0 && opts.max_radix != 4 && opts.max_radix != 8) {
    err = "max radix must be 4 or 8";
    return false;
  }
  if (opts.threads != 0 && opts.threads != (uint32_t)MP_THREADS_NARROW && opts.threads != (uint32_t)MP_THREADS_WIDE) {
    err = "workgroup size must be 512 or 1024";
    return false;
  }
  const bool wide = opts.threads == (uint32_t)MP_THREADS_WIDE;
  const uint32_t ls_cap = opts.ls_cap;
  const int maxr = opts.max_radix ? (int)opts.max_radix : 8;
  const uint64_t n = uint64_t(1) << ln, h = n / 2;
  P = DevPlan{};
  P.centre_nt = wide ? MP_THREADS_WIDE : MP_THREADS_NARROW;
  P.ln = ln;
  P.L = ln - 1;
  P.h = (uint32_t)h;
  T.tw61.clear();
  T.tw31.clear();
  const int L = P.L;
  // -LS caps the points of the LDS-staged centre block; it also selects the
  // multi-pass plan when the whole residue would otherwise stay in one
  // workgroup.
  int c = CENTRE_LOG_C;
  if (ls_cap) {
    c = 0;
    while ((2u << c) <= ls_cap && c < CENTRE_LOG_C) ++c;
    c = std::max(c, 2);
  }
  if (L <= SMALL_LOG_H && !(ls_cap && L >= c + 2)) {
    P.small = true;
    // A single workgroup's transform is bound by latency: radix-4 steps
    // measured faster than radix-8 up to h = 1024 (13.2 vs 15.4 us/iter at
    // n = 512, 17.3 vs 20.1 at n = 2048), radix-8 faster at h = 2048
    // (25.4 vs 30.5 at n = 4096) and h = 4096 (42.3 vs 51.5 at n = 8192).
    // Single tests of n = 8192 take the multi-pass profile (26.4 us); the
    // single-workgroup plan pays at that length in batches, where it holds
    // one CU per test: 0.19 against 0.26 us per test-iteration for 256 tests.
    P.outer = make_pass(0, L, T.tw61, T.tw31, opts.max_radix ? maxr : (L <= 10 ? 4 : 8));
    P.nseg = (uint32_t)(h / SEG_ELEMS);
  } else {
    if (L < c + 2) {
      err = "plan error: transform too short for the multi-pass plan";
      return false;
    }
    P.small = false;
    const int a = L - c;
    int a_outer = a, a_mid = 0;
    if (a > 9) {
      a_mid = a / 2;
      a_outer = a - a_mid;
    }
    P.outer = make_pass(0, a_outer, T.tw61, T.tw31, maxr);
    P.logS = L - a_outer;
    P.logW = std::min(11 - a_outer, P.logS);
    if (opts.carry_pairs) {
      int lw = 0;
      while ((2u << lw) <= opts.carry_pairs) ++lw;
      P.logW = std::max(2, std::min(lw, P.logW));
    }
    if (a_outer > PASS_MAX_LOG_R || a_mid > PASS_MAX_LOG_R || (1 << P.logW) < SEG_ELEMS) {
      err = "plan error: strided passes too tall (local stage cap too small for this length)";
      return false;
    }
    if (a_mid) {
      P.has_mid = true;
      P.mid = make_pass(a_outer, a_mid, T.tw61, T.tw31, maxr);
      P.mid_logS = L - a_outer - a_mid;
      P.mid_logW = std::min(11 - a_mid, P.mid_logS);
    }
    // a 1024-thread centre holds 4 elements per thread: radix-4 steps
    P.centre = make_pass(L - c, c, T.tw61, T.tw31, wide ? 4 : maxr);
    P.nseg = (uint32_t)(h >> P.logW);
  }
  // centre twiddles: w_h^bitrev(j) for position 2j (the reference's w[h/2+j])
  T.sq61.resize(h / 2);
  T.sq31.resize(h / 2);
  {
    const int bits = L - 1;
    const e61 w61 = root61(L);
    const e31 w31 = root31(L);
    std::vector<e61> p61(h / 2);
    std::vector<e31> p31(h / 2);
    e61 c61{1, 0};
    e31 c31{1, 0};
    for (uint64_t k = 0; k < h / 2; ++k) {
      p61[k] = c61;
      p31[k] = c31;
      c61 = mul(c61, w61);
      c31 = mul(c31, w31);
    }
    // stored as 1 + w (the centre's squaring uses u1^2 (1 + w)), M61 in split form
    const e61 one61{1, 0};
    const e31 one31{1, 0};
    for (uint64_t j = 0; j < h / 2; ++j) {
      const uint32_t br = bits > 0 ? bitrev((uint32_t)j, bits) : 0u;
      T.sq61[j] = split31(add(one61, p61[br]));
      T.sq31[j] = add(one31, p31[br]);
    }
  }
  // step twiddles: M61 in the split form of mul_tw
  for (e61& w : T.tw61) w = split31(w);
  // per-workgroup centre records (contiguous, so a workgroup stages its
  // twiddles with one batched copy and loads its register twiddles with
  // coalesced loads)
  if (!P.small) {
    const Pass& cp = P.centre;
    const int c = cp.A;
    const uint32_t ncentre = (uint32_t)((h >> c) / 2);
    uint32_t nE = 0;  // twiddles of the steps before the last (LDS)
    for (int si = 0; si + 1 < cp.nsteps; ++si)
      nE += 2u * (1u << (cp.steps[si].ell - cp.ell0)) * (uint32_t)(cp.steps[si].radix - 1);
    const Step& last = cp.steps[cp.nsteps - 1];
    const int lr = log2i(last.radix);
    const CentreSlots cs = centre_slots(P.centre_nt);
    const int per_thread = ((2 << (c - lr)) + cs.role_threads - 1) / cs.role_threads;
    if (last.radix < 4 || nE > (uint32_t)CENTRE_EARLY_TW || per_thread * (last.radix - 1) > cs.step_slots ||
        (2 << (c - 1)) > cs.sq_slots * cs.role_threads) {
      err = "plan error: centre twiddles do not fit the workgroup's slots";
      return false;
    }
    P.crec = nE + (uint32_t)((cs.step_slots + cs.sq_slots) * cs.role_threads);
    T.ctw61.clear();
    T.ctw31.clear();
    T.ctw61.reserve((size_t)ncentre * P.crec);
    T.ctw31.reserve((size_t)ncentre * P.crec);
    for (uint32_t g = 0; g < ncentre; ++g) {
      uint32_t bl[2];
      centre_blocks(g, bl[0], bl[1]);
      for (int si = 0; si + 1 < cp.nsteps; ++si) {
        const Step& st = cp.steps[si];
        const int lam = st.ell - cp.ell0;
        const uint32_t span = (1u << lam) * (uint32_t)(st.radix - 1);
        for (int b = 0; b < 2; ++b) {
          const uint32_t g0 = st.twoff + (bl[b] << lam) * (uint32_t)(st.radix - 1);
          T.ctw61.insert(T.ctw61.end(), T.tw61.begin() + g0, T.tw61.begin() + g0 + span);
          T.ctw31.insert(T.ctw31.end(), T.tw31.begin() + g0, T.tw31.begin() + g0 + span);
        }
      }
      centre_register_twiddles(cp, cs, g, bl, T.tw61, T.sq61, T.ctw61);
      centre_register_twiddles(cp, cs, g, bl, T.tw31, T.sq31, T.ctw31);
    }
  }
  return true;
}

bool make_digit_gen(uint32_t p, const DevPlan& P, DigitGen& D, std::string& err) {
  const int ln = P.ln;
  if (transform_log2(p) != ln) {
    err = "exponent does not match the plan's transform length";
    return false;
  }
  if ((p & 1u) == 0) {
    err = "exponent must be odd";
    return false;
  }
  const uint64_t n = uint64_t(1) << ln;
  const uint32_t q = p >> ln;
  // The exactness argument for the carry needs 8 digits of at least 12 bits.
  if (!P.small && q < 12) {
    err = "plan error: digit width below 12 bits on the multi-pass plan";
    return false;
  }
  // The single-workgroup kernels compiled for n >= 16 carry in 64-bit words
  // only (16 <= q <= 43); every exponent of those lengths has 16 <= q <= 42.
  if (ln >= 4 && (q < 16 || q > 43)) {
    err = "plan error: digit width outside 16..44 bits at n >= 16";
    return false;
  }
  // Digit widths ceil(p(j+1)/n) - ceil(pj/n) and weight shifts
  // lr*(n - (pj mod n)) with 2^lr an n-th root of 2 in each field
  // (mersenne2_gpu.cpp:3921-3944).
  const uint64_t lr61 = ((uint64_t(1) << 60) >> ln) % 61u;
  const uint64_t lr31 = ((uint64_t(1) << 30) >> ln) % 31u;
  D = DigitGen{};
  D.p = p;
  D.nmask = (uint32_t)(n - 1);
  D.s = (uint32_t)(p & (n - 1));
  D.q = q;
  D.lr61 = (uint32_t)lr61;
  D.lr31 = (uint32_t)lr31;
  D.d61 = (uint32_t)((lr61 * D.s) % 61u);
  D.d31 = (uint32_t)((lr31 * D.s) % 31u);
  D.k61 = (uint32_t)((ln + 1) % 61);
  D.k31 = (uint32_t)((ln + 1) % 31);
  // The kernels restart the walk at a segment's first digit (dwalk_at) and
  // step it across the segment (dwalk_next). Check both against the closed
  // forms: over every digit up to n = 2^16; above that, over the first and
  // last 2^16 digits and a 64-digit window at every multiple of 4096
  // (about n/64 + 2^17 digits: the walk's correctness does not depend on the
  // digit index beyond what these cover, and the check is most of a test's
  // set-up time at large n).
  // The carrying kernels also take a segment's metadata (seg_meta: shifts,
  // unweighting rotations and width bits of its 8 digits, each count possibly
  // holding its modulus for 0) from its first digit; it is checked at every
  // segment start inside the same windows (all multiples of 8).
  auto check = [&](uint64_t j0, uint64_t len) -> bool {
    DWalk w = dwalk_at(D, (uint32_t)j0);
    SegMeta M{};
    uint64_t pj = (uint64_t)p * j0;            // p j
    uint64_t b0 = j0 ? ((pj - 1) >> ln) + 1 : 0;  // ceil(p j / n)
    for (uint64_t j = j0; j < j0 + len; ++j) {
      const uint64_t b1 = ((pj + p - 1) >> ln) + 1;
      const uint32_t wd = (uint32_t)(b1 - b0);
      const uint64_t r = pj & (n - 1);
      const uint32_t w61 = j ? (uint32_t)((lr61 * (n - r)) % 61u) : 0u;
      const uint32_t w31 = j ? (uint32_t)((lr31 * (n - r)) % 31u) : 0u;
      if ((wd != q && wd != q + 1) || dwalk_width(D, w) != wd || dwalk_w61(w) != w61 || dwalk_w31(w) != w31) {
        err = "digit walk mismatch at digit " + std::to_string(j);
        return false;
      }
      if ((j & (2 * SEG_ELEMS - 1)) == 0) M = seg_meta(D, (uint32_t)j);
      const uint32_t m = M.m[j & (2 * SEG_ELEMS - 1)];
      const uint32_t u61 = meta_u61(m), u31 = meta_u31(m);  // left rotations by -(w + ln + 1)
      if (meta_w61(m) > 61u || meta_w31(m) > 31u || u61 > 61u || u31 > 31u || meta_w61(m) % 61u != w61 ||
          meta_w31(m) % 31u != w31 || (u61 + w61 + D.k61) % 61u != 0 || (u31 + w31 + D.k31) % 31u != 0 ||
          meta_wide(m) != wd - q) {
        err = "segment metadata mismatch at digit " + std::to_string(j);
        return false;
      }
      dwalk_next(D, w);
      pj += p;
      b0 = b1;
    }
    return true;
  };
  constexpr uint64_t FULL = uint64_t(1) << 16;
  if (n <= FULL) return check(0, n);
  if (!check(0, FULL) || !check(n - FULL, FULL)) return false;
  for (uint64_t j0 = FULL; j0 < n - FULL; j0 += 4096)
    if (!check(j0, 64)) return false;
  return true;
}

std::string describe_plan(const DevPlan& P) {
  std::ostringstream o;
  auto pass = [&](const char* name, const Pass& ps) {
    o << name << ": levels " << ps.ell0 << ".." << (ps.ell0 + ps.A - 1) << " radix";
    for (int i = 0; i < ps.nsteps; ++i) o << (i ? "," : " ") << ps.steps[i].radix;
    o << "\n";
  };
  o << "plan: h=" << P.h << " (n=" << 2ull * P.h << "), ";
  if (P.small) o << "single-workgroup LDS-resident\n";
  else o << "multi-pass\n";
  if (P.small) {
    pass("  k_small", P.outer);
  } else {
    pass("  k_outer_fwd/k_outer_inv", P.outer);
    o << "    columns S=" << (1u << P.logS) << ", W=" << (1u << P.logW) << " per workgroup, "
      << ((1u << P.logS) >> P.logW) << " workgroups, carry segments=" << P.nseg << "\n";
    if (P.has_mid) {
      pass("  k_mid<fwd/inv>", P.mid);
      o << "    columns S=" << (1u << P.mid_logS) << ", W=" << (1u << P.mid_logW) << "\n";
    }
    pass("  k_centre", P.centre);
    o << "    " << ((P.h >> P.centre.A) / 2) << " workgroups of 2 blocks x " << (1u << P.centre.A) << ", "
      << P.centre_nt << " threads\n";
  }
  return o.str();
}

namespace {

// Non-negative big integer as 64-bit words.
struct Big {
  std::vector<uint64_t> w;
  explicit Big(size_t nw) : w(nw, 0) {}
  void add_at(uint64_t bit, unsigned __int128 v) {
    size_t i = bit / 64;
    const unsigned sh = bit % 64;
    uint64_t parts[3] = {(uint64_t)(v << sh), 0, 0};
    if (sh) {
      parts[1] = (uint64_t)(v >> (64 - sh));
      parts[2] = (uint64_t)((v >> 64) >> (64 - sh));
    } else {
      parts[1] = (uint64_t)(v >> 64);
    }
    unsigned __int128 c = 0;
    for (int k = 0; k < 3 || c; ++k, ++i) {
      if (i >= w.size()) w.push_back(0);
      const unsigned __int128 s = (unsigned __int128)w[i] + (k < 3 ? parts[k] : 0) + c;
      w[i] = (uint64_t)s;
      c = s >> 64;
      if (k >= 2 && !c) break;
    }
  }
  // value mod 2^p - 1, canonical (0 for M_p)
  void mod_mersenne(uint32_t p) {
    const size_t pw = p / 64;
    const unsigned pb = p % 64;
    for (;;) {
      bool high = false;
      for (size_t i = pw; i < w.size(); ++i) {
        const uint64_t x = (i == pw) ? (pb ? w[i] >> pb : w[i]) : w[i];
        if (x) {
          high = true;
          break;
        }
      }
      if (!high) break;
      // hi = value >> p ; value &= 2^p - 1 ; value += hi
      std::vector<uint64_t> hi(w.size() - pw + 1, 0);
      for (size_t i = 0; pw + i < w.size(); ++i) {
        const uint64_t lo = pb ? (w[pw + i] >> pb) : w[pw + i];
        const uint64_t up = (pb && pw + i + 1 < w.size()) ? (w[pw + i + 1] << (64 - pb)) : 0;
        hi[i] = lo | up;
      }
      if (pb)
        w[pw] &= (uint64_t(1) << pb) - 1;
      else if (pw < w.size())
        w[pw] = 0;
      for (size_t i = pw + 1; i < w.size(); ++i) w[i] = 0;
      unsigned __int128 c = 0;
      for (size_t i = 0; i < w.size(); ++i) {
        const unsigned __int128 s = (unsigned __int128)w[i] + (i < hi.size() ? hi[i] : 0) + c;
        w[i] = (uint64_t)s;
        c = s >> 64;
      }
      if (c) w.push_back((uint64_t)c);
    }
    // all ones below 2^p -> 0
    bool ones = true;
    for (size_t i = 0; i < pw && ones; ++i) ones = (w[i] == ~0ull);
    if (ones && pb) ones = ((w[pw] & ((uint64_t(1) << pb) - 1)) == ((uint64_t(1) << pb) - 1));
    if (ones) std::fill(w.begin(), w.end(), 0);
  }
  bool ge(const Big& o) const {
    const size_t m = std::max(w.size(), o.w.size());
    for (size_t k = m; k-- > 0;) {
      const uint64_t a = k < w.size() ? w[k] : 0, b = k < o.w.size() ? o.w[k] : 0;
      if (a != b) return a > b;
    }
    return true;
  }
  void sub(const Big& o) {  // requires *this >= o
    unsigned __int128 borrow = 0;
    for (size_t i = 0; i < w.size(); ++i) {
      const unsigned __int128 b = (unsigned __int128)(i < o.w.size() ? o.w[i] : 0) + borrow;
      borrow = (unsigned __int128)w[i] < b ? 1 : 0;
      w[i] = (uint64_t)((unsigned __int128)w[i] - b);
    }
  }
};

}  // namespace

std::vector<uint8_t> residue_from_state(uint32_t p, int ln, const int64_t* digits, uint64_t ndigits,
                                        const Carry128* carries, const uint64_t* carry_digit, uint64_t ncarries) {
  const size_t nw = p / 64 + 4;
  Big pos(nw), neg(0);  // neg is sized on its first term
  auto neg_at = [&](uint64_t bit, unsigned __int128 v) {
    if (neg.w.empty()) neg.w.assign(nw, 0);
    neg.add_at(bit, v);
  };
  // Digits that are all non-negative and below 2^47 (every state the kernels
  // leave: digit widths are at most 45 bits) are summed in one streaming
  // pass: a 128-bit accumulator holds the bits from word `wi` on, and each
  // digit is added at its bit offset (< 64) in it. Between two completed
  // words at most 64 digits arrive, each below 2^110 at its offset, so the
  // accumulator stays below 2^117.
  bool stream = true;
  for (uint64_t j = 0; j < ndigits && stream; ++j) stream = digits[j] >= 0 && (digits[j] >> 47) == 0;
  if (stream) {
    // Invariant: digit j's offset in the accumulator, ceil(p j / n) - 64 wi,
    // is below 64; the next digit starts at most 46 bits further on, so after
    // each digit at most one word is complete. Branch-free: the current word
    // is always written and the accumulator moves on when it is complete.
    unsigned __int128 acc = 0;
    uint64_t wi = 0, pj = 0;
    const uint64_t n1 = (uint64_t(1) << ln) - 1;
    for (uint64_t j = 0; j < ndigits; ++j) {
      const uint64_t d = (uint64_t)digits[j];
      const uint64_t sh = ((pj + n1) >> ln) - 64 * wi;
      acc += ((unsigned __int128)((d >> 1) >> (63 - sh)) << 64) | (d << sh);
      pj += p;
      pos.w[wi] = (uint64_t)acc;
      const uint64_t next = (((pj + n1) >> ln) - 64 * wi) >> 6;  // 0 or 1
      acc = next ? (acc >> 64) : acc;
      wi += next;
    }
    for (; acc; acc >>= 64) pos.w[wi++] = (uint64_t)acc;
  } else {
    for (uint64_t j = 0; j < ndigits; ++j) {
      const int64_t d = digits[j];
      if (d > 0) pos.add_at(digit_bitpos(p, ln, j), (unsigned __int128)(uint64_t)d);
      if (d < 0) neg_at(digit_bitpos(p, ln, j), (unsigned __int128)(uint64_t)(-(d + 1)) + 1u);
    }
  }
  for (uint64_t s = 0; s < ncarries; ++s) {
    const __int128 v = (__int128)(((unsigned __int128)(uint64_t)carries[s].hi << 64) | carries[s].lo);
    if (v == 0) continue;
    const uint64_t b = digit_bitpos(p, ln, carry_digit[s]);
    if (v > 0) pos.add_at(b, (unsigned __int128)v);
    else neg_at(b, (unsigned __int128)(-(v + 1)) + 1u);
  }
  pos.mod_mersenne(p);
  bool any_neg = false;
  for (uint64_t x : neg.w) any_neg |= x != 0;
  if (any_neg) neg.mod_mersenne(p);
  if (any_neg && !pos.ge(neg)) {  // pos += M_p
    Big m(nw);
    for (uint32_t b = 0; b < p; ++b) m.w[b / 64] |= uint64_t(1) << (b % 64);
    Big t = pos;
    // t = pos + M_p
    unsigned __int128 c = 0;
    t.w.resize(std::max(t.w.size(), m.w.size()) + 1, 0);
    for (size_t i = 0; i < t.w.size(); ++i) {
      const unsigned __int128 s = (unsigned __int128)t.w[i] + (i < m.w.size() ? m.w[i] : 0) + c;
      t.w[i] = (uint64_t)s;
      c = s >> 64;
    }
    pos = t;
  }
  if (any_neg) {
    pos.sub(neg);
    pos.mod_mersenne(p);
  }
  const size_t nbytes = (p + 7) / 8;
  std::vector<uint8_t> out(nbytes, 0);
  for (size_t i = 0; i < nbytes; i += 8) {
    const uint64_t x = pos.w[i / 8];
    for (size_t k = 0; k < 8 && i + k < nbytes; ++k) out[i + k] = (uint8_t)(x >> (8 * k));
  }
  return out;
}

void digits_from_residue(uint32_t p, int ln, const uint8_t* r, size_t len, int64_t* digits) {
  // digit j: the wd <= 47 bits from bit b0, read as one little-endian 64-bit
  // window starting at byte b0 / 8 (bytes past len read as 0)
  const uint64_t n = uint64_t(1) << ln;
  for (uint64_t j = 0; j < n; ++j) {
    const uint64_t b0 = digit_bitpos(p, ln, j), wd = digit_bitpos(p, ln, j + 1) - b0;
    const uint64_t at = b0 / 8;
    uint64_t x = 0;
    if (at + 8 <= len) {
      memcpy(&x, r + at, 8);
    } else {
      for (uint64_t k = 0; at + k < len && k < 8; ++k) x |= (uint64_t)r[at + k] << (8 * k);
    }
    digits[j] = (int64_t)((x >> (b0 % 8)) & ((uint64_t(1) << wd) - 1));
  }
}

}  // namespace m2
