// gfx950 HIP kernels for one IBDWT squaring of x mod 2^p - 1 over
// GF(M61^2) x GF(M31^2).
//
// Data between iterations: the digits as signed 64-bit pairs (digits 2e and
// 2e+1 of element e) and G, the carries not yet applied. On the small plan the
// digits live in z61 in natural order. On the multi-pass plan they live in
// their own buffer zd, strip-major (all rows of one outer-pass column strip
// contiguous) like G, so the outer passes read and write them with full cache
// lines; a separate buffer, because the spectrum in z61/z31 is in natural
// element order and a workgroup's strip overlaps other workgroups' rows.
//
// Carry rule (exact, no waits between workgroups): the digits are cut into
// segments; every segment is normalised with carry-in 0, then receives the
// carry-out g of the previous segment (cyclically: 2^p = 1 mod M_p) and
// ripples it upward, but its top digit keeps whatever is left instead of
// passing it on. The represented value is unchanged, every digit but a
// segment's top one is < 2^width, and the top one is <= 2^width as long as
// the bits below it exceed the bits of g (8 digits of >= 12 bits each). No
// digit is negative (LL's -2 included, see seg_normalise), so convolution
// terms stay in [0, 2^(ln+2q+3) - 2^(2q+2)], below M61*M31 (crt61_31).
//
// Twiddles: a workgroup copies the twiddle blocks of its interior radix steps
// into LDS with coalesced loads at the start. The centre keeps the twiddles of
// its last forward step, its squaring and its first inverse step in
// registers (one coalesced record slot per thread); a step that reads or
// writes global memory directly at a pass edge (edge_step) loads its own
// twiddles from the global tables.
#include <hip/hip_ext.h>

#include "m2_arith.h"
#include "m2_digits.h"
#include "m2_internal.h"

namespace m2 {

typedef __int128 i128;

__device__ __forceinline__ int64_t to_i64(uint64_t x) { return (int64_t)x; }

// LDS data index with one pad slot per 8 elements. Radix steps access
// elements 8^k apart; the pad spreads those (and 8-element butterfly groups)
// over distinct banks.
__device__ __forceinline__ int PADX(int i) { return i + (i >> 3); }
constexpr int padded(int n) { return n + n / 8; }

// Threads per workgroup of the strided (outer and middle) passes: 256 per
// field. The centre pass takes its count NT as a template parameter:
// MP_THREADS_NARROW or MP_THREADS_WIDE, half per field.
constexpr int MP_THREADS = MP_THREADS_NARROW;
template <int NT>
constexpr int centre_tw_slots() {
  return centre_slots(NT).step_slots + centre_slots(NT).sq_slots;
}

// Signed digit -> canonical field residues.
__device__ __forceinline__ uint64_t digit61(int64_t d) { return d >= 0 ? (uint64_t)d : P61 - (uint64_t)(-d); }
// x mod P31 in [0, P31] for |x| < 2^62 (digits are below 2^46): one fold
// (x mod 2^31 + x >> 31 < 2 P31) and an unsigned min.
__device__ __forceinline__ uint32_t mod31_u64(uint64_t x) {
  const uint32_t r = ((uint32_t)x & P31) + (uint32_t)(x >> 31);
  return umin32(r, r - P31);
}
__device__ __forceinline__ uint32_t digit31(int64_t d) {
  return d >= 0 ? mod31_u64((uint64_t)d) : neg31(mod31_u64((uint64_t)(-d)));
}

// Digit masks 2^q - 1 and 2^(q+1) - 1 of a test (uniform per test).
struct WidthMasks {
  uint64_t q, q1;
  __device__ __forceinline__ explicit WidthMasks(const DigitGen& D)
      : q((1ull << D.q) - 1ull), q1((2ull << D.q) - 1ull) {}
  __device__ __forceinline__ uint64_t of(uint32_t m) const { return meta_wide(m) ? q1 : q; }
};

// Unweight (the rotations in the digit's metadata) and Garner CRT (crt61_31:
// x in [0, M) read unsigned, a narrow band below M as x - M), plus `ones`
// (the digit's all-ones value 2^width - 1 in LL tests, else 0; see
// seg_normalise).
__device__ __forceinline__ i128 crt_digit(uint64_t s, uint32_t t, uint32_t meta, uint64_t ones) {
  return crt61_31(rot61(s, meta_u61(meta)), rot31(t, meta_u31(meta)), ones);
}

// acc >> w (arithmetic) for 0 <= w < 64, as two 64-bit shifts of the words.
__device__ __forceinline__ i128 shr_w(i128 acc, uint32_t w) {
  const uint64_t lo = (uint64_t)acc;
  const int64_t hi = (int64_t)(acc >> 64);
  const uint64_t clo = (lo >> w) | (((uint64_t)hi << 1) << (63u - w));
  return (i128)((unsigned __int128)(uint64_t)(hi >> w) << 64 | clo);
}

// Per-test buffers reach the kernels through the test descriptors, i.e. as
// pointers loaded from memory, which the compiler treats as generic (flat)
// addresses. They are device (global) memory: casting them to the global
// address space gives global loads and stores, which do not wait on the LDS
// counter as flat ones do.
#define M2_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ M2_GLOBAL T* gptr(T* p) {
  return (M2_GLOBAL T*)p;
}

// Element copies between global memory and registers/LDS (structure
// assignment does not cross address spaces). Stores of spectra and digits are
// streaming (nontemporal): their reader is the next launch, on any XCD, and
// so they need not stay in this XCD's L2 (every shape measured 1.5-2% faster
// per iteration at n = 2^20..2^24, batched or not).
__device__ __forceinline__ e61 gld(const M2_GLOBAL e61* p) { return e61{p->a, p->b}; }
__device__ __forceinline__ e31 gld(const M2_GLOBAL e31* p) { return e31{p->a, p->b}; }
__device__ __forceinline__ Carry128 gld(const M2_GLOBAL Carry128* p) { return Carry128{p->lo, p->hi}; }
__device__ __forceinline__ void gst(M2_GLOBAL e61* p, e61 v) {
  __builtin_nontemporal_store(v.a, &p->a);
  __builtin_nontemporal_store(v.b, &p->b);
}
__device__ __forceinline__ void gst(M2_GLOBAL e31* p, e31 v) {
  __builtin_nontemporal_store(v.a, &p->a);
  __builtin_nontemporal_store(v.b, &p->b);
}
__device__ __forceinline__ void gst(M2_GLOBAL Carry128* p, Carry128 v) { p->lo = v.lo; p->hi = v.hi; }
// Streaming loads of a kernel's spectrum (NTM). Measured per kernel shape:
// they pay where one workgroup per CU runs the whole grid (the n = 2^20
// single test's fused outer pass and 1024-thread centre: 49.2 -> 45.4
// us/iteration) and cost 1-3% where two workgroups per CU overlap (batches,
// n >= 2^21), so only those kernels use them.
template <bool NTM>
__device__ __forceinline__ e61 gld_s(const M2_GLOBAL e61* p) {
  if constexpr (NTM) return e61{__builtin_nontemporal_load(&p->a), __builtin_nontemporal_load(&p->b)};
  else return gld(p);
}
template <bool NTM>
__device__ __forceinline__ e31 gld_s(const M2_GLOBAL e31* p) {
  if constexpr (NTM) return e31{__builtin_nontemporal_load(&p->a), __builtin_nontemporal_load(&p->b)};
  else return gld(p);
}

template <class CP>
__device__ __forceinline__ i128 load_carry(CP G, uint32_t i) {
  const Carry128 c = G[i];
  return (((i128)c.hi) << 64) | (i128)c.lo;
}
template <class CP>
__device__ __forceinline__ void store_carry(CP G, uint32_t i, i128 v) {
  G[i].lo = (uint64_t)v;
  G[i].hi = (int64_t)(v >> 64);
}

// Carry segments: 4 elements (8 digits), first digit j0. In every kernel
// that carries, each thread owns at most one segment (a workgroup holds at
// most 2048 elements, 512 segments, with 512 or 1024 threads) and keeps its 8 digits
// in registers from the normalisation to the weighting, so the digits do not
// go through LDS between the carry steps.
static_assert(PASS_MAX_ELEMS / SEG_ELEMS <= MP_THREADS_NARROW, "one carry segment per thread");
static_assert((1 << PASS_MAX_LOG_R) - 1 <= MP_THREADS_NARROW && CENTRE_EARLY_TW <= MP_THREADS_NARROW,
              "one staged twiddle per thread");
struct SegDigits {
  int64_t d[2 * SEG_ELEMS];
};

// Unweight + CRT + normalisation of the segment's spectrum at LDS index idx0
// into registers, with carry-in `carry`; returns the carry-out.
//
// LL tests subtract 2 as carry-in -2 at digit 0, and (ll_ones = all ones)
// add the all-ones digit vector, whose value is 2^p - 1 = 0 (mod 2^p - 1), to
// the terms: every digit's sum is then the term plus 2^width - 1 and digit
// 0's is the term plus 2^width - 3 >= 0 (width >= 2), so no sum, digit or
// carry is ever negative and the next squaring's terms stay in the range the
// CRT reads (crt61_31). PRP tests pass ll_ones = 0 and carry-in 0.
//
// Tests with 16 <= q <= 31 or 32 <= q <= 43 (the same for every lane) take
// the 64-bit forms (norm_digit32, norm_digit64); the carry-in is then 0 or -2.
// GENERAL = false leaves out the 128-bit form of other widths: transform
// lengths n >= 16 only have exponents with 16 <= q <= 42 (build_plan checks).
template <bool GENERAL = true>
__device__ __forceinline__ i128 seg_normalise(const e61* L61, const e31* L31, int idx0, const SegMeta& M,
                                              const WidthMasks& wm, uint32_t q, i128 carry, uint64_t ll_ones,
                                              SegDigits& s) {
  if (norm32_applies(q)) {
    uint64_t cA = (uint64_t)(int64_t)carry, cB = 0;
#pragma unroll
    for (int i = 0; i < SEG_ELEMS; ++i) {
      const e61 v = L61[PADX(idx0 + i)];
      const e31 y = L31[PADX(idx0 + i)];
      const uint32_t ma = M.m[2 * i], mb = M.m[2 * i + 1];
      s.d[2 * i] = (int64_t)norm_digit32(v.a, y.a, ma, q, wm.of(ma) & ll_ones, cA, cB);
      s.d[2 * i + 1] = (int64_t)norm_digit32(v.b, y.b, mb, q, wm.of(mb) & ll_ones, cA, cB);
    }
    return ((i128)cB << 32) + (i128)cA;
  }
  if (!GENERAL || norm64_applies(q)) {
    uint64_t c = (uint64_t)(int64_t)carry;
#pragma unroll
    for (int i = 0; i < SEG_ELEMS; ++i) {
      const e61 v = L61[PADX(idx0 + i)];
      const e31 y = L31[PADX(idx0 + i)];
      const uint32_t ma = M.m[2 * i], mb = M.m[2 * i + 1];
      s.d[2 * i] = (int64_t)norm_digit64(v.a, y.a, ma, q, wm.of(ma) & ll_ones, c);
      s.d[2 * i + 1] = (int64_t)norm_digit64(v.b, y.b, mb, q, wm.of(mb) & ll_ones, c);
    }
    return (i128)c;
  }
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) {
    const e61 v = L61[PADX(idx0 + i)];
    const e31 y = L31[PADX(idx0 + i)];
    const uint32_t ma = M.m[2 * i], mb = M.m[2 * i + 1];
    const uint64_t ka = wm.of(ma), kb = wm.of(mb);
    const i128 xa = crt_digit(v.a, y.a, ma, ka & ll_ones) + carry;
    s.d[2 * i] = (int64_t)((uint64_t)xa & ka);
    carry = shr_w(xa, q + meta_wide(ma));
    const i128 xb = crt_digit(v.b, y.b, mb, kb & ll_ones) + carry;
    s.d[2 * i + 1] = (int64_t)((uint64_t)xb & kb);
    carry = shr_w(xb, q + meta_wide(mb));
  }
  return carry;
}

// Add carry c at the segment's first digit; the top digit absorbs the rest.
// A wave skips the remaining digits once no lane has a carry left. Tests with
// 16 <= q <= 31 have digits below 2^32 and carries in [0, 2^78) here, which
// ripple in the 64-bit form of the normalisation (ripple_digit32); with
// 32 <= q <= 43 carries are below 2^61 (ripple_digit64).
template <bool GENERAL = true>
__device__ __forceinline__ void seg_ripple(SegDigits& s, const SegMeta& M, const WidthMasks& wm, uint32_t q, i128 c) {
  if (norm32_applies(q)) {
    uint64_t cA = (uint64_t)c & 0xFFFFFFFFull, cB = (uint64_t)(c >> 32);
#pragma unroll
    for (int i = 0; i < 2 * SEG_ELEMS - 1; ++i)
      if ((cA | cB) != 0) s.d[i] = (int64_t)ripple_digit32((uint64_t)s.d[i], M.m[i], q, cA, cB);
    s.d[2 * SEG_ELEMS - 1] += (int64_t)(cA + (cB << 32));
    return;
  }
  if (!GENERAL || norm64_applies(q)) {  // carries in [0, 2^61)
    uint64_t cc = (uint64_t)c;
#pragma unroll
    for (int i = 0; i < 2 * SEG_ELEMS - 1; ++i)
      if (cc != 0) s.d[i] = (int64_t)ripple_digit64((uint64_t)s.d[i], M.m[i], q, cc);
    s.d[2 * SEG_ELEMS - 1] += (int64_t)cc;
    return;
  }
#pragma unroll
  for (int i = 0; i < 2 * SEG_ELEMS - 1; ++i) {
    if (c != 0) {
      const i128 acc = (i128)s.d[i] + c;
      s.d[i] = (int64_t)((uint64_t)acc & wm.of(M.m[i]));
      c = shr_w(acc, q + meta_wide(M.m[i]));
    }
  }
  s.d[2 * SEG_ELEMS - 1] += (int64_t)c;
}

// The segment's digits -> weighted field elements at LDS index idx0. With
// nonneg (tests with 16 <= q <= 43) every digit is in [0, 2^61).
template <bool NONNEG = false>
__device__ __forceinline__ void seg_weight_as(const SegDigits& s, e61* L61, e31* L31, int idx0, const SegMeta& M) {
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) {
    const int64_t da = s.d[2 * i], db = s.d[2 * i + 1];
    const uint32_t ma = M.m[2 * i], mb = M.m[2 * i + 1];
    if (NONNEG) {
      L61[PADX(idx0 + i)] = {rot61((uint64_t)da, meta_w61(ma)), rot61((uint64_t)db, meta_w61(mb))};
      L31[PADX(idx0 + i)] = {rot31(mod31_u64((uint64_t)da), meta_w31(ma)), rot31(mod31_u64((uint64_t)db), meta_w31(mb))};
    } else {
      L61[PADX(idx0 + i)] = {rot61(digit61(da), meta_w61(ma)), rot61(digit61(db), meta_w61(mb))};
      L31[PADX(idx0 + i)] = {rot31(digit31(da), meta_w31(ma)), rot31(digit31(db), meta_w31(mb))};
    }
  }
}
template <bool GENERAL = true>
__device__ __forceinline__ void seg_weight(const SegDigits& s, e61* L61, e31* L31, int idx0, const SegMeta& M,
                                           uint32_t q) {
  if (!GENERAL || norm32_applies(q) || norm64_applies(q)) seg_weight_as<true>(s, L61, L31, idx0, M);
  else seg_weight_as<false>(s, L61, L31, idx0, M);
}

// Segment digits to / from global memory (digit pairs as int64, 64 contiguous
// bytes per segment).
__device__ __forceinline__ void seg_store(M2_GLOBAL e61* dst, const SegDigits& s) {
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) gst(dst + i, e61{(uint64_t)s.d[2 * i], (uint64_t)s.d[2 * i + 1]});
}
__device__ __forceinline__ void seg_load(const M2_GLOBAL e61* src, SegDigits& s) {
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) {
    const e61 v = gld(src + i);
    s.d[2 * i] = to_i64(v.a);
    s.d[2 * i + 1] = to_i64(v.b);
  }
}
__device__ __forceinline__ i128 carry_at(const M2_GLOBAL Carry128* G, uint32_t i) {
  const Carry128 c = gld(G + i);
  return (((i128)c.hi) << 64) | (i128)c.lo;
}

// ---------------------------------------------------------------------------
// Twiddles in LDS. For step si of a pass, block b of the workgroup (with
// block index j0[b] at the pass's first level) uses the contiguous global
// range twoff + (j0[b] << lam) * (r-1) .. + (1 << lam) * (r-1); it is copied to
// LDS at offset lo + b * (1 << lam) * (r-1), lo = sum of the earlier steps.
__device__ __forceinline__ int twiddle_span(const Step& st, int lam) { return (1 << lam) * (st.radix - 1); }

template <int B, class G61, class G31, class Src, class Dst>
__device__ __forceinline__ void gather2(e61* L61, e31* L31, G61 g61, G31 g31, int count, Src src, Dst dst);

// A flat range of count <= blockDim.x twiddles (the tables of a pass that
// starts at level 0, or a centre record's LDS part), one entry per thread:
// load() issues the global loads into registers, store() writes LDS later,
// so that their latency overlaps the kernel's data loads instead of adding
// to them.
struct FlatTwiddles {
  uint64_t a = 0, b = 0;
  uint32_t c = 0, d = 0;
  __device__ __forceinline__ void load(const e61* __restrict__ g61, const e31* __restrict__ g31, int count) {
    if (count <= 0) return;
    const int i = min((int)threadIdx.x, count - 1);
    a = g61[i].a;
    b = g61[i].b;
    c = g31[i].a;
    d = g31[i].b;
  }
  __device__ __forceinline__ void store(e61* T61, e31* T31, int count) const {
    if ((int)threadIdx.x < count) {
      T61[threadIdx.x] = e61{a, b};
      T31[threadIdx.x] = e31{c, d};
    }
  }
};
// The flat range of an outer pass (levels 0..A-1): its steps' tables lie back
// to back, 2^A - 1 <= 511 entries, the layout stage_pass_twiddles gives.
__device__ __forceinline__ int outer_twiddles(FlatTwiddles& ft, const Pass& ps, const e61* __restrict__ tw61,
                                              const e31* __restrict__ tw31) {
  const int n = (1 << ps.A) - 1;
  ft.load(tw61 + ps.steps[0].twoff, tw31 + ps.steps[0].twoff, n);
  return n;
}

__device__ __forceinline__ int stage_pass_twiddles(e61* T61, e31* T31, const Pass& ps, int nb, const uint32_t* j0,
                                                   const e61* __restrict__ tw61, const e31* __restrict__ tw31) {
  if (ps.ell0 == 0 && nb == 1 && j0[0] == 0) {
    // a pass that starts at level 0 uses its steps' whole tables, which lie
    // back to back: one flat copy of 2^A - 1 entries
    const int n = (1 << ps.A) - 1;
    const e61* g61 = tw61 + ps.steps[0].twoff;
    const e31* g31 = tw31 + ps.steps[0].twoff;
    gather2<1>(T61, T31, g61, g31, n, [](int i) { return (uint32_t)i; }, [](int i) { return i; });
    return n;
  }
  int lo = 0;
  for (int si = 0; si < ps.nsteps; ++si) {
    const Step st = ps.steps[si];
    const int lam = st.ell - ps.ell0;
    const int span = twiddle_span(st, lam);
    for (int b = 0; b < nb; ++b) {
      const uint32_t g0 = st.twoff + (j0[b] << lam) * (uint32_t)(st.radix - 1);
      for (int k = threadIdx.x; k < span; k += blockDim.x) {
        T61[lo + b * span + k] = tw61[g0 + k];
        T31[lo + b * span + k] = tw31[g0 + k];
      }
    }
    lo += nb * span;
  }
  return lo;
}

// End of a radix step: a workgroup barrier, or, with WAVE (each field's whole
// transform is one wave's), only an ordering point for the wave's own LDS
// accesses: a wave's LDS instructions complete in issue order, so its later
// reads see its earlier writes as long as the compiler keeps them in order.
template <bool WAVE>
__device__ __forceinline__ void step_sync() {
  if (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One radix step over data in LDS. M61 values leave forward steps folded
// once (<= P61 + 7) and inverse steps reduced (<= P61); twiddles are in the
// split form of mul_tw. The LDS holds nb blocks of R = 2^logR rows
// by W = 2^logW columns (row-major); every column of a block is an
// independent transform. Step `lam` (relative level) with radix r: the
// butterflies combine rows blk*(R>>lam) + off + e*d, e < r, d = R >> (lam +
// log2 r); the twist tau^e of block blk is at Ts[b * span + blk*(r-1) + e-1].
// Thread roles. With SPLIT (the 512-thread multi-pass kernels) the first
// half of the workgroup works on the M61 field and the second half on the
// M31 field, so each SIMD holds one wave of each and runs one while the other
// waits on LDS; without SPLIT every thread does both fields.
template <bool SPLIT>
struct Lanes {
  int tid, n;
  bool f61, f31;
  __device__ __forceinline__ Lanes() {
    if (SPLIT) {
      n = (int)blockDim.x >> 1;
      f61 = (int)threadIdx.x < n;
      f31 = !f61;
      tid = f61 ? (int)threadIdx.x : (int)threadIdx.x - n;
    } else {
      n = (int)blockDim.x;
      f61 = f31 = true;
      tid = (int)threadIdx.x;
    }
  }
};

// TWIST = false is the step at level 0, whose only block has tau = 1: no
// twiddle products; the inverse then reduces its M61 outputs itself (the
// carry step's unweighting rotates 61-bit values).
template <bool FWD, int RADIX, bool SPLIT, bool TWIST = true, bool WAVE = false>
__device__ __forceinline__ void lds_step(e61* L61, e31* L31, int logR, int logW, int nb, int lam, const e61* Ts61,
                                         const e31* Ts31) {
  constexpr int LR = RADIX == 8 ? 3 : (RADIX == 4 ? 2 : 1);
  const Lanes<SPLIT> ln;
  const int logd = logR - lam - LR;
  const int logU = logR - LR + logW;  // butterflies per block (log2)
  const int total = nb << logU;
  const int span = (1 << lam) * (RADIX - 1);
  for (int t = ln.tid; t < total; t += ln.n) {
    const int b = t >> logU;
    const int u = t & ((1 << logU) - 1);
    const int col = u & ((1 << logW) - 1);
    const int g = u >> logW;
    const int off = g & ((1 << logd) - 1);
    const int blk = g >> logd;
    const int row0 = (blk << (logR - lam)) + off;
    const int base = (b << (logR + logW)) + (row0 << logW) + col;
    const int stride = 1 << (logd + logW);
    const int tw = b * span + blk * (RADIX - 1);
    if (ln.f61) {
      const e61* T61 = Ts61 + tw;
      e61 v[RADIX];
#pragma unroll
      for (int e = 0; e < RADIX; ++e) v[e] = L61[PADX(base + e * stride)];
      if (FWD) {
        if (TWIST) {
#pragma unroll
          for (int e = 1; e < RADIX; ++e) v[e] = mul_tw(v[e], T61[e - 1]);
        }
        if (RADIX == 8) dft8_fwd_lazy(v);
        if (RADIX == 4) dft4_fwd_lazy(v);
        if (RADIX == 2) dft2_fwd_lazy(v[0], v[1]);
      } else {
        if (RADIX == 8) dft8_inv_lazy(v);
        if (RADIX == 4) dft4_inv_lazy(v);
        if (RADIX == 2) dft2_inv_lazy(v[0], v[1]);
        if (TWIST) {
#pragma unroll
          for (int e = 1; e < RADIX; ++e) v[e] = mulconj_tw(v[e], T61[e - 1]);
        } else {
          // the lazy inverse outputs are folded once (<= P61 + 7; output 0 of
          // radix 4 and 8 already reduced), so one more fold reduces them
#pragma unroll
          for (int e = RADIX == 2 ? 0 : 1; e < RADIX; ++e) v[e] = {fold61(v[e].a), fold61(v[e].b)};
        }
      }
#pragma unroll
      for (int e = 0; e < RADIX; ++e) L61[PADX(base + e * stride)] = v[e];
    }
    if (ln.f31) {
      const e31* T31 = Ts31 + tw;
      e31 x[RADIX];
#pragma unroll
      for (int e = 0; e < RADIX; ++e) x[e] = L31[PADX(base + e * stride)];
      if (FWD) {
        if (TWIST) {
#pragma unroll
          for (int e = 1; e < RADIX; ++e) x[e] = mul_tw(x[e], T31[e - 1]);
        }
        if (RADIX == 8) dft8_fwd(x);
        if (RADIX == 4) dft4_fwd(x);
        if (RADIX == 2) dft2(x[0], x[1]);
      } else {
        if (RADIX == 8) dft8_inv(x);
        if (RADIX == 4) dft4_inv(x);
        if (RADIX == 2) dft2(x[0], x[1]);
        if (TWIST) {
#pragma unroll
          for (int e = 1; e < RADIX; ++e) x[e] = mulconj_tw(x[e], T31[e - 1]);
        }
      }
#pragma unroll
      for (int e = 0; e < RADIX; ++e) L31[PADX(base + e * stride)] = x[e];
    }
  }
  step_sync<WAVE>();
}

// All steps of a pass, forward (in order) or inverse (reversed), twiddles in
// LDS as laid out by stage_pass_twiddles. LEVEL0: the pass may start at level
// 0, whose step runs without twiddle products. WAVE: see step_sync.
template <bool FWD, bool SPLIT, bool LEVEL0 = false, bool WAVE = false>
__device__ __forceinline__ void lds_pass(e61* L61, e31* L31, int logW, int nb, const Pass& ps, const e61* T61,
                                         const e31* T31, int k0 = 0, int k1 = MAX_STEPS) {
  int offs[MAX_STEPS];
  int lo = 0;
  for (int si = 0; si < ps.nsteps; ++si) {
    offs[si] = lo;
    lo += nb * twiddle_span(ps.steps[si], ps.steps[si].ell - ps.ell0);
  }
  for (int k = k0; k < min(k1, ps.nsteps); ++k) {
    const int si = FWD ? k : ps.nsteps - 1 - k;
    const Step st = ps.steps[si];
    const int lam = st.ell - ps.ell0;
    if (LEVEL0 && st.ell == 0) {
      if (st.radix == 8) lds_step<FWD, 8, SPLIT, false, WAVE>(L61, L31, ps.A, logW, nb, lam, T61, T31);
      else if (st.radix == 4) lds_step<FWD, 4, SPLIT, false, WAVE>(L61, L31, ps.A, logW, nb, lam, T61, T31);
      else lds_step<FWD, 2, SPLIT, false, WAVE>(L61, L31, ps.A, logW, nb, lam, T61, T31);
      continue;
    }
    const e61* t61 = T61 + offs[si];
    const e31* t31 = T31 + offs[si];
    if (st.radix == 8) lds_step<FWD, 8, SPLIT, true, WAVE>(L61, L31, ps.A, logW, nb, lam, t61, t31);
    else if (st.radix == 4) lds_step<FWD, 4, SPLIT, true, WAVE>(L61, L31, ps.A, logW, nb, lam, t61, t31);
    else lds_step<FWD, 2, SPLIT, true, WAVE>(L61, L31, ps.A, logW, nb, lam, t61, t31);
  }
}

// body(k) for k = 0, 1, ... < KMAX until it returns false: unrolled for
// KMAX <= 2 (so that a register-slot getter indexed by k resolves at compile
// time), a loop otherwise (code size).
template <int KMAX, class F>
__device__ __forceinline__ void for_k(F&& body) {
  if constexpr (KMAX <= 2) {
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      if (!body(k)) break;
  } else {
#pragma unroll 1
    for (int k = 0; k < KMAX; ++k)
      if (!body(k)) break;
  }
}

// Squaring of blocks that pair with themselves: block 0 (positions 0..Cb-1,
// mirror formula, k = 0 special) at LDS offset 0, and, when nself = 2, block 1
// (positions Cb..2Cb-1, k pairs with 3Cb-1-k) at LDS offset Cb. The centre
// twiddle of pair q (block 0's Cb/2 even positions, then block 1's) is
// s61(k, q) / s31(k, q), k being the thread's iteration: from LDS, or from
// the registers the centre keeps for iteration k (k < 4: a thread takes at
// most KMAX pairs, nself * Cb/2 <= KMAX x threads per field). WAVE: see
// step_sync.
template <bool SPLIT, bool WAVE = false, int KMAX = 4, class S61, class S31>
__device__ __forceinline__ void square_self(e61* L61, e31* L31, int logCb, int nself, S61 s61, S31 s31) {
  const Lanes<SPLIT> ln;
  const int Cb = 1 << logCb;
  const int half = Cb >> 1;
  for_k<KMAX>([&](int k) {
    const int q = ln.tid + k * ln.n;
    if (q >= nself * half) return false;
    const int blk = q >= half ? 1 : 0;
    const int j = q - blk * half;
    const int kk = 2 * j;
    const int kidx = blk * Cb + kk;
    const int mkidx = blk ? Cb + (Cb - 1 - kk) : (kk ? (3 << (31 - __clz(kk))) - kk - 1 : 0);
    if (blk == 0 && kk == 0) {
      if (ln.f61) {
        e61 a = L61[PADX(0)], a2 = a;
        square_pair(a, a2, s61(k, q));
        const e61 z1 = red(L61[PADX(1)]);
        L61[PADX(0)] = a;
        L61[PADX(1)] = sqr(add(z1, z1));
      }
      if (ln.f31) {
        e31 b = L31[PADX(0)], b2 = b;
        square_pair(b, b2, s31(k, q));
        const e31 y1 = L31[PADX(1)];
        L31[PADX(0)] = b;
        L31[PADX(1)] = sqr(add(y1, y1));
      }
    } else {
      if (ln.f61) {
        e61 a = L61[PADX(kidx)], c = L61[PADX(mkidx)];
        square_pair(a, c, s61(k, q));
        L61[PADX(kidx)] = a;
        L61[PADX(mkidx)] = c;
      }
      if (ln.f31) {
        e31 b = L31[PADX(kidx)], d = L31[PADX(mkidx)];
        square_pair(b, d, s31(k, q));
        L31[PADX(kidx)] = b;
        L31[PADX(mkidx)] = d;
      }
    }
    return true;
  });
  step_sync<WAVE>();
}

// Squaring across two LDS blocks B (at 0) and B2 = 3*2^a - 1 - B (at Cb):
// position B*Cb + o pairs with B2*Cb + (Cb - 1 - o). Twiddle index ti: block
// B's even positions (Cb/2) followed by block B2's (Cb/2); s61/s31 as above.
template <bool SPLIT, int KMAX = 4, class S61, class S31>
__device__ __forceinline__ void square_cross(e61* L61, e31* L31, int logCb, S61 s61, S31 s31) {
  const Lanes<SPLIT> ln;
  const int Cb = 1 << logCb;
  for_k<KMAX>([&](int k) {
    const int o = ln.tid + k * ln.n;
    if (o >= Cb) return false;
    int kidx, mkidx, ti;
    if ((o & 1) == 0) {
      kidx = o;
      mkidx = Cb + (Cb - 1 - o);
      ti = o >> 1;
    } else {
      kidx = Cb + (Cb - 1 - o);
      mkidx = o;
      ti = (Cb >> 1) + ((Cb - 1 - o) >> 1);
    }
    if (ln.f61) {
      e61 a = L61[PADX(kidx)], c = L61[PADX(mkidx)];
      square_pair(a, c, s61(k, ti));
      L61[PADX(kidx)] = a;
      L61[PADX(mkidx)] = c;
    }
    if (ln.f31) {
      e31 b = L31[PADX(kidx)], d = L31[PADX(mkidx)];
      square_pair(b, d, s31(k, ti));
      L31[PADX(kidx)] = b;
      L31[PADX(mkidx)] = d;
    }
    return true;
  });
  __syncthreads();
}

// A pass's first step with its inputs read straight from global memory
// (LOAD), or its last step with its outputs written straight there (!LOAD):
// the same butterflies, LDS layout and values as lds_step (SPLIT lanes), with
// the thread's own elements at global index gidx(LDS index) and its twiddles
// read from the global tables at twi(b, blk) + e - 1. A first step so needs
// no barrier before it (each wave starts on its own loads as they arrive),
// and a last step no LDS write-back and store loop after it. NTM: the
// spectrum is read with streaming loads (gld_s).
template <int NT, bool FWD, int RADIX, bool LOAD, bool NTM = false, class GI, class TI>
__device__ __forceinline__ void edge_step(e61* L61, e31* L31, int logR, int logW, int nb, int lam,
                                          M2_GLOBAL e61* z61, M2_GLOBAL e31* z31, GI gidx,
                                          const e61* __restrict__ tw61, const e31* __restrict__ tw31, TI twi) {
  constexpr int LR = RADIX == 8 ? 3 : (RADIX == 4 ? 2 : 1);
  // butterflies per thread (radix 8 on 1024 threads: one, on half the threads)
  constexpr int KMAX = PASS_MAX_ELEMS / RADIX / (NT / 2) > 0 ? PASS_MAX_ELEMS / RADIX / (NT / 2) : 1;
  const Lanes<true> ln;
  const int logd = logR - lam - LR;
  const int logU = logR - LR + logW;
  const int total = nb << logU;
  const int stride = 1 << (logd + logW);
  int base[KMAX], tw[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int t = ln.tid + k * ln.n;
    const int b = t >> logU;
    const int u = t & ((1 << logU) - 1);
    const int col = u & ((1 << logW) - 1);
    const int g = u >> logW;
    const int off = g & ((1 << logd) - 1);
    const int blk = g >> logd;
    base[k] = (b << (logR + logW)) + (((blk << (logR - lam)) + off) << logW) + col;
    tw[k] = (int)twi(b, blk);
  }
  if (ln.f61) {
    e61 v[KMAX][RADIX];
    if (LOAD) {
      // every load first, so their latencies overlap
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
        if (ln.tid + k * ln.n < total) {
#pragma unroll
          for (int e = 0; e < RADIX; ++e) v[k][e] = gld_s<NTM>(z61 + gidx(base[k] + e * stride));
        }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (ln.tid + k * ln.n >= total) break;
      if (!LOAD) {
#pragma unroll
        for (int e = 0; e < RADIX; ++e) v[k][e] = L61[PADX(base[k] + e * stride)];
      }
      const M2_GLOBAL e61* T = gptr(tw61 + tw[k]);
      if (FWD) {
#pragma unroll
        for (int e = 1; e < RADIX; ++e) v[k][e] = mul_tw(v[k][e], gld(T + e - 1));
        if (RADIX == 8) dft8_fwd_lazy(v[k]);
        if (RADIX == 4) dft4_fwd_lazy(v[k]);
        if (RADIX == 2) dft2_fwd_lazy(v[k][0], v[k][1]);
      } else {
        if (RADIX == 8) dft8_inv_lazy(v[k]);
        if (RADIX == 4) dft4_inv_lazy(v[k]);
        if (RADIX == 2) dft2_inv_lazy(v[k][0], v[k][1]);
#pragma unroll
        for (int e = 1; e < RADIX; ++e) v[k][e] = mulconj_tw(v[k][e], gld(T + e - 1));
      }
#pragma unroll
      for (int e = 0; e < RADIX; ++e) {
        if (LOAD) L61[PADX(base[k] + e * stride)] = v[k][e];
        else gst(z61 + gidx(base[k] + e * stride), v[k][e]);
      }
    }
  } else {
    e31 x[KMAX][RADIX];
    if (LOAD) {
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
        if (ln.tid + k * ln.n < total) {
#pragma unroll
          for (int e = 0; e < RADIX; ++e) x[k][e] = gld_s<NTM>(z31 + gidx(base[k] + e * stride));
        }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (ln.tid + k * ln.n >= total) break;
      if (!LOAD) {
#pragma unroll
        for (int e = 0; e < RADIX; ++e) x[k][e] = L31[PADX(base[k] + e * stride)];
      }
      const M2_GLOBAL e31* T = gptr(tw31 + tw[k]);
      if (FWD) {
#pragma unroll
        for (int e = 1; e < RADIX; ++e) x[k][e] = mul_tw(x[k][e], gld(T + e - 1));
        if (RADIX == 8) dft8_fwd(x[k]);
        if (RADIX == 4) dft4_fwd(x[k]);
        if (RADIX == 2) dft2(x[k][0], x[k][1]);
      } else {
        if (RADIX == 8) dft8_inv(x[k]);
        if (RADIX == 4) dft4_inv(x[k]);
        if (RADIX == 2) dft2(x[k][0], x[k][1]);
#pragma unroll
        for (int e = 1; e < RADIX; ++e) x[k][e] = mulconj_tw(x[k][e], gld(T + e - 1));
      }
#pragma unroll
      for (int e = 0; e < RADIX; ++e) {
        if (LOAD) L31[PADX(base[k] + e * stride)] = x[k][e];
        else gst(z31 + gidx(base[k] + e * stride), x[k][e]);
      }
    }
  }
}

// edge_step for step si of pass ps (radix dispatch). No barrier: the caller
// adds it where the LDS results are read by other threads.
template <int NT, bool FWD, bool LOAD, class GI, class TI>
__device__ __forceinline__ void edge_pass_step(e61* L61, e31* L31, int logW, int nb, const Pass& ps, int si,
                                               M2_GLOBAL e61* z61, M2_GLOBAL e31* z31, GI gidx,
                                               const e61* tw61, const e31* tw31, TI twi) {
  const Step st = ps.steps[si];
  const int lam = st.ell - ps.ell0;
  if (st.radix == 8) edge_step<NT, FWD, 8, LOAD>(L61, L31, ps.A, logW, nb, lam, z61, z31, gidx, tw61, tw31, twi);
  else if (st.radix == 4) edge_step<NT, FWD, 4, LOAD>(L61, L31, ps.A, logW, nb, lam, z61, z31, gidx, tw61, tw31, twi);
  else edge_step<NT, FWD, 2, LOAD>(L61, L31, ps.A, logW, nb, lam, z61, z31, gidx, tw61, tw31, twi);
}

// tw[first + k] for a runtime k < N, by selects, so that the register array
// is never indexed dynamically (which would move it to scratch memory).
template <int N, int NS>
__device__ __forceinline__ e61 pick_slot(const e61 (&tw)[NS], int first, int k) {
  e61 r = tw[first];
#pragma unroll
  for (int j = 1; j < N; ++j) {
    e61 c = tw[first + j];
    // opaque copies: keeps the compiler from turning the selects back into a
    // dynamically indexed load
    asm volatile("" : "+v"(c.a), "+v"(c.b));
    r.a = k == j ? c.a : r.a;
    r.b = k == j ? c.b : r.b;
  }
  return r;
}

// The centre's last forward step (FWD) or first inverse step with this
// thread's twiddles in registers: iteration k's butterfly uses slots
// k (r - 1) .. k (r - 1) + r - 2 (M61 lanes: tw[s]; M31 lanes: the e31
// packed in tw[s].a). One column, nb blocks, as lds_step with logW = 0.
// Radix 8 (one butterfly per thread) or 4 (at most two on 512 threads, one
// on 1024).
template <bool FWD, int RADIX, int NT>
__device__ __forceinline__ void reg_step(e61* L61, e31* L31, int logR, int nb, int lam,
                                         const e61 (&tw)[centre_tw_slots<NT>()]) {
  constexpr int LR = RADIX == 8 ? 3 : (RADIX == 4 ? 2 : 1);
  constexpr int KMAX = centre_slots(NT).step_slots / (RADIX - 1);
  static_assert(KMAX >= 1, "a thread's last-step twiddles fit its slots");
  const Lanes<true> ln;
  const int logd = logR - lam - LR;
  const int logU = logR - LR;
  const int total = nb << logU;
  static_assert(RADIX == 8 || RADIX == 4, "the centre's last step is radix 8 or 4");
  // twiddle e (1 <= e < RADIX) of iteration k (radix 4: two iterations)
  auto twk = [&](int k, int e) -> e61 {
    if (KMAX == 1) return tw[e - 1];
    const e61 lo = tw[e - 1];
    e61 hi = tw[RADIX - 1 + e - 1];
    asm volatile("" : "+v"(hi.a), "+v"(hi.b));
    return e61{k ? hi.a : lo.a, k ? hi.b : lo.b};
  };
#pragma unroll 1
  for (int k = 0; k < KMAX; ++k) {
    const int t = ln.tid + k * ln.n;
    if (t >= total) break;
    const int b = t >> logU;
    const int u = t & ((1 << logU) - 1);
    const int off = u & ((1 << logd) - 1);
    const int blk = u >> logd;
    const int base = (b << logR) + (blk << (logR - lam)) + off;
    const int stride = 1 << logd;
    if (ln.f61) {
      e61 v[RADIX];
#pragma unroll
      for (int e = 0; e < RADIX; ++e) v[e] = L61[PADX(base + e * stride)];
      if (FWD) {
#pragma unroll
        for (int e = 1; e < RADIX; ++e) v[e] = mul_tw(v[e], twk(k, e));
        if (RADIX == 8) dft8_fwd_lazy(v);
        if (RADIX == 4) dft4_fwd_lazy(v);
        if (RADIX == 2) dft2_fwd_lazy(v[0], v[1]);
      } else {
        if (RADIX == 8) dft8_inv_lazy(v);
        if (RADIX == 4) dft4_inv_lazy(v);
        if (RADIX == 2) dft2_inv_lazy(v[0], v[1]);
#pragma unroll
        for (int e = 1; e < RADIX; ++e) v[e] = mulconj_tw(v[e], twk(k, e));
      }
#pragma unroll
      for (int e = 0; e < RADIX; ++e) L61[PADX(base + e * stride)] = v[e];
    } else {
      e31 x[RADIX];
#pragma unroll
      for (int e = 0; e < RADIX; ++e) x[e] = L31[PADX(base + e * stride)];
      auto w31 = [&](int e) {
        const e61 w = twk(k, e);
        return e31{(uint32_t)w.a, (uint32_t)(w.a >> 32)};
      };
      if (FWD) {
#pragma unroll
        for (int e = 1; e < RADIX; ++e) x[e] = mul_tw(x[e], w31(e));
        if (RADIX == 8) dft8_fwd(x);
        if (RADIX == 4) dft4_fwd(x);
        if (RADIX == 2) dft2(x[0], x[1]);
      } else {
        if (RADIX == 8) dft8_inv(x);
        if (RADIX == 4) dft4_inv(x);
        if (RADIX == 2) dft2(x[0], x[1]);
#pragma unroll
        for (int e = 1; e < RADIX; ++e) x[e] = mulconj_tw(x[e], w31(e));
      }
#pragma unroll
      for (int e = 0; e < RADIX; ++e) L31[PADX(base + e * stride)] = x[e];
    }
  }
  __syncthreads();
}

template <bool FWD, int NT>
__device__ __forceinline__ void centre_last_step(e61* L61, e31* L31, const Pass& ps, int nb,
                                                 const e61 (&tw)[centre_tw_slots<NT>()]) {
  const Step st = ps.steps[ps.nsteps - 1];
  const int lam = st.ell - ps.ell0;
  // the centre has at least 2 levels, so its last step is radix 8 or 4, and
  // radix 4 on 1024 threads (the host refuses other plans)
  if constexpr (NT == MP_THREADS_NARROW) {
    if (st.radix == 8) {
      reg_step<FWD, 8, NT>(L61, L31, ps.A, nb, lam, tw);
      return;
    }
  }
  reg_step<FWD, 4, NT>(L61, L31, ps.A, nb, lam, tw);
}

// Copies count elements of both fields, global index src(i) -> LDS index
// dst(i), with each thread issuing B loads before its LDS stores so that the
// loads' latencies overlap instead of adding up.
template <int B, class G61, class G31, class Src, class Dst>
__device__ __forceinline__ void gather2(e61* L61, e31* L31, G61 g61, G31 g31, int count, Src src, Dst dst) {
  // Loads are unconditional (in the last, partial batch lanes past the end
  // re-read the last element) so that the staged values stay in registers;
  // full batches store unmasked.
  const int step = B * (int)blockDim.x;
  for (int base = 0; base < count; base += step) {
    const bool full = base + step <= count;
    uint64_t a[B], b[B];
    uint32_t c[B], d[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int i = base + k * (int)blockDim.x + (int)threadIdx.x;
      const uint32_t j = src(full ? i : min(i, count - 1));
      a[k] = g61[j].a;
      b[k] = g61[j].b;
      c[k] = g31[j].a;
      d[k] = g31[j].b;
    }
    if (full) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const int j = dst(base + k * (int)blockDim.x + (int)threadIdx.x);
        L61[j] = e61{a[k], b[k]};
        L31[j] = e31{c[k], d[k]};
      }
    } else {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const int i = base + k * (int)blockDim.x + (int)threadIdx.x;
        if (i < count) {
          const int j = dst(i);
          L61[j] = e61{a[k], b[k]};
          L31[j] = e31{c[k], d[k]};
        }
      }
    }
  }
}

// The same for one e61 array.
template <int B, class G61, class Src, class Dst>
__device__ __forceinline__ void gather1(e61* L61, G61 g61, int count, Src src, Dst dst) {
  const int step = B * (int)blockDim.x;
  for (int base = 0; base < count; base += step) {
    const bool full = base + step <= count;
    uint64_t a[B], b[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int i = base + k * (int)blockDim.x + (int)threadIdx.x;
      const uint32_t j = src(full ? i : min(i, count - 1));
      a[k] = g61[j].a;
      b[k] = g61[j].b;
    }
    if (full) {
#pragma unroll
      for (int k = 0; k < B; ++k) L61[dst(base + k * (int)blockDim.x + (int)threadIdx.x)] = e61{a[k], b[k]};
    } else {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const int i = base + k * (int)blockDim.x + (int)threadIdx.x;
        if (i < count) L61[dst(i)] = e61{a[k], b[k]};
      }
    }
  }
}

__device__ __forceinline__ void stage_range(e61* S61, e31* S31, const e61* __restrict__ g61,
                                            const e31* __restrict__ g31, uint32_t g0, int count) {
  for (int k = threadIdx.x; k < count; k += blockDim.x) {
    S61[k] = g61[g0 + k];
    S31[k] = g31[g0 + k];
  }
}

// ---------------------------------------------------------------------------
// Small plan: one workgroup holds the whole residue (h <= 2^LOGH) and runs
// `iters` iterations without leaving LDS. Digits are in natural order. NTH
// threads, half per field (M61 and M31 waves side by side, as in the
// multi-pass kernels): 512 for h <= 2^SMALL_LOG_H (at h = 2^12 each thread
// owns two carry segments and does two butterflies per radix-8 step), or,
// for h <= 2^8, 128 —
// one wave per field, which then holds its field's whole transform, so the
// radix steps and the squaring need no workgroup barrier (step_sync), only
// the carry, which reads both fields, does; its LDS (17 KB) also lets several
// tests of a batch share a CU. (At h = 2^9 a lane would take two to four
// butterflies per step in turn: 20.3 against 13.4 us/iteration with 512
// threads.)
constexpr int SMALL_THREADS = 512;
constexpr int SMALL_WAVE_LOG_H = 8;
constexpr int SMALL_WAVE_THREADS = 128;
// Up to h = 2^11 every twiddle of the small plan is staged in LDS (data 54 KB
// padded, twiddles 72 KB). At h = 2^12 the padded data (108 KB) and the
// carries (16 KB) leave room for SMALL_LDS_TW twiddles: the tables of the
// leading steps (radix 8: 511 entries, radix 4: 1023) are staged, while the
// last step's table and the squaring twiddles are read from global memory,
// where all tests of the length share them (they stay in L2).
constexpr int SMALL_LDS_LOG_H = 11;
constexpr int SMALL_LDS_TW = 1024;

// The single-workgroup plan's default transform of 2^L points, its steps at
// compile time: radix 4 up to L = 10 (a radix-2 step first when L is odd),
// radix 8 above (a radix-2 or radix-4 step first for L mod 3 = 1 or 2), as
// make_pass lays them out. Step K at relative level LAM has its twiddles at
// offset OFF of the flat table ((r - 1) 2^LAM entries per step, back to back)
// and is run when K0 <= K < K1; level 0 has no twiddle products. Forward runs
// the steps in order, inverse in reverse.
constexpr int small_maxr(int L) { return L <= 10 ? 4 : 8; }
constexpr int small_first_radix(int L) {
  return small_maxr(L) == 4 ? (L % 2 ? 2 : 4) : (L % 3 == 1 ? 2 : (L % 3 == 2 ? 4 : 8));
}
template <bool FWD, int L, bool WAVE, int K0, int K1, int K = 0, int LAM = 0, int OFF = 0>
__device__ __forceinline__ void small_steps(e61* L61, e31* L31, const e61* T61, const e31* T31) {
  constexpr int R = K == 0 ? small_first_radix(L) : small_maxr(L);
  constexpr int LR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  constexpr bool MORE = LAM + LR < L;
  if constexpr (!FWD && MORE)
    small_steps<FWD, L, WAVE, K0, K1, K + 1, LAM + LR, OFF + (1 << LAM) * (R - 1)>(L61, L31, T61, T31);
  if constexpr (K >= K0 && K < K1)
    lds_step<FWD, R, true, (LAM != 0), WAVE>(L61, L31, L, 0, 1, LAM, T61 + OFF, T31 + OFF);
  if constexpr (FWD && MORE)
    small_steps<FWD, L, WAVE, K0, K1, K + 1, LAM + LR, OFF + (1 << LAM) * (R - 1)>(L61, L31, T61, T31);
}
// Steps of that plan, and how many of its leading steps have their tables
// within the first `cap` entries.
constexpr int small_nsteps(int L) {
  return small_maxr(L) == 4 ? (L + 1) / 2 : (L + 2) / 3;
}
constexpr int small_lds_steps(int L, int cap) {
  int k = 0, lam = 0, lo = 0;
  while (lam < L) {
    const int r = k == 0 ? small_first_radix(L) : small_maxr(L);
    const int span = (1 << lam) * (r - 1);
    if (lo + span > cap) break;
    lo += span;
    lam += r == 8 ? 3 : (r == 4 ? 2 : 1);
    ++k;
  }
  return k;
}

// CL > 0: the transform is known at compile time to be 2^CL points in the
// default plan (small_steps), so its step geometry is constant; CL = 0 reads
// the steps from the pass.
template <int LOGH, int NTH, int CL = 0>
__global__ void __launch_bounds__(NTH) k_small(const TestDesc* __restrict__ tests, const e61* __restrict__ tw61,
                                               const e31* __restrict__ tw31, const e61* __restrict__ sq61,
                                               const e31* __restrict__ sq31, Pass ps) {
  constexpr bool WAVE = NTH == SMALL_WAVE_THREADS;
  constexpr int NSEG = (1 << LOGH) / SEG_ELEMS;
  constexpr int SPT = NSEG > NTH ? NSEG / NTH : 1;  // carry segments per thread
  static_assert(NSEG <= SPT * NTH, "every carry segment has a thread");
  static_assert(!WAVE || LOGH <= SMALL_WAVE_LOG_H, "one wave per field holds the transform");
  constexpr bool GTW = LOGH > SMALL_LDS_LOG_H;  // part of the twiddles read from global memory
  // the compiled-in plans of n >= 16 (CL >= 3) only meet 16 <= q <= 42: their
  // carry has no 128-bit form (fewer registers, less spilling at h = 4096)
  constexpr bool GEN = CL < 3;
  constexpr int TN = GTW ? SMALL_LDS_TW : (1 << LOGH);
  constexpr int SN = GTW ? 1 : (1 << (LOGH - 1));
  // squaring pairs per thread (h/2 pairs over NTH/2 threads per field)
  constexpr int SQK = (1 << (LOGH - 1)) / (NTH / 2) > 4 ? (1 << (LOGH - 1)) / (NTH / 2) : 4;
  __shared__ e61 L61[padded(1 << LOGH)];
  __shared__ e31 L31[padded(1 << LOGH)];
  __shared__ e61 T61[TN];
  __shared__ e31 T31[TN];
  __shared__ e61 S61[SN];
  __shared__ e31 S31[SN];
  __shared__ Carry128 gs[NSEG];
  static_assert(CL == 0 || CL <= LOGH, "a constant plan of at most 2^LOGH points");
  const TestDesc td = tests[blockIdx.x];
  M2_GLOBAL e61* __restrict__ z61 = gptr(td.z61);
  const DigitGen D = td.dgen;
  const int logH = CL ? CL : ps.A;
  const int h = 1 << logH;
  const int nseg = h / SEG_ELEMS;
  // steps [0, kl) take their twiddles from LDS, the rest (GTW) from global
  // memory at the same offsets
  int kl = ps.nsteps;
  const e61* G61 = tw61 + ps.steps[0].twoff;
  const e31* G31 = tw31 + ps.steps[0].twoff;
  if constexpr (GTW) {
    int lo = 0;
    kl = 0;
    for (; kl < ps.nsteps; ++kl) {
      const int sp = twiddle_span(ps.steps[kl], ps.steps[kl].ell);
      if (lo + sp > TN) break;
      lo += sp;
    }
    gather2<1>(T61, T31, G61, G31, lo, [](int i) { return (uint32_t)i; }, [](int i) { return i; });
  } else {
    const uint32_t j0[1] = {0};
    stage_pass_twiddles(T61, T31, ps, 1, j0, tw61, tw31);
    stage_range(S61, S31, sq61, sq31, 0, h >> 1);
  }
  // thread t owns carry segments t SPT .. t SPT + SPT - 1 (segment s: digits
  // 8s..8s+7) for the whole launch: their digits stay in registers between
  // iterations
  const int s0 = threadIdx.x * SPT;
  SegDigits sd[SPT];
  SegMeta M[SPT];
#pragma unroll
  for (int i = 0; i < SPT; ++i) {
    if (s0 + i < nseg) seg_load(z61 + (s0 + i) * SEG_ELEMS, sd[i]);
    M[i] = seg_meta(D, 2u * (uint32_t)((s0 + i) * SEG_ELEMS));
  }
  const WidthMasks wm(D);
  for (uint32_t it = 0; it < td.iters; ++it) {
#pragma unroll
    for (int i = 0; i < SPT; ++i)
      if (s0 + i < nseg) seg_weight<GEN>(sd[i], L61, L31, (s0 + i) * SEG_ELEMS, M[i], D.q);
    __syncthreads();
    // constant plan: steps [0, KL) from LDS, the rest (GTW) from global memory
    constexpr int NSC = CL ? small_nsteps(CL) : 0, KL = GTW ? small_lds_steps(CL, TN) : NSC;
    if constexpr (CL != 0) small_steps<true, CL, WAVE, 0, KL>(L61, L31, T61, T31);
    else lds_pass<true, true, true, WAVE>(L61, L31, 0, 1, ps, T61, T31, 0, kl);
    if constexpr (GTW) {
      if constexpr (CL != 0) small_steps<true, CL, WAVE, KL, NSC>(L61, L31, G61, G31);
      else lds_pass<true, true, true, WAVE>(L61, L31, 0, 1, ps, G61, G31, kl);
      square_self<true, WAVE, SQK>(L61, L31, logH, 1, [&](int, int q) { return sq61[q]; },
                                   [&](int, int q) { return sq31[q]; });
      if constexpr (CL != 0) small_steps<false, CL, WAVE, KL, NSC>(L61, L31, G61, G31);
      else lds_pass<false, true, true, WAVE>(L61, L31, 0, 1, ps, G61, G31, 0, ps.nsteps - kl);
    } else {
      square_self<true, WAVE, SQK>(L61, L31, logH, 1, [&](int, int q) { return S61[q]; },
                                   [&](int, int q) { return S31[q]; });
    }
    if constexpr (CL != 0) small_steps<false, CL, WAVE, 0, KL>(L61, L31, T61, T31);
    else lds_pass<false, true, true, WAVE>(L61, L31, 0, 1, ps, T61, T31, ps.nsteps - kl);
    if (WAVE) __syncthreads();  // the carry reads both fields
#pragma unroll
    for (int i = 0; i < SPT; ++i) {
      const int s = s0 + i;
      if (s < nseg) {
        const i128 c0 = (td.mode == 0 && s == 0) ? (i128)(-2) : (i128)0;
        store_carry(gs, s, seg_normalise<GEN>(L61, L31, s * SEG_ELEMS, M[i], wm, D.q, c0, td.mode == 0 ? ~0ull : 0ull,
                                         sd[i]));
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SPT; ++i)
      if (s0 + i < nseg) seg_ripple<GEN>(sd[i], M[i], wm, D.q, load_carry(gs, (s0 + i + nseg - 1) % nseg));
  }
#pragma unroll
  for (int i = 0; i < SPT; ++i)
    if (s0 + i < nseg) seg_store(z61 + (s0 + i) * SEG_ELEMS, sd[i]);
}

// ---------------------------------------------------------------------------
// Large plan, outer pass: levels [0, A) over columns of stride S = 2^logS.
// Workgroup cg owns the W = 2^logW adjacent columns [cg W, cg W + W) of all R
// rows; row r of the strip is a carry segment run (W/4 sub-segments chained
// inside the workgroup). Digits of element (r, cg W + c) live at
// z61[(cg R + r) W + c] between iterations; the carry leaving strip-row (r, cg)
// is G[cg R + r] and enters the next strip-row in digit order.
__device__ __forceinline__ uint32_t prev_strip_carry(uint32_t cg, uint32_t r, uint32_t R, uint32_t NS) {
  // producer of the carry into strip-row (r, cg): (r, cg-1), or (r-1, NS-1), or (R-1, NS-1)
  if (cg > 0) return (cg - 1) * R + r;
  return (NS - 1) * R + (r ? r - 1 : R - 1);
}


// Workgroups are dealt to the 8 XCDs round-robin by their linear index, and
// each XCD has its own L2. Neighbouring column strips share cache lines (a
// strip row is 64 bytes of M61 and 32 of M31 data), so strip c is given to
// workgroup index b with c = (b mod 8) (n/8) + b/8: each XCD then works on
// one contiguous range of strips and whole lines stay in one L2. (Needs n a
// multiple of 8; gridDim.x is n, so the XCD of a workgroup does not depend
// on blockIdx.y.)
constexpr uint32_t NUM_XCD = 8;
__device__ __forceinline__ uint32_t xcd_strip(uint32_t b, uint32_t n) {
  return (n % NUM_XCD) ? b : (b % NUM_XCD) * (n / NUM_XCD) + b / NUM_XCD;
}

// A strided pass over R rows stages R - 1 twiddles; with R <= 2^9 the outer
// and middle passes fit two workgroups per CU (74 KB of LDS each), so one
// workgroup's loads and stores overlap another's arithmetic.
constexpr int OUTER_TW = 1 << PASS_MAX_LOG_R;

// CA > 0: the pass is known at compile time to be A = CA levels over W =
// 2^CLOGW columns of stride S = 2^CLOGS (the shapes of OUTER_SHAPES), in
// radix-4, 8 steps (A = 5), radix-8, 8 (6), radix-2, 8, 8 (7), radix-4, 8, 8
// (8) or radix-8, 8, 8 (9), so its step geometry and index arithmetic are
// constants; CA = 0 reads it from the arguments.
template <int CA = 0, int CLOGW = 0, int CLOGS = 0>
__global__ void __launch_bounds__(MP_THREADS) k_outer_fwd(const TestDesc* __restrict__ tests,
                                                          const e61* __restrict__ tw61, const e31* __restrict__ tw31,
                                                          Pass ps, int logS, int logW) {
  __shared__ e61 L61[padded(PASS_MAX_ELEMS)];
  __shared__ e31 L31[padded(PASS_MAX_ELEMS)];
  __shared__ e61 T61[OUTER_TW];
  __shared__ e31 T31[OUTER_TW];
  const TestDesc td = tests[blockIdx.y];
  const M2_GLOBAL e61* __restrict__ zd = gptr(td.zd);
  M2_GLOBAL e61* __restrict__ z61 = gptr(td.z61);
  M2_GLOBAL e31* __restrict__ z31 = gptr(td.z31);
  const M2_GLOBAL Carry128* __restrict__ G = gptr(td.G);
  static_assert(CA == 0 || (CA >= 5 && CA <= 9), "the constant forms are radix 4, 8; 8, 8; 2, 8, 8; 4, 8, 8; 8, 8, 8");
  if (CA) logW = CLOGW, logS = CLOGS;
  const DigitGen D = td.dgen;
  const int R = 1 << (CA ? CA : ps.A), W = 1 << logW;
  const uint32_t NS = (1u << logS) >> logW, cg = xcd_strip(blockIdx.x, NS);
  const uint32_t col0 = cg << logW;
  // each thread takes one carry segment: its digits straight from the
  // strip-major digit buffer (64 contiguous bytes), the carry entering its
  // strip row if it is the row's first segment, then the weights into LDS
  FlatTwiddles ft;
  const int ntw = outer_twiddles(ft, ps, tw61, tw31);
  {
    const int subs = W / SEG_ELEMS;
    const int it = threadIdx.x;
    if (it < R * subs) {
      const int r = it / subs, v = it - r * subs;
      const int idx0 = r * W + v * SEG_ELEMS;
      const uint32_t e0 = ((uint32_t)r << logS) + col0 + (uint32_t)(v * SEG_ELEMS);
      SegDigits sd;
      seg_load(zd + (size_t)cg * (R * W) + idx0, sd);
      const SegMeta M = seg_meta(D, 2u * e0);
      if (v == 0) seg_ripple(sd, M, WidthMasks(D), D.q, carry_at(G, prev_strip_carry(cg, (uint32_t)r, R, NS)));
      seg_weight(sd, L61, L31, idx0, M, D.q);
    }
  }
  ft.store(T61, T31, ntw);
  __syncthreads();
  // the last step (not level 0 when there are two or more) writes global
  // memory directly
  auto gidx = [&](int i) { return ((uint32_t)(i >> logW) << logS) + col0 + (uint32_t)(i & (W - 1)); };
  if constexpr (CA == 5) {
    // relative levels 0 (no twiddles) and 2; spans 3, 28 back to back
    lds_step<true, 4, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
    const uint32_t off = ps.steps[1].twoff;
    edge_step<MP_THREADS, true, 8, false>(L61, L31, CA, CLOGW, 1, 2, z61, z31, gidx, tw61, tw31,
                                          [&](int, int blk) { return off + (uint32_t)blk * 7u; });
    return;
  } else if constexpr (CA == 6) {
    // relative levels 0 (no twiddles) and 3; spans 7, 56 back to back
    lds_step<true, 8, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
    const uint32_t off = ps.steps[1].twoff;
    edge_step<MP_THREADS, true, 8, false>(L61, L31, CA, CLOGW, 1, 3, z61, z31, gidx, tw61, tw31,
                                          [&](int, int blk) { return off + (uint32_t)blk * 7u; });
    return;
  } else if constexpr (CA == 7) {
    // relative levels 0 (no twiddles), 1 and 4; spans 1, 14, 112 back to back
    lds_step<true, 2, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
    lds_step<true, 8, true>(L61, L31, CA, CLOGW, 1, 1, T61 + 1, T31 + 1);
    const uint32_t off = ps.steps[2].twoff;
    edge_step<MP_THREADS, true, 8, false>(L61, L31, CA, CLOGW, 1, 4, z61, z31, gidx, tw61, tw31,
                                          [&](int, int blk) { return off + (uint32_t)blk * 7u; });
    return;
  } else if constexpr (CA == 8) {
    // relative levels 0 (no twiddles), 2 and 5; spans 3, 28, 224 back to back
    lds_step<true, 4, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
    lds_step<true, 8, true>(L61, L31, CA, CLOGW, 1, 2, T61 + 3, T31 + 3);
    const uint32_t off = ps.steps[2].twoff;
    edge_step<MP_THREADS, true, 8, false>(L61, L31, CA, CLOGW, 1, 5, z61, z31, gidx, tw61, tw31,
                                          [&](int, int blk) { return off + (uint32_t)blk * 7u; });
    return;
  } else if constexpr (CA == 9) {
    // relative levels 0 (no twiddles), 3 and 6; spans 7, 56, 448 back to back
    lds_step<true, 8, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
    lds_step<true, 8, true>(L61, L31, CA, CLOGW, 1, 3, T61 + 7, T31 + 7);
    const uint32_t off = ps.steps[2].twoff;
    edge_step<MP_THREADS, true, 8, false>(L61, L31, CA, CLOGW, 1, 6, z61, z31, gidx, tw61, tw31,
                                          [&](int, int blk) { return off + (uint32_t)blk * 7u; });
    return;
  }
  const int ns = ps.nsteps;
  const bool edges = ns >= 2;
  lds_pass<true, true, true>(L61, L31, logW, 1, ps, T61, T31, 0, edges ? ns - 1 : ns);
  if (edges) {
    const Step sl = ps.steps[ns - 1];
    edge_pass_step<MP_THREADS, true, false>(L61, L31, logW, 1, ps, ns - 1, z61, z31, gidx, tw61, tw31,
                                [&](int, int blk) { return sl.twoff + (uint32_t)blk * (uint32_t)(sl.radix - 1); });
    return;
  }
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    gst(z61 + gidx(i), L61[PADX(i)]);
    gst(z31 + gidx(i), L31[PADX(i)]);
  }
}

// CA > 0: the pass is known at compile time to be A = CA levels over W =
// 2^CLOGW columns of stride S = 2^CLOGS (the shapes of OUTER_SHAPES), in
// radix-4, 8 steps (A = 5), radix-8, 8 (6), radix-2, 8, 8 (7), radix-4, 8, 8
// (8) or radix-8, 8, 8 (9), so its step geometry and index arithmetic are
// constants; CA = 0 reads it from the arguments.
template <int CA = 0, int CLOGW = 0, int CLOGS = 0>
__global__ void __launch_bounds__(MP_THREADS) k_outer_inv(const TestDesc* __restrict__ tests,
                                                          const e61* __restrict__ tw61, const e31* __restrict__ tw31,
                                                          Pass ps, int logS, int logW) {
  __shared__ e61 L61[padded(PASS_MAX_ELEMS)];
  __shared__ e31 L31[padded(PASS_MAX_ELEMS)];
  __shared__ e61 T61[OUTER_TW];
  __shared__ e31 T31[OUTER_TW];
  __shared__ Carry128 gs[PASS_MAX_ELEMS / SEG_ELEMS];
  const TestDesc td = tests[blockIdx.y];
  const M2_GLOBAL e61* __restrict__ z61 = gptr(td.z61);
  const M2_GLOBAL e31* __restrict__ z31 = gptr(td.z31);
  M2_GLOBAL e61* __restrict__ zd = gptr(td.zd);
  M2_GLOBAL Carry128* __restrict__ G = gptr(td.G);
  const DigitGen D = td.dgen;
  const int mode = td.mode;
  static_assert(CA == 0 || (CA >= 5 && CA <= 9), "the constant forms are radix 4, 8; 8, 8; 2, 8, 8; 4, 8, 8; 8, 8, 8");
  if (CA) logW = CLOGW, logS = CLOGS;
  const int R = 1 << (CA ? CA : ps.A), W = 1 << logW;
  const uint32_t cg = xcd_strip(blockIdx.x, (1u << logS) >> logW);
  const uint32_t col0 = cg << logW;
  FlatTwiddles ft;
  const int ntw = outer_twiddles(ft, ps, tw61, tw31);
  gather2<4>(L61, L31, z61, z31, R * W,
             [&](int i) { return ((uint32_t)(i >> logW) << logS) + col0 + (uint32_t)(i & (W - 1)); },
             [](int i) { return PADX(i); });
  ft.store(T61, T31, ntw);
  __syncthreads();
  if constexpr (CA == 5) {
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 2, T61 + 3, T31 + 3);
    lds_step<false, 4, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
  } else if constexpr (CA == 6) {
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 3, T61 + 7, T31 + 7);
    lds_step<false, 8, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
  } else if constexpr (CA == 7) {
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 4, T61 + 15, T31 + 15);
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 1, T61 + 1, T31 + 1);
    lds_step<false, 2, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
  } else if constexpr (CA == 8) {
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 5, T61 + 31, T31 + 31);
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 2, T61 + 3, T31 + 3);
    lds_step<false, 4, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
  } else if constexpr (CA == 9) {
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 6, T61 + 63, T31 + 63);
    lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 3, T61 + 7, T31 + 7);
    lds_step<false, 8, true, false>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
  } else {
    lds_pass<false, true, true>(L61, L31, logW, 1, ps, T61, T31);
  }
  // thread `it` owns carry segment it: normalise into registers, take the
  // carry of the segment before it in the strip row, write the row's last
  // carry to G; the digits go back through LDS so that the store to the
  // strip-major digit buffer is one 16-byte element per lane (a segment per
  // lane, 64 bytes apart, measured slower)
  const int subs = W / SEG_ELEMS;
  const int it = threadIdx.x;
  const bool mine = it < R * subs;
  const int r = it / subs, v = it - r * subs;
  const int idx0 = r * W + v * SEG_ELEMS;
  const uint32_t e0 = ((uint32_t)r << logS) + col0 + (uint32_t)(v * SEG_ELEMS);
  SegDigits sd;
  const SegMeta M = seg_meta(D, 2u * e0);
  const WidthMasks wm(D);
  if (mine) {
    const i128 c0 = (mode == 0 && e0 == 0) ? (i128)(-2) : (i128)0;
    store_carry(gs, it, seg_normalise(L61, L31, idx0, M, wm, D.q, c0, mode == 0 ? ~0ull : 0ull, sd));
  }
  __syncthreads();
  if (mine) {
    if (v > 0) seg_ripple(sd, M, wm, D.q, load_carry(gs, it - 1));
    if (v == subs - 1) gst(G + cg * (uint32_t)R + (uint32_t)r, gs[it]);
#pragma unroll
    for (int i = 0; i < SEG_ELEMS; ++i) L61[PADX(idx0 + i)] = e61{(uint64_t)sd.d[2 * i], (uint64_t)sd.d[2 * i + 1]};
  }
  __syncthreads();
  M2_GLOBAL e61* dst = zd + (size_t)cg * (R * W);
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) gst(dst + i, L61[PADX(i)]);
}

// The inverse outer pass of one iteration and the forward outer pass of the
// next in one launch, for a single test whose NS strips fit one workgroup per
// CU: the digits stay in LDS in between. The carry entering a strip row comes
// from the left neighbour strip (or, for strip 0, the last strip one row up).
// Each strip row's carry is published as two 64-bit words, each holding 56
// bits of the value under an 8-bit tag, the launch count mod 256; a 64-bit
// store is single-copy atomic, so a reader that sees both words tagged with
// its own launch's count has the whole value, with no flag, fence or cache
// write-back between writer and reader. The words are written and read with
// device-scope (relaxed) atomics, which go to the coherent level instead of
// an XCD's own L2. The wait is bounded: after spin_limit polls a reader gives
// up and raises sync[NS], which the host reports.
__device__ __forceinline__ void carry_put(M2_GLOBAL uint64_t* cf, uint32_t i, i128 c, uint32_t tag) {
  const uint64_t t = (uint64_t)tag << 56, m = (1ull << 56) - 1ull;
  uint64_t* w = (uint64_t*)(cf + 2 * i);
  __hip_atomic_store(w, ((uint64_t)c & m) | t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 1, ((uint64_t)(c >> 56) & m) | t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ i128 carry_take(M2_GLOBAL uint64_t* cf, uint32_t i, uint32_t tag, uint32_t limit,
                                           uint32_t* fault) {
  uint64_t* w = (uint64_t*)(cf + 2 * i);
  for (uint32_t n = 0;;) {
    const uint64_t w0 = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t w1 = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(w0 >> 56) == tag && (uint32_t)(w1 >> 56) == tag) {
      const int64_t hi = (int64_t)(w1 << 8) >> 8;  // upper 56 bits, sign-extended
      return ((i128)hi << 56) | (i128)(w0 & ((1ull << 56) - 1ull));
    }
    if (++n >= limit) {
      __hip_atomic_store(fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return 0;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// CA > 0: the pass is known at compile time to be A = CA levels in radix-8
// steps over W = 2^CLOGW columns of stride S = 2^CLOGS (the n = 2^20 single
// test: 9, 2, 10), so its step geometry and index arithmetic are constants,
// and its spectrum is read with streaming loads (gld_s);
// CA = 0 reads the geometry from the arguments.
template <int CA, int CLOGW, int CLOGS>
__global__ void __launch_bounds__(MP_THREADS) k_outer_fused(const TestDesc* __restrict__ tests,
                                                     const e61* __restrict__ tw61, const e31* __restrict__ tw31,
                                                     Pass ps, int logS_arg, int logW_arg, uint32_t spin_limit) {
  static_assert(CA == 0 || CA == 9, "the constant form covers three radix-8 steps");
  const int logS = CA ? CLOGS : logS_arg, logW = CA ? CLOGW : logW_arg;
  __shared__ e61 L61[padded(PASS_MAX_ELEMS)];
  __shared__ e31 L31[padded(PASS_MAX_ELEMS)];
  __shared__ e61 T61[OUTER_TW];
  __shared__ e31 T31[OUTER_TW];
  __shared__ Carry128 gs[PASS_MAX_ELEMS / SEG_ELEMS];
  const TestDesc td = tests[blockIdx.y];
  M2_GLOBAL e61* __restrict__ z61 = gptr(td.z61);
  M2_GLOBAL e31* __restrict__ z31 = gptr(td.z31);
  M2_GLOBAL uint64_t* __restrict__ cf = gptr(td.cf);
  M2_GLOBAL uint32_t* gsync = gptr(td.sync);
  uint32_t* sync = (uint32_t*)gsync;  // generic pointer to global memory, for the atomic builtins
  const DigitGen D = td.dgen;
  const int mode = td.mode;
  const int A = CA ? CA : ps.A;
  const int R = 1 << A, W = 1 << logW;
  const uint32_t NS = (1u << logS) >> logW, cg = xcd_strip(blockIdx.x, NS);
  const uint32_t col0 = cg << logW;
  // this launch's count: every thread reads the strip's counter before thread
  // 0 advances it (after the first barrier), and uses it only after the
  // inverse pass, so the load's latency is hidden behind the data loads
  // the M61 waves (first half) carry the longer dependency chains: they get
  // the SIMD's issue slots first
  if (threadIdx.x < blockDim.x / 2) __builtin_amdgcn_s_setprio(2);
  const uint32_t epoch = __hip_atomic_load(sync + cg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
  auto gidx = [&](int i) { return ((uint32_t)(i >> logW) << logS) + col0 + (uint32_t)(i & (W - 1)); };
  // the inverse pass's first step (the forward's last) reads global memory
  // directly, and the forward's last step writes it directly
  const int ns = CA ? 3 : ps.nsteps;
  const bool edges = ns >= 2;
  const Step sl = ps.steps[ns - 1];
  auto twl = [&](int, int blk) { return sl.twoff + (uint32_t)blk * (uint32_t)(sl.radix - 1); };
  FlatTwiddles ft;
  const int ntw = outer_twiddles(ft, ps, tw61, tw31);
  if (CA) edge_step<MP_THREADS, false, 8, true, true>(L61, L31, 9, CLOGW, 1, 6, z61, z31, gidx, tw61, tw31, twl);
  else if (edges)
    edge_pass_step<MP_THREADS, false, true>(L61, L31, logW, 1, ps, ns - 1, z61, z31, gidx, tw61, tw31, twl);
  else gather2<4>(L61, L31, z61, z31, R * W, gidx, [](int i) { return PADX(i); });
  ft.store(T61, T31, ntw);
  __syncthreads();
  // inverse outer pass, unweight + CRT, segment carries (as k_outer_inv);
  // thread `it` owns carry segment it and keeps its digits in registers
  // until the weighting. Constant form: steps at relative levels 6 (edge),
  // 3 and 0 (untwisted); twiddle blocks at 0 (level 0: 7), 7 (level 3: 56).
  if (CA) {
    lds_step<false, 8, true>(L61, L31, 9, CLOGW, 1, 3, T61 + 7, T31 + 7);
    lds_step<false, 8, true, false>(L61, L31, 9, CLOGW, 1, 0, T61, T31);
  } else {
    lds_pass<false, true, true>(L61, L31, logW, 1, ps, T61, T31, edges ? 1 : 0);
  }
  const int subs = W / SEG_ELEMS;
  const int it = threadIdx.x;
  const bool mine = it < R * subs;
  const int r = it / subs, v = it - r * subs;
  const int idx0 = r * W + v * SEG_ELEMS;
  const uint32_t e0 = ((uint32_t)r << logS) + col0 + (uint32_t)(v * SEG_ELEMS);
  SegDigits sd;
  const SegMeta M = seg_meta(D, 2u * e0);
  const WidthMasks wm(D);
  if (mine) {
    const i128 c0 = (mode == 0 && e0 == 0) ? (i128)(-2) : (i128)0;
    store_carry(gs, it, seg_normalise(L61, L31, idx0, M, wm, D.q, c0, mode == 0 ? ~0ull : 0ull, sd));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_store(sync + cg, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // spin_limit 0: report a time-out without waiting (the host's way to
    // exercise the report path)
    if (spin_limit == 0) __hip_atomic_store(sync + NS, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // the carry of the segment before this one in its strip row; the row's
  // last carry is published; the row's first segment then takes the carry
  // the neighbour strip published; weight and the forward outer pass follow
  if (mine) {
    if (v > 0) seg_ripple(sd, M, wm, D.q, load_carry(gs, it - 1));
    if (v == subs - 1) carry_put(cf, cg * (uint32_t)R + (uint32_t)r, load_carry(gs, it), epoch & 0xFFu);
    if (v == 0)
      seg_ripple(sd, M, wm, D.q,
                 carry_take(cf, prev_strip_carry(cg, r, R, NS), epoch & 0xFFu, spin_limit, sync + NS));
    seg_weight(sd, L61, L31, idx0, M, D.q);
  }
  __syncthreads();
  if (CA) {
    lds_step<true, 8, true, false>(L61, L31, 9, CLOGW, 1, 0, T61, T31);
    lds_step<true, 8, true>(L61, L31, 9, CLOGW, 1, 3, T61 + 7, T31 + 7);
    edge_step<MP_THREADS, true, 8, false, true>(L61, L31, 9, CLOGW, 1, 6, z61, z31, gidx, tw61, tw31, twl);
    return;
  }
  lds_pass<true, true, true>(L61, L31, logW, 1, ps, T61, T31, 0, edges ? ns - 1 : ns);
  if (edges) {
    edge_pass_step<MP_THREADS, true, false>(L61, L31, logW, 1, ps, ns - 1, z61, z31, gidx, tw61, tw31, twl);
    return;
  }
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    gst(z61 + gidx(i), L61[PADX(i)]);
    gst(z31 + gidx(i), L31[PADX(i)]);
  }
}

// Middle pass: levels [ell0, ell0 + A) with no carry work. Level-ell0 block
// j (of R * S elements) is split into S/W column groups.
// CA > 0: the pass is known at compile time to be A = CA levels in two steps,
// radix 8, 8 over W = 2^CLOGW columns of stride S = 2^CLOGS (n = 2^23 and
// 2^24: 6, 5, 10) or radix 4, 8 (n = 2^21 and 2^22: 5, 6, 10), so its step
// geometry and index arithmetic are constants; CA = 0 reads it from the
// arguments.
template <bool FWD, int CA = 0, int CLOGW = 0, int CLOGS = 0>
__global__ void __launch_bounds__(MP_THREADS) k_mid(const TestDesc* __restrict__ tests, const e61* __restrict__ tw61,
                                             const e31* __restrict__ tw31, Pass ps, int logS, int logW) {
  static_assert(CA == 0 || CA == 5 || CA == 6, "the constant forms are two steps: radix 4, 8 or 8, 8");
  __shared__ e61 L61[padded(PASS_MAX_ELEMS)];
  __shared__ e31 L31[padded(PASS_MAX_ELEMS)];
  __shared__ e61 T61[OUTER_TW];
  __shared__ e31 T31[OUTER_TW];
  const TestDesc td = tests[blockIdx.y];
  M2_GLOBAL e61* __restrict__ z61 = gptr(td.z61);
  M2_GLOBAL e31* __restrict__ z31 = gptr(td.z31);
  const int A = CA ? CA : ps.A, lw = CA ? CLOGW : logW, ls = CA ? CLOGS : logS;
  const int R = 1 << A, W = 1 << lw;
  const uint32_t NS = (1u << ls) >> lw;
  const uint32_t bx = xcd_strip(blockIdx.x, gridDim.x);
  const uint32_t jb = bx / NS, cg = bx - jb * NS;
  const uint32_t base = (jb << (A + ls)) + (cg << lw);
  const uint32_t j0[1] = {jb};
  // loads and stores go through LDS (direct global loads or stores at the
  // pass edges measured slower here, where two workgroups share a CU)
  auto gidx = [&](int i) { return base + ((uint32_t)(i >> lw) << ls) + (uint32_t)(i & (W - 1)); };
  gather2<4>(L61, L31, z61, z31, R * W, gidx, [](int i) { return PADX(i); });
  stage_pass_twiddles(T61, T31, ps, 1, j0, tw61, tw31);
  __syncthreads();
  if constexpr (CA == 5) {
    // relative levels 0 and 2; their twiddle spans (3 and 28) at offsets 0 and 3
    if constexpr (FWD) {
      lds_step<true, 4, true>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
      lds_step<true, 8, true>(L61, L31, CA, CLOGW, 1, 2, T61 + 3, T31 + 3);
    } else {
      lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 2, T61 + 3, T31 + 3);
      lds_step<false, 4, true>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
    }
  } else if constexpr (CA == 6) {
    // relative levels 0 and 3; their twiddle spans (7 and 56) at offsets 0 and 7
    if constexpr (FWD) {
      lds_step<true, 8, true>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
      lds_step<true, 8, true>(L61, L31, CA, CLOGW, 1, 3, T61 + 7, T31 + 7);
    } else {
      lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 3, T61 + 7, T31 + 7);
      lds_step<false, 8, true>(L61, L31, CA, CLOGW, 1, 0, T61, T31);
    }
  } else {
    lds_pass<FWD, true>(L61, L31, logW, 1, ps, T61, T31);
  }
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    gst(z61 + gidx(i), L61[PADX(i)]);
    gst(z31 + gidx(i), L31[PADX(i)]);
  }
}

// Centre pass: workgroup g squares the pair of contiguous blocks (B, B2) of
// C = 2^logC elements that the real-data pairing links; workgroup 0 takes
// blocks 0 and 1, which pair with themselves. Forward levels [ps.ell0, L),
// square, inverse. There are h / (2C) workgroups. The twiddles of all but the
// last step are staged in LDS; each thread loads the twiddles of its own
// last-step butterflies and squaring pairs straight into registers, so the
// workgroup needs 66 KB of LDS and two fit on a CU. CC > 0: the pass is
// known at compile time to be 2^CC-element blocks, on 1024 threads in five
// radix-4 steps with direct loads and stores at its edges (the n = 2^20
// single test), on 512 threads in radix-2, 8, 8, 8 steps (n >= 2^21, and
// batches), so its step geometry is constant; CC = 0 reads it from the
// arguments.
// The 512-thread forms are held to 128 VGPRs: two workgroups per CU.
template <int NT, int CC>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == MP_THREADS_NARROW ? 4 : 1))) k_centre(const TestDesc* __restrict__ tests, const e61* __restrict__ ctw61,
                                                const e31* __restrict__ ctw31, uint32_t crec, Pass ps,
                                                int direct_in) {
  // CC = 10: the step geometry of a 2^10-element block compiled in; on 1024
  // threads five radix-4 steps with direct first loads, on 512 threads radix
  // 2, 8, 8, 8 steps after a gather into LDS
  static_assert(CC == 0 || (CC == 10 && (NT == MP_THREADS_WIDE || NT == MP_THREADS_NARROW)),
                "the constant forms are the 2^10 centres");
  constexpr bool WIDE = NT == MP_THREADS_WIDE;
  // the constant 1024-thread form (the n = 2^20 single test, one workgroup
  // per CU) reads its spectrum and its register twiddles with streaming
  // loads (gld_s)
  constexpr bool STREAM = WIDE && CC != 0;
  __shared__ e61 L61[padded(2 << CENTRE_LOG_C)];
  __shared__ e31 L31[padded(2 << CENTRE_LOG_C)];
  __shared__ e61 T61[CENTRE_EARLY_TW];
  __shared__ e31 T31[CENTRE_EARLY_TW];
  const TestDesc td = tests[blockIdx.y];
  M2_GLOBAL e61* __restrict__ z61 = gptr(td.z61);
  M2_GLOBAL e31* __restrict__ z31 = gptr(td.z31);
  const int logC = CC ? CC : ps.A, C = 1 << logC;
  const uint32_t g = blockIdx.x;
  uint32_t B = 0, B2 = 1;
  const int nb = 2;
  if (g > 0) {
    const uint32_t a = 32u - __clz(g);  // octave: blocks [2^a, 2^(a+1))
    B = (1u << a) + (g - (1u << (a - 1)));
    B2 = 3u * (1u << a) - 1u - B;
  }
  auto gidx = [&](int i) { return ((i < C ? B : B2) << logC) + (uint32_t)(i & (C - 1)); };
  const e61* r61 = ctw61 + (size_t)g * crec;
  const e31* r31 = ctw31 + (size_t)g * crec;
  // With two steps or more, the last inverse step writes global memory
  // directly, and, with direct_in (a grid of at most one workgroup per CU,
  // where no other workgroup's arithmetic overlaps the loads), the first
  // forward step reads it directly too; step 0's twiddles lead the record
  // (block b's span, then block b+1's).
  const int ns = CC ? (WIDE ? 5 : 4) : ps.nsteps;
  const bool edges = ns >= 2;
  const bool edge_in = CC ? WIDE : edges && direct_in;
  const Step s0 = ps.steps[0];
  auto tw0 = [&](int b, int blk) {
    if (CC) return (uint32_t)(WIDE ? b * 3 + blk * 3 : b + blk);  // step 0: relative level 0, radix 4 or 2
    return (uint32_t)(b * twiddle_span(s0, s0.ell - ps.ell0) + blk * (s0.radix - 1));
  };
  // the LDS part of the record (nE <= CENTRE_EARLY_TW <= NT entries) is
  // loaded before the data and written to LDS after it; constant form: the
  // spans of steps 0..3 (3, 12, 48, 192 per block, two blocks) at offsets 0,
  // 6, 30 and 126, 510 in all (1024 threads); steps 0..2 (1, 14, 112) at
  // offsets 0, 2 and 30, 254 in all (512 threads)
  int nE = 0;
  if (CC) nE = WIDE ? 510 : 254;
  else
    for (int si = 0; si + 1 < ns; ++si) nE += nb * twiddle_span(ps.steps[si], ps.steps[si].ell - ps.ell0);
  // register twiddles: slot s of thread t at record entry nE + s * RT + t,
  // RT threads per field (coalesced); M31 lanes keep their e31 packed in .a.
  // The 1024-thread centre issues these loads first, so that they overlap the
  // data loads (its 5 slots fit beside the first step's registers); the
  // 512-thread centre's 11 slots would push it past 128 VGPRs, which two
  // workgroups per CU need, so it issues them after the first step.
  constexpr CentreSlots CS = centre_slots(NT);
  constexpr int NSLOT = centre_tw_slots<NT>();
  const Lanes<true> ln;
  // the M61 waves (first half) carry the longer dependency chains: they get
  // the SIMD's issue slots first
  if (threadIdx.x < blockDim.x / 2) __builtin_amdgcn_s_setprio(2);
  e61 tw[NSLOT];
  auto load_reg_twiddles = [&]() {
    if (ln.f61) {
      const M2_GLOBAL e61* src = gptr(r61 + nE) + ln.tid;
#pragma unroll
      for (int k = 0; k < NSLOT; ++k) tw[k] = gld_s<STREAM>(src + k * CS.role_threads);
    } else {
      const M2_GLOBAL e31* src = gptr(r31 + nE) + ln.tid;
#pragma unroll
      for (int k = 0; k < NSLOT; ++k) {
        const e31 w = gld_s<STREAM>(src + k * CS.role_threads);
        tw[k] = e61{(uint64_t)w.a | ((uint64_t)w.b << 32), 0};
      }
    }
  };
  if constexpr (NT == MP_THREADS_WIDE) load_reg_twiddles();
  FlatTwiddles ft;
  ft.load(r61, r31, nE);
  if constexpr (CC != 0 && WIDE) edge_step<NT, true, 4, true, STREAM>(L61, L31, CC, 0, 2, 0, z61, z31, gidx, r61, r31, tw0);
  else if (!CC && edge_in) edge_pass_step<NT, true, true>(L61, L31, 0, nb, ps, 0, z61, z31, gidx, r61, r31, tw0);
  else gather2<PASS_MAX_ELEMS / NT>(L61, L31, z61, z31, nb << logC, gidx, [](int i) { return PADX(i); });
  ft.store(T61, T31, nE);
  if constexpr (NT != MP_THREADS_WIDE) load_reg_twiddles();
  __syncthreads();
  if constexpr (CC != 0 && WIDE) {
    lds_step<true, 4, true>(L61, L31, CC, 0, 2, 2, T61 + 6, T31 + 6);
    lds_step<true, 4, true>(L61, L31, CC, 0, 2, 4, T61 + 30, T31 + 30);
    lds_step<true, 4, true>(L61, L31, CC, 0, 2, 6, T61 + 126, T31 + 126);
    reg_step<true, 4, NT>(L61, L31, CC, 2, 8, tw);
  } else if constexpr (CC != 0) {
    lds_step<true, 2, true>(L61, L31, CC, 0, 2, 0, T61, T31);
    lds_step<true, 8, true>(L61, L31, CC, 0, 2, 1, T61 + 2, T31 + 2);
    lds_step<true, 8, true>(L61, L31, CC, 0, 2, 4, T61 + 30, T31 + 30);
    reg_step<true, 8, NT>(L61, L31, CC, 2, 7, tw);
  } else {
    lds_pass<true, true>(L61, L31, 0, nb, ps, T61, T31, edge_in ? 1 : 0, ns - 1);
    centre_last_step<true, NT>(L61, L31, ps, nb, tw);
  }
  auto s61 = [&](int k, int) { return pick_slot<CS.sq_slots>(tw, CS.step_slots, k); };
  auto s31 = [&](int k, int) {
    const e61 w = pick_slot<CS.sq_slots>(tw, CS.step_slots, k);
    return e31{(uint32_t)w.a, (uint32_t)(w.a >> 32)};
  };
  if (g == 0) square_self<true, false, CS.sq_slots>(L61, L31, logC, 2, s61, s31);
  else square_cross<true, CS.sq_slots>(L61, L31, logC, s61, s31);
  if constexpr (CC != 0 && WIDE) {
    reg_step<false, 4, NT>(L61, L31, CC, 2, 8, tw);
    lds_step<false, 4, true>(L61, L31, CC, 0, 2, 6, T61 + 126, T31 + 126);
    lds_step<false, 4, true>(L61, L31, CC, 0, 2, 4, T61 + 30, T31 + 30);
    lds_step<false, 4, true>(L61, L31, CC, 0, 2, 2, T61 + 6, T31 + 6);
    edge_step<NT, false, 4, false, STREAM>(L61, L31, CC, 0, 2, 0, z61, z31, gidx, r61, r31, tw0);
    return;
  }
  if constexpr (CC != 0 && !WIDE) {
    reg_step<false, 8, NT>(L61, L31, CC, 2, 7, tw);
    lds_step<false, 8, true>(L61, L31, CC, 0, 2, 4, T61 + 30, T31 + 30);
    lds_step<false, 8, true>(L61, L31, CC, 0, 2, 1, T61 + 2, T31 + 2);
    edge_step<NT, false, 2, false>(L61, L31, CC, 0, 2, 0, z61, z31, gidx, r61, r31, tw0);
    return;
  }
  centre_last_step<false, NT>(L61, L31, ps, nb, tw);
  lds_pass<false, true>(L61, L31, 0, nb, ps, T61, T31, 1, edges ? ns - 1 : ns);
  if (edges) {
    edge_pass_step<NT, false, false>(L61, L31, 0, nb, ps, 0, z61, z31, gidx, r61, r31, tw0);
    return;
  }
  for (int i = threadIdx.x; i < (nb << logC); i += blockDim.x) {
    gst(z61 + gidx(i), L61[PADX(i)]);
    gst(z31 + gidx(i), L31[PADX(i)]);
  }
}

// ---------------------------------------------------------------------------
// Host-side launchers, used by the engine.

// P.L when P's single-workgroup transform is the default plan of its length
// (small_steps), which has a compiled-in instance k_small<·, ·, L> for every
// length of the plan (L = 2..12), else 0.
int small_const(const DevPlan& P) {
  const Pass& o = P.outer;
  const int L = P.L;
  if (L < 2 || L > SMALL_LOG_H || o.ell0 != 0 || o.A != L || o.nsteps != small_nsteps(L)) return 0;
  int ell = 0;
  for (int k = 0; k < o.nsteps; ++k) {
    const int r = k == 0 ? small_first_radix(L) : small_maxr(L);
    if (o.steps[k].radix != r || o.steps[k].ell != ell) return 0;
    ell += r == 8 ? 3 : (r == 4 ? 2 : 1);
  }
  return L;
}

// k_small instance for P: wave-sized up to 2^8 points, 512 threads above;
// the default plans with their steps compiled in
template <int I = 2, class F>
hipError_t small_dispatch(const DevPlan& P, F&& go, int cl = -1) {
  constexpr int WL = SMALL_WAVE_LOG_H, WT = SMALL_WAVE_THREADS, LL = SMALL_LDS_LOG_H, T = SMALL_THREADS;
  static_assert(SMALL_LOG_H <= LL + 1, "one kernel for the lengths past the all-LDS ones");
  if (cl < 0) cl = small_const(P);
  if constexpr (I <= SMALL_LOG_H) {
    if (cl == I) {
      if constexpr (I <= WL) return go(k_small<WL, WT, I>, WT);
      else if constexpr (I <= LL) return go(k_small<LL, T, I>, T);
      else return go(k_small<SMALL_LOG_H, T, I>, T);
    }
    return small_dispatch<I + 1>(P, go, cl);
  } else {
    if (P.L <= WL) return go(k_small<WL, WT>, WT);
    if (P.L <= LL) return go(k_small<LL, T>, T);
    return go(k_small<SMALL_LOG_H, T>, T);
  }
}

hipError_t launch_small(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st) {
  return small_dispatch(P, [&](auto kernel, int nth) {
    hipLaunchKernelGGL(kernel, dim3(ntests), dim3(nth), 0, st, tests, P.tw61, P.tw31, P.sq61, P.sq31, P.outer);
    return hipGetLastError();
  });
}

// Polls of a neighbour's launch counter before the fused pass gives up
// (a poll is a device-scope load plus a short sleep, about a microsecond:
// about a second in all). Settable for tests (m2_debug_fused_spin_limit).
uint32_t g_fused_spin_limit = 1u << 20;

namespace {
// Whether the centre of P has the step geometry that k_centre<NT, 10> has
// compiled in: 2^10-element blocks in five radix-4 steps whose first step
// loads directly (1024 threads), or in radix-2, 8, 8, 8 steps after a gather
// (512 threads).
template <int NT>
bool centre_const(const DevPlan& P, int direct_in) {
  constexpr bool wide = NT == MP_THREADS_WIDE;
  static const int radix_w[] = {4, 4, 4, 4, 4}, lam_w[] = {0, 2, 4, 6, 8};
  static const int radix_n[] = {2, 8, 8, 8}, lam_n[] = {0, 1, 4, 7};
  const int ns = wide ? 5 : 4;
  if (P.centre.A != 10 || P.centre.nsteps != ns || (direct_in != 0) != wide) return false;
  for (int i = 0; i < ns; ++i)
    if (P.centre.steps[i].radix != (wide ? radix_w : radix_n)[i] ||
        P.centre.steps[i].ell - P.centre.ell0 != (wide ? lam_w : lam_n)[i])
      return false;
  return true;
}

// Outer-pass shapes compiled into k_outer_fwd/k_outer_inv<A, logW, logS>: the
// default (radix-8) plans' outer passes of n = 2^16 .. 2^24. The steps are
// those make_pass gives A levels: radix 4, 8 (A = 5), 8, 8 (6), 2, 8, 8 (7),
// 4, 8, 8 (8), 8, 8, 8 (9), the first at level 0.
struct OuterShape {
  int A, logW, logS;
};
constexpr OuterShape OUTER_SHAPES[] = {
    {5, 6, 10}, {6, 5, 10}, {7, 4, 10}, {8, 3, 10}, {9, 2, 10},  // n = 2^16 .. 2^20
    {5, 6, 15}, {6, 5, 15}, {6, 5, 16}, {7, 4, 16},              // n = 2^21 .. 2^24
};
constexpr int N_OUTER_SHAPES = sizeof(OUTER_SHAPES) / sizeof(OUTER_SHAPES[0]);

// Index + 1 of P's outer-pass shape in OUTER_SHAPES, or 0.
int outer_const(const DevPlan& P) {
  const Pass& o = P.outer;
  static const int first_radix[10] = {0, 0, 0, 0, 0, 4, 8, 2, 4, 8};
  if (o.ell0 != 0 || o.A < 5 || o.A > 9 || o.nsteps != (o.A + 2) / 3 || o.steps[0].ell != 0 ||
      o.steps[0].radix != first_radix[o.A])
    return 0;
  for (int k = 1; k < o.nsteps; ++k)
    if (o.steps[k].radix != 8 || o.steps[k].ell != o.steps[k - 1].ell + (k == 1 ? (o.A - 3 * (o.nsteps - 1)) : 3))
      return 0;
  for (int i = 0; i < N_OUTER_SHAPES; ++i)
    if (OUTER_SHAPES[i].A == o.A && OUTER_SHAPES[i].logW == P.logW && OUTER_SHAPES[i].logS == P.logS) return i + 1;
  return 0;
}

// The kernel instances for shape index oc (outer_const) or the generic ones.
template <bool FWD, int I = 0>
auto outer_kernel(int oc) {
  if constexpr (I == N_OUTER_SHAPES) {
    (void)oc;
    return FWD ? k_outer_fwd<> : k_outer_inv<>;
  } else {
    constexpr OuterShape sh = OUTER_SHAPES[I];
    if (oc == I + 1) return FWD ? k_outer_fwd<sh.A, sh.logW, sh.logS> : k_outer_inv<sh.A, sh.logW, sh.logS>;
    return outer_kernel<FWD, I + 1>(oc);
  }
}

// Launches the kernels of one iteration of every test in `tests` through
// launch(name, kernel, grid, block, args...); grid.y indexes the tests. NT is
// the centre's workgroup size, P.centre_nt.
template <int NT, class F>
hipError_t iteration_nt(const DevPlan& P, const TestDesc* tests, uint32_t nt, F&& launch) {
  hipError_t e;
  const uint32_t ngroups_outer = (1u << P.logS) >> P.logW;
  const dim3 blk(MP_THREADS), cblk(NT);
  // the outer passes of the default plans of n = 2^16 .. 2^24 have kernels
  // with their geometry built in (outer_const)
  const int oc = outer_const(P);
  auto outer_fwd = outer_kernel<true>(oc);
  auto outer_inv = outer_kernel<false>(oc);
  if ((e = launch("k_outer_fwd", outer_fwd, dim3(ngroups_outer, nt), blk, tests, P.tw61, P.tw31, P.outer, P.logS,
                  P.logW)) != hipSuccess)
    return e;
  const uint32_t nmid = P.has_mid ? (1u << P.outer.A) * ((1u << P.mid_logS) >> P.mid_logW) : 0;
  // the middle passes of n = 2^23 and 2^24 (6 levels in two radix-8 steps,
  // W = 32, S = 1024) and of n = 2^21 and 2^22 (5 levels in radix-4, 8
  // steps, W = 64, S = 1024) have kernels with that geometry built in
  const Pass& mp = P.mid;
  const bool mtwo = P.has_mid && mp.nsteps == 2 && mp.steps[0].ell == mp.ell0 && mp.steps[1].radix == 8 &&
                    P.mid_logS == 10;
  const bool m6 = mtwo && mp.A == 6 && mp.steps[0].radix == 8 && mp.steps[1].ell == mp.ell0 + 3 && P.mid_logW == 5;
  const bool m5 = mtwo && mp.A == 5 && mp.steps[0].radix == 4 && mp.steps[1].ell == mp.ell0 + 2 && P.mid_logW == 6;
  auto mid_fwd = m6 ? k_mid<true, 6, 5, 10> : m5 ? k_mid<true, 5, 6, 10> : k_mid<true>;
  auto mid_inv = m6 ? k_mid<false, 6, 5, 10> : m5 ? k_mid<false, 5, 6, 10> : k_mid<false>;
  if (P.has_mid && (e = launch("k_mid_fwd", mid_fwd, dim3(nmid, nt), blk, tests, P.tw61, P.tw31, P.mid, P.mid_logS,
                               P.mid_logW)) != hipSuccess)
    return e;
  const uint32_t ncentre = (P.h >> P.centre.A) / 2;
  const int direct_in = ncentre * nt <= P.cus ? 1 : 0;
  if ((e = launch("k_centre", centre_const<NT>(P, direct_in) ? k_centre<NT, 10> : k_centre<NT, 0>, dim3(ncentre, nt),
                  cblk, tests, P.ctw61, P.ctw31, P.crec, P.centre, direct_in)) != hipSuccess)
    return e;
  if (P.has_mid && (e = launch("k_mid_inv", mid_inv, dim3(nmid, nt), blk, tests, P.tw61, P.tw31, P.mid, P.mid_logS,
                               P.mid_logW)) != hipSuccess)
    return e;
  return launch("k_outer_inv", outer_inv, dim3(ngroups_outer, nt), blk, tests, P.tw61, P.tw31, P.outer, P.logS,
                P.logW);
}
template <class F>
hipError_t iteration(const DevPlan& P, const TestDesc* tests, uint32_t nt, F&& launch) {
  if (P.centre_nt == MP_THREADS_WIDE) return iteration_nt<MP_THREADS_WIDE>(P, tests, nt, launch);
  return iteration_nt<MP_THREADS_NARROW>(P, tests, nt, launch);
}


// k >= 1 iterations of a single test (no middle pass) with the outer passes of
// consecutive iterations fused: the run starts and ends with the digits.
template <int NT, class F>
hipError_t run_fused_nt(const DevPlan& P, const TestDesc* tests, uint64_t k, F&& launch) {
  hipError_t e;
  // the outer passes of n = 2^20 (levels 0..8 in three radix-8 steps, W = 4,
  // S = 1024) have kernels with that geometry built in
  const int oc = outer_const(P);
  const bool c20 = oc > 0 && OUTER_SHAPES[oc - 1].A == 9 && OUTER_SHAPES[oc - 1].logW == 2 &&
                   OUTER_SHAPES[oc - 1].logS == 10;
  const uint32_t ns = outer_strips(P);
  const uint32_t ncentre = (P.h >> P.centre.A) / 2;
  const int direct_in = ncentre <= P.cus ? 1 : 0;
  // and so has its 1024-thread centre (2^10-element blocks, five radix-4
  // steps, one workgroup per CU so that the first step loads directly)
  auto centre = centre_const<NT>(P, direct_in) ? k_centre<NT, 10> : k_centre<NT, 0>;
  const dim3 blk(MP_THREADS), cblk(NT);
  if ((e = launch("k_outer_fwd", c20 ? k_outer_fwd<9, 2, 10> : k_outer_fwd<>, dim3(ns, 1), blk, tests, P.tw61,
                  P.tw31, P.outer, P.logS, P.logW)) != hipSuccess)
    return e;
  for (uint64_t i = 0; i + 1 < k; ++i) {
    if ((e = launch("k_centre", centre, dim3(ncentre, 1), cblk, tests, P.ctw61, P.ctw31, P.crec, P.centre,
                    direct_in)) != hipSuccess)
      return e;
    if ((e = launch("k_outer_fused", c20 ? k_outer_fused<9, 2, 10> : k_outer_fused<0, 0, 0>, dim3(ns, 1), blk, tests,
                    P.tw61, P.tw31, P.outer, P.logS, P.logW, g_fused_spin_limit)) != hipSuccess)
      return e;
  }
  if ((e = launch("k_centre", centre, dim3(ncentre, 1), cblk, tests, P.ctw61, P.ctw31, P.crec, P.centre,
                  direct_in)) != hipSuccess)
    return e;
  return launch("k_outer_inv", c20 ? k_outer_inv<9, 2, 10> : k_outer_inv<>, dim3(ns, 1), blk, tests, P.tw61, P.tw31,
                P.outer, P.logS, P.logW);
}
template <class F>
hipError_t run_fused(const DevPlan& P, const TestDesc* tests, uint64_t k, F&& launch) {
  if (P.centre_nt == MP_THREADS_WIDE) return run_fused_nt<MP_THREADS_WIDE>(P, tests, k, launch);
  return run_fused_nt<MP_THREADS_NARROW>(P, tests, k, launch);
}
}  // namespace

hipError_t launch_run_fused(const DevPlan& P, const TestDesc* tests, uint64_t k, hipStream_t st) {
  if (P.small || P.has_mid || k == 0) return hipErrorInvalidValue;
  return run_fused(P, tests, k, [&](const char*, auto kernel, dim3 grid, dim3 block, auto... args) -> hipError_t {
    hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    return hipGetLastError();
  });
}

hipError_t launch_run_fused_events(const DevPlan& P, const TestDesc* tests, uint64_t k, hipStream_t st,
                                   hipEvent_t* starts, hipEvent_t* stops, int max_events, int* count,
                                   const char** names) {
  if (P.small || P.has_mid || k == 0) return hipErrorInvalidValue;
  int n = 0;
  auto launch = [&](const char* name, auto kernel, dim3 grid, dim3 block, auto... args) -> hipError_t {
    if (n >= max_events) return hipErrorInvalidValue;
    names[n] = name;
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, starts[n], stops[n], 0, args...);
    ++n;
    return hipGetLastError();
  };
  hipError_t e = run_fused(P, tests, k, launch);
  *count = n;
  return e;
}

hipError_t launch_iteration(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st) {
  return iteration(P, tests, ntests, [&](const char*, auto kernel, dim3 grid, dim3 block, auto... args) -> hipError_t {
    hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    return hipGetLastError();
  });
}

hipError_t launch_iteration_events(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st,
                                   hipEvent_t* starts, hipEvent_t* stops, int max_events, int* count,
                                   const char** names) {
  int k = 0;
  auto launch = [&](const char* name, auto kernel, dim3 grid, dim3 block, auto... args) -> hipError_t {
    if (k >= max_events) return hipErrorInvalidValue;
    names[k] = name;
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, starts[k], stops[k], 0, args...);
    ++k;
    return hipGetLastError();
  };
  hipError_t e = iteration(P, tests, ntests, launch);
  *count = k;
  return e;
}

hipError_t launch_small_events(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st,
                               hipEvent_t start, hipEvent_t stop) {
  return small_dispatch(P, [&](auto kernel, int nth) {
    hipExtLaunchKernelGGL(kernel, dim3(ntests), dim3(nth), 0, st, start, stop, 0, tests, P.tw61, P.tw31, P.sq61,
                          P.sq31, P.outer);
    return hipGetLastError();
  });
}

// Issue-rate probe for v_mad_u64_u32, the 32x32->64 multiply-add every field
// product is built from: each thread runs eight independent chains
// x <- lo32(x) * k + x, one instruction per step.
constexpr int MAD_CHAINS = 8;

__global__ void __launch_bounds__(256) k_mad_rate(uint64_t* out, uint32_t iters, uint32_t k) {
  uint64_t x[MAD_CHAINS];
#pragma unroll
  for (int c = 0; c < MAD_CHAINS; ++c) x[c] = threadIdx.x * 131u + c * 7u + blockIdx.x;
  for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < MAD_CHAINS; ++c) x[c] = (uint64_t)(uint32_t)x[c] * k + x[c];
  }
  uint64_t s = 0;
#pragma unroll
  for (int c = 0; c < MAD_CHAINS; ++c) s ^= x[c];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

hipError_t launch_mad_rate(uint64_t* out, uint32_t blocks, uint32_t iters, hipStream_t st, uint64_t* lane_ops) {
  hipLaunchKernelGGL(k_mad_rate, dim3(blocks), dim3(256), 0, st, out, iters, 0x9E3779B9u);
  *lane_ops = (uint64_t)blocks * 256u * iters * MAD_CHAINS;
  return hipGetLastError();
}

}  // namespace m2
