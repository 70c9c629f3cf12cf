// gfx950 HIP kernels for one IBDWT squaring of x mod 2^p - 1 over
// GF(M61^2) x GF(M31^2).
//
// Data between iterations: z61[h] holds the digits as signed 64-bit pairs
// (digit 2k in .a, digit 2k+1 in .b) and G holds the carries not yet applied
// (one per carry segment row of the outer pass). During an iteration z61/z31
// hold the weighted spectrum.
//
// Carry rule (exact, no inter-workgroup waits): the digits are cut into
// segments; every segment is normalised with carry-in 0, then receives the
// carry-out g of the previous segment (cyclically: 2^p = 1 mod M_p) and
// ripples it upward, but its top digit keeps whatever is left instead of
// passing it on. The represented value is unchanged, every digit but a
// segment's top one is < 2^width, and the top one is <= 2^width as long as
// the bits below it exceed the bits of g (8 digits of >= 12 bits each).
// Convolution terms then stay below n * 2^(2q+2) <= 2^91 < M61*M31.
#include "m2_arith.h"
#include "m2_internal.h"

namespace m2 {

typedef __int128 i128;

// M61*M31, and the point above which a CRT value is read as negative (LL's -2
// can make a digit -1, and convolution terms slightly negative).
#define M2_CRT_M (((i128)P61) * (i128)P31)
#define M2_CRT_NEG (((i128)3) << 90)

__device__ __forceinline__ int64_t to_i64(uint64_t x) { return (int64_t)x; }

// Signed digit -> canonical field residues.
__device__ __forceinline__ uint64_t digit61(int64_t d) { return d >= 0 ? (uint64_t)d : P61 - (uint64_t)(-d); }
__device__ __forceinline__ uint32_t mod31_u64(uint64_t x) {
  uint64_t r = (x & P31) + (x >> 31);
  r = (r & P31) + (r >> 31);
  return (uint32_t)(r >= P31 ? r - P31 : r);
}
__device__ __forceinline__ uint32_t digit31(int64_t d) {
  return d >= 0 ? mod31_u64((uint64_t)d) : neg31(mod31_u64((uint64_t)(-d)));
}

// Unweight (x 2^-(w + ln + 1), which also removes 1/h and the centre's 4) and
// Garner CRT: x = t + M31 * ((s - t) * (2^31 + 1) mod M61), then signed.
__device__ __forceinline__ i128 crt_digit(uint64_t s, uint32_t t, uint32_t w61, uint32_t w31, uint32_t k1) {
  uint32_t a61 = w61 + k1;
  a61 -= (a61 >= 61u) ? 61u : 0u;
  a61 -= (a61 >= 61u) ? 61u : 0u;
  uint32_t a31 = w31 + k1;
  a31 -= (a31 >= 31u) ? 31u : 0u;
  a31 -= (a31 >= 31u) ? 31u : 0u;
  a31 -= (a31 >= 31u) ? 31u : 0u;
  s = a61 ? rot61(s, 61u - a61) : s;
  t = a31 ? rot31(t, 31u - a31) : t;
  uint64_t u = sub61(s, t);
  u = add61(u, rot61(u, 31));
  i128 x = (((i128)u) << 31) - (i128)u + (i128)t;
  if (x >= M2_CRT_NEG) x -= M2_CRT_M;
  return x;
}

__device__ __forceinline__ int64_t norm_digit(i128 acc, uint32_t w, i128& carry) {
  const uint64_t mask = w ? ((~0ull) >> (64u - w)) : 0ull;
  carry = acc >> w;
  return (int64_t)((uint64_t)acc & mask);
}

// Add carry c at the first of the 8 digits d[0..7]; the top digit absorbs the rest.
__device__ __forceinline__ void ripple8(int64_t* d, const uint32_t* w, i128 c) {
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (c != 0) {
      i128 acc = (i128)d[i] + c;
      d[i] = norm_digit(acc, w[i], c);
    }
  }
  d[7] += (int64_t)c;
}

// ---------------------------------------------------------------------------
// One radix step over data in LDS. The LDS holds nb blocks of R = 2^logR rows
// by W = 2^logW columns (row-major); every column of a block is an
// independent transform whose block index at the pass's first level is j0[b].
// Step `lam` (relative level) with radix r: butterflies combine rows
// blk*(R>>lam) + off + e*d, e < r, d = R >> (lam + log2 r); the twist tau^e
// comes from the step's table at global block index (j0 << lam) + blk.
template <bool FWD, int RADIX>
__device__ __forceinline__ void lds_step(e61* L61, e31* L31, int logR, int logW, int nb, const uint32_t* j0, int lam,
                                         uint32_t twoff, const e61* __restrict__ tw61, const e31* __restrict__ tw31) {
  constexpr int LR = RADIX == 8 ? 3 : (RADIX == 4 ? 2 : 1);
  const int logd = logR - lam - LR;
  const int logU = logR - LR + logW;  // butterflies per block (log2)
  const int total = nb << logU;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int b = t >> logU;
    const int u = t & ((1 << logU) - 1);
    const int col = u & ((1 << logW) - 1);
    const int g = u >> logW;
    const int off = g & ((1 << logd) - 1);
    const int blk = g >> logd;
    const int row0 = (blk << (logR - lam)) + off;
    const uint32_t j = (j0[b] << lam) + (uint32_t)blk;
    const int base = (b << (logR + logW)) + (row0 << logW) + col;
    const int stride = 1 << (logd + logW);
    const e61* T61 = tw61 + twoff + j * (RADIX - 1);
    const e31* T31 = tw31 + twoff + j * (RADIX - 1);
    e61 v[RADIX];
    e31 x[RADIX];
#pragma unroll
    for (int e = 0; e < RADIX; ++e) {
      v[e] = L61[base + e * stride];
      x[e] = L31[base + e * stride];
    }
    if (FWD) {
#pragma unroll
      for (int e = 1; e < RADIX; ++e) {
        v[e] = mul(v[e], T61[e - 1]);
        x[e] = mul(x[e], T31[e - 1]);
      }
      if (RADIX == 8) { dft8_fwd(v); dft8_fwd(x); }
      if (RADIX == 4) { dft4_fwd(v); dft4_fwd(x); }
      if (RADIX == 2) { dft2(v[0], v[1]); dft2(x[0], x[1]); }
    } else {
      if (RADIX == 8) { dft8_inv(v); dft8_inv(x); }
      if (RADIX == 4) { dft4_inv(v); dft4_inv(x); }
      if (RADIX == 2) { dft2(v[0], v[1]); dft2(x[0], x[1]); }
#pragma unroll
      for (int e = 1; e < RADIX; ++e) {
        v[e] = mulconj(v[e], T61[e - 1]);
        x[e] = mulconj(x[e], T31[e - 1]);
      }
    }
#pragma unroll
    for (int e = 0; e < RADIX; ++e) {
      L61[base + e * stride] = v[e];
      L31[base + e * stride] = x[e];
    }
  }
  __syncthreads();
}

template <bool FWD>
__device__ __forceinline__ void lds_pass(e61* L61, e31* L31, int logW, int nb, const uint32_t* j0, const Pass& ps,
                                         const e61* __restrict__ tw61, const e31* __restrict__ tw31) {
  for (int si = 0; si < ps.nsteps; ++si) {
    const Step st = FWD ? ps.steps[si] : ps.steps[ps.nsteps - 1 - si];
    const int lam = st.ell - ps.ell0;
    if (st.radix == 8) lds_step<FWD, 8>(L61, L31, ps.A, logW, nb, j0, lam, st.twoff, tw61, tw31);
    else if (st.radix == 4) lds_step<FWD, 4>(L61, L31, ps.A, logW, nb, j0, lam, st.twoff, tw61, tw31);
    else lds_step<FWD, 2>(L61, L31, ps.A, logW, nb, j0, lam, st.twoff, tw61, tw31);
  }
}

// Real-data pairing and squaring of one pair of spectrum values (the
// reference's square_half, mersenne2_gpu.cpp:856-870): with
// u0 = z_k + conj(z_mk), u1 = z_k - conj(z_mk), v0 = u0^2 - u1^2 tw,
// v1 = 2 u0 u1:  z_k <- v0 + v1,  z_mk <- conj(v0 - v1).
template <class E>
__device__ __forceinline__ void square_pair(E& zk, E& zmk, E tw) {
  const E c = conj(zmk);
  const E u0 = add(zk, c), u1 = sub(zk, c);
  const E v0 = sub(sqr(u0), mul(sqr(u1), tw));
  const E v1 = mul(u0, add(u1, u1));
  zk = add(v0, v1);
  zmk = conj(sub(v0, v1));
}

// Squaring inside one LDS block that pairs with itself: block 0 (positions
// 0..Cb-1, pairs via the mirror formula, k = 0 special) or block 1
// (positions Cb..2Cb-1, k pairs with 3Cb-1-k).
__device__ __forceinline__ void square_self(e61* L61, e31* L31, int logCb, uint32_t B, const e61* __restrict__ sq61,
                                            const e31* __restrict__ sq31) {
  const int Cb = 1 << logCb;
  for (int j = threadIdx.x; j < (Cb >> 1); j += blockDim.x) {
    const int k = 2 * j;
    int kidx, mkidx;
    uint32_t kg;
    if (B == 0) {
      kidx = k;
      mkidx = k ? (3 << (31 - __clz(k))) - k - 1 : 0;
      kg = (uint32_t)k;
    } else {
      kidx = k;
      mkidx = Cb - 1 - k;
      kg = (uint32_t)Cb + (uint32_t)k;
    }
    const e61 t61 = sq61[kg >> 1];
    const e31 t31 = sq31[kg >> 1];
    if (B == 0 && k == 0) {
      e61 a = L61[0], a2 = a;
      e31 b = L31[0], b2 = b;
      square_pair(a, a2, t61);
      square_pair(b, b2, t31);
      const e61 z1 = L61[1];
      const e31 y1 = L31[1];
      L61[0] = a;
      L31[0] = b;
      L61[1] = sqr(add(z1, z1));
      L31[1] = sqr(add(y1, y1));
    } else {
      e61 a = L61[kidx], c = L61[mkidx];
      e31 b = L31[kidx], d = L31[mkidx];
      square_pair(a, c, t61);
      square_pair(b, d, t31);
      L61[kidx] = a;
      L61[mkidx] = c;
      L31[kidx] = b;
      L31[mkidx] = d;
    }
  }
  __syncthreads();
}

// Squaring across two LDS blocks B (at 0) and B2 = 3*2^a - 1 - B (at Cb):
// position B*Cb + o pairs with B2*Cb + (Cb - 1 - o).
__device__ __forceinline__ void square_cross(e61* L61, e31* L31, int logCb, uint32_t B, uint32_t B2,
                                             const e61* __restrict__ sq61, const e31* __restrict__ sq31) {
  const int Cb = 1 << logCb;
  for (int o = threadIdx.x; o < Cb; o += blockDim.x) {
    int kidx, mkidx;
    uint32_t kg;
    if ((o & 1) == 0) {
      kidx = o;
      mkidx = Cb + (Cb - 1 - o);
      kg = (B << logCb) + (uint32_t)o;
    } else {
      kidx = Cb + (Cb - 1 - o);
      mkidx = o;
      kg = (B2 << logCb) + (uint32_t)(Cb - 1 - o);
    }
    e61 a = L61[kidx], c = L61[mkidx];
    e31 b = L31[kidx], d = L31[mkidx];
    square_pair(a, c, sq61[kg >> 1]);
    square_pair(b, d, sq31[kg >> 1]);
    L61[kidx] = a;
    L61[mkidx] = c;
    L31[kidx] = b;
    L31[mkidx] = d;
  }
  __syncthreads();
}

// Load 8 digits (4 elements) of a segment from LDS (int64 pairs in e61 slots).
__device__ __forceinline__ void seg_load(const e61* L61, int idx0, int64_t* d) {
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) {
    const e61 v = L61[idx0 + i];
    d[2 * i] = to_i64(v.a);
    d[2 * i + 1] = to_i64(v.b);
  }
}
__device__ __forceinline__ void seg_store(e61* L61, int idx0, const int64_t* d) {
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) L61[idx0 + i] = {(uint64_t)d[2 * i], (uint64_t)d[2 * i + 1]};
}

// Unweight + CRT + zero-carry-in normalisation of one segment (elements at
// LDS idx0.. with digits starting at global digit j0); returns its carry-out.
__device__ __forceinline__ i128 seg_normalise(e61* L61, const e31* L31, int idx0, uint32_t j0,
                                              const DigitInfo* __restrict__ dig, uint32_t k1, i128 carry) {
  int64_t d[2 * SEG_ELEMS];
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) {
    const e61 v = L61[idx0 + i];
    const e31 y = L31[idx0 + i];
    const DigitInfo da = dig[j0 + 2 * i], db = dig[j0 + 2 * i + 1];
    i128 xa = crt_digit(v.a, y.a, da.w61, da.w31, k1);
    d[2 * i] = norm_digit(xa + carry, da.width, carry);
    i128 xb = crt_digit(v.b, y.b, db.w61, db.w31, k1);
    d[2 * i + 1] = norm_digit(xb + carry, db.width, carry);
  }
  seg_store(L61, idx0, d);
  return carry;
}

__device__ __forceinline__ void seg_ripple(e61* L61, int idx0, uint32_t j0, const DigitInfo* __restrict__ dig, i128 c) {
  if (c == 0) return;
  int64_t d[2 * SEG_ELEMS];
  uint32_t w[2 * SEG_ELEMS];
  seg_load(L61, idx0, d);
#pragma unroll
  for (int i = 0; i < 2 * SEG_ELEMS; ++i) w[i] = dig[j0 + i].width;
  ripple8(d, w, c);
  seg_store(L61, idx0, d);
}

// Digits of a segment -> weighted field elements in place (L61 digits -> L61/L31).
__device__ __forceinline__ void seg_weight(e61* L61, e31* L31, int idx0, uint32_t j0,
                                           const DigitInfo* __restrict__ dig) {
#pragma unroll
  for (int i = 0; i < SEG_ELEMS; ++i) {
    const e61 v = L61[idx0 + i];
    const int64_t da = to_i64(v.a), db = to_i64(v.b);
    const DigitInfo ia = dig[j0 + 2 * i], ib = dig[j0 + 2 * i + 1];
    L61[idx0 + i] = {rot61(digit61(da), ia.w61), rot61(digit61(db), ib.w61)};
    L31[idx0 + i] = {rot31(digit31(da), ia.w31), rot31(digit31(db), ib.w31)};
  }
}

__device__ __forceinline__ i128 load_carry(const Carry128* G, uint32_t i) {
  const Carry128 c = G[i];
  return (((i128)c.hi) << 64) | (i128)c.lo;
}
__device__ __forceinline__ void store_carry(Carry128* G, uint32_t i, i128 v) {
  G[i].lo = (uint64_t)v;
  G[i].hi = (int64_t)(v >> 64);
}

// ---------------------------------------------------------------------------
// Small plan: one workgroup holds the whole residue (h <= 2^SMALL_LOG_H) and
// runs `iters` iterations without leaving LDS.
__global__ void __launch_bounds__(256) k_small(e61* __restrict__ z61, const DigitInfo* __restrict__ dig,
                                               const e61* __restrict__ tw61, const e31* __restrict__ tw31,
                                               const e61* __restrict__ sq61, const e31* __restrict__ sq31, Pass ps,
                                               int ln, int mode, uint32_t iters) {
  __shared__ e61 L61[1 << SMALL_LOG_H];
  __shared__ e31 L31[1 << SMALL_LOG_H];
  __shared__ Carry128 gs[(1 << SMALL_LOG_H) / SEG_ELEMS];
  const int logH = ln - 1;
  const int h = 1 << logH;
  const int nseg = h / SEG_ELEMS;
  const uint32_t k1 = (uint32_t)ln + 1u;
  const uint32_t j0[1] = {0};
  for (int i = threadIdx.x; i < h; i += blockDim.x) L61[i] = z61[i];
  __syncthreads();
  for (uint32_t it = 0; it < iters; ++it) {
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) seg_weight(L61, L31, s * SEG_ELEMS, 2u * s * SEG_ELEMS, dig);
    __syncthreads();
    lds_pass<true>(L61, L31, 0, 1, j0, ps, tw61, tw31);
    square_self(L61, L31, logH, 0, sq61, sq31);
    lds_pass<false>(L61, L31, 0, 1, j0, ps, tw61, tw31);
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) {
      const i128 c0 = (mode == 0 && s == 0) ? (i128)(-2) : (i128)0;
      store_carry(gs, s, seg_normalise(L61, L31, s * SEG_ELEMS, 2u * s * SEG_ELEMS, dig, k1, c0));
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nseg; s += blockDim.x)
      seg_ripple(L61, s * SEG_ELEMS, 2u * s * SEG_ELEMS, dig, load_carry(gs, (s + nseg - 1) % nseg));
    __syncthreads();
  }
  for (int i = threadIdx.x; i < h; i += blockDim.x) z61[i] = L61[i];
}

// ---------------------------------------------------------------------------
// Large plan, outer pass: levels [0, A) over columns of stride S = 2^logS.
// A workgroup owns W = 2^logW adjacent columns (R x W elements); its rows are
// carry segments of W elements (W/4 sub-segments chained inside the
// workgroup); the carry of row r leaves through G[r * NS + cg] and enters the
// next row segment in digit order in the next forward pass.
__global__ void __launch_bounds__(256) k_outer_fwd(e61* __restrict__ z61, e31* __restrict__ z31,
                                                   const DigitInfo* __restrict__ dig, const e61* __restrict__ tw61,
                                                   const e31* __restrict__ tw31, const Carry128* __restrict__ G,
                                                   Pass ps, int logS, int logW) {
  __shared__ e61 L61[PASS_MAX_ELEMS];
  __shared__ e31 L31[PASS_MAX_ELEMS];
  const int R = 1 << ps.A, W = 1 << logW;
  const uint32_t cg = blockIdx.x, NS = (1u << logS) >> logW, NSEG = (uint32_t)R * NS;
  const uint32_t col0 = cg << logW;
  const uint32_t j0[1] = {0};
  // prologue: digits -> (carry-in) -> weighted elements
  const int subs = W / SEG_ELEMS;
  for (int it = threadIdx.x; it < R * subs; it += blockDim.x) {
    const int r = it / subs, v = it - r * subs;
    const int idx0 = r * W + v * SEG_ELEMS;
    const uint32_t e0 = ((uint32_t)r << logS) + col0 + (uint32_t)(v * SEG_ELEMS);
#pragma unroll
    for (int i = 0; i < SEG_ELEMS; ++i) L61[idx0 + i] = z61[e0 + i];
    if (v == 0) {
      const uint32_t s = (uint32_t)r * NS + cg;
      seg_ripple(L61, idx0, 2u * e0, dig, load_carry(G, (s + NSEG - 1) % NSEG));
    }
    seg_weight(L61, L31, idx0, 2u * e0, dig);
  }
  __syncthreads();
  lds_pass<true>(L61, L31, logW, 1, j0, ps, tw61, tw31);
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    const uint32_t e = ((uint32_t)(i >> logW) << logS) + col0 + (uint32_t)(i & (W - 1));
    z61[e] = L61[i];
    z31[e] = L31[i];
  }
}

__global__ void __launch_bounds__(256) k_outer_inv(e61* __restrict__ z61, const e31* __restrict__ z31,
                                                   const DigitInfo* __restrict__ dig, const e61* __restrict__ tw61,
                                                   const e31* __restrict__ tw31, Carry128* __restrict__ G, Pass ps,
                                                   int logS, int logW, int ln, int mode) {
  __shared__ e61 L61[PASS_MAX_ELEMS];
  __shared__ e31 L31[PASS_MAX_ELEMS];
  __shared__ Carry128 gs[PASS_MAX_ELEMS / SEG_ELEMS];
  const int R = 1 << ps.A, W = 1 << logW;
  const uint32_t cg = blockIdx.x, NS = (1u << logS) >> logW;
  const uint32_t col0 = cg << logW;
  const uint32_t k1 = (uint32_t)ln + 1u;
  const uint32_t j0[1] = {0};
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    const uint32_t e = ((uint32_t)(i >> logW) << logS) + col0 + (uint32_t)(i & (W - 1));
    L61[i] = z61[e];
    L31[i] = z31[e];
  }
  __syncthreads();
  lds_pass<false>(L61, L31, logW, 1, j0, ps, tw61, tw31);
  const int subs = W / SEG_ELEMS;
  for (int it = threadIdx.x; it < R * subs; it += blockDim.x) {
    const int r = it / subs, v = it - r * subs;
    const int idx0 = r * W + v * SEG_ELEMS;
    const uint32_t e0 = ((uint32_t)r << logS) + col0 + (uint32_t)(v * SEG_ELEMS);
    const i128 c0 = (mode == 0 && e0 == 0) ? (i128)(-2) : (i128)0;
    store_carry(gs, it, seg_normalise(L61, L31, idx0, 2u * e0, dig, k1, c0));
  }
  __syncthreads();
  for (int it = threadIdx.x; it < R * subs; it += blockDim.x) {
    const int r = it / subs, v = it - r * subs;
    const int idx0 = r * W + v * SEG_ELEMS;
    const uint32_t e0 = ((uint32_t)r << logS) + col0 + (uint32_t)(v * SEG_ELEMS);
    if (v > 0) seg_ripple(L61, idx0, 2u * e0, dig, load_carry(gs, it - 1));
    if (v == subs - 1) G[(uint32_t)r * NS + cg] = gs[it];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    const uint32_t e = ((uint32_t)(i >> logW) << logS) + col0 + (uint32_t)(i & (W - 1));
    z61[e] = L61[i];
  }
}

// Middle pass: levels [ell0, ell0 + A) with no carry work. Level-ell0 block
// j (of R * S elements) is split into S/W column groups.
template <bool FWD>
__global__ void __launch_bounds__(256) k_mid(e61* __restrict__ z61, e31* __restrict__ z31,
                                             const e61* __restrict__ tw61, const e31* __restrict__ tw31, Pass ps,
                                             int logS, int logW) {
  __shared__ e61 L61[PASS_MAX_ELEMS];
  __shared__ e31 L31[PASS_MAX_ELEMS];
  const int R = 1 << ps.A, W = 1 << logW;
  const uint32_t NS = (1u << logS) >> logW;
  const uint32_t jb = blockIdx.x / NS, cg = blockIdx.x - jb * NS;
  const uint32_t base = (jb << (ps.A + logS)) + (cg << logW);
  const uint32_t j0[1] = {jb};
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    const uint32_t e = base + ((uint32_t)(i >> logW) << logS) + (uint32_t)(i & (W - 1));
    L61[i] = z61[e];
    L31[i] = z31[e];
  }
  __syncthreads();
  lds_pass<FWD>(L61, L31, logW, 1, j0, ps, tw61, tw31);
  for (int i = threadIdx.x; i < R * W; i += blockDim.x) {
    const uint32_t e = base + ((uint32_t)(i >> logW) << logS) + (uint32_t)(i & (W - 1));
    z61[e] = L61[i];
    z31[e] = L31[i];
  }
}

// Centre pass: workgroup g squares the pair of contiguous blocks (B, B2) of
// C = 2^logC elements that the real-data pairing links (g = 0, 1: blocks 0
// and 1 pair with themselves); forward levels [ps.ell0, L), square, inverse.
__global__ void __launch_bounds__(256) k_centre(e61* __restrict__ z61, e31* __restrict__ z31,
                                                const e61* __restrict__ tw61, const e31* __restrict__ tw31,
                                                const e61* __restrict__ sq61, const e31* __restrict__ sq31, Pass ps) {
  __shared__ e61 L61[2 << CENTRE_LOG_C];
  __shared__ e31 L31[2 << CENTRE_LOG_C];
  const int logC = ps.A, C = 1 << logC;
  const uint32_t g = blockIdx.x;
  uint32_t B, B2;
  int nb;
  if (g < 2) {
    B = g;
    B2 = g;
    nb = 1;
  } else {
    const uint32_t gp = g - 1;  // >= 1
    const uint32_t a = 32u - __clz(gp);  // octave: blocks [2^a, 2^(a+1))
    B = (1u << a) + (gp - (1u << (a - 1)));
    B2 = 3u * (1u << a) - 1u - B;
    nb = 2;
  }
  const uint32_t j0[2] = {B, B2};
  for (int i = threadIdx.x; i < (nb << logC); i += blockDim.x) {
    const uint32_t e = ((i < C ? B : B2) << logC) + (uint32_t)(i & (C - 1));
    L61[i] = z61[e];
    L31[i] = z31[e];
  }
  __syncthreads();
  lds_pass<true>(L61, L31, 0, nb, j0, ps, tw61, tw31);
  if (nb == 1) square_self(L61, L31, logC, B, sq61, sq31);
  else square_cross(L61, L31, logC, B, B2, sq61, sq31);
  lds_pass<false>(L61, L31, 0, nb, j0, ps, tw61, tw31);
  for (int i = threadIdx.x; i < (nb << logC); i += blockDim.x) {
    const uint32_t e = ((i < C ? B : B2) << logC) + (uint32_t)(i & (C - 1));
    z61[e] = L61[i];
    z31[e] = L31[i];
  }
}

// ---------------------------------------------------------------------------
// Host-side launchers (one per kernel), used by the engine.

hipError_t launch_small(const DevPlan& P, uint32_t iters, hipStream_t st) {
  hipLaunchKernelGGL(k_small, dim3(1), dim3(256), 0, st, P.z61, P.dig, P.tw61, P.tw31, P.sq61, P.sq31, P.outer,
                     P.ln, P.mode, iters);
  return hipGetLastError();
}

hipError_t launch_iteration(const DevPlan& P, hipStream_t st) {
  const uint32_t ngroups_outer = (1u << P.logS) >> P.logW;
  hipLaunchKernelGGL(k_outer_fwd, dim3(ngroups_outer), dim3(256), 0, st, P.z61, P.z31, P.dig, P.tw61, P.tw31, P.G,
                     P.outer, P.logS, P.logW);
  if (P.has_mid) {
    const uint32_t nmid = (1u << P.outer.A) * ((1u << P.mid_logS) >> P.mid_logW);
    hipLaunchKernelGGL(k_mid<true>, dim3(nmid), dim3(256), 0, st, P.z61, P.z31, P.tw61, P.tw31, P.mid, P.mid_logS,
                       P.mid_logW);
  }
  const uint32_t ncentre = (P.h >> P.centre.A) / 2 + 1;
  hipLaunchKernelGGL(k_centre, dim3(ncentre), dim3(256), 0, st, P.z61, P.z31, P.tw61, P.tw31, P.sq61, P.sq31,
                     P.centre);
  if (P.has_mid) {
    const uint32_t nmid = (1u << P.outer.A) * ((1u << P.mid_logS) >> P.mid_logW);
    hipLaunchKernelGGL(k_mid<false>, dim3(nmid), dim3(256), 0, st, P.z61, P.z31, P.tw61, P.tw31, P.mid, P.mid_logS,
                       P.mid_logW);
  }
  hipLaunchKernelGGL(k_outer_inv, dim3(ngroups_outer), dim3(256), 0, st, P.z61, P.z31, P.dig, P.tw61, P.tw31, P.G,
                     P.outer, P.logS, P.logW, P.ln, P.mode);
  return hipGetLastError();
}

hipError_t launch_iteration_events(const DevPlan& P, hipStream_t st, hipEvent_t* ev, int max_events, int* count,
                                   const char** names) {
  int k = 0;
  auto mark = [&](const char* name) -> hipError_t {
    if (k >= max_events) return hipErrorInvalidValue;
    names[k] = name;
    return hipEventRecord(ev[k++], st);
  };
  hipError_t e;
  const uint32_t ngroups_outer = (1u << P.logS) >> P.logW;
  hipLaunchKernelGGL(k_outer_fwd, dim3(ngroups_outer), dim3(256), 0, st, P.z61, P.z31, P.dig, P.tw61, P.tw31, P.G,
                     P.outer, P.logS, P.logW);
  if ((e = mark("k_outer_fwd")) != hipSuccess) return e;
  const uint32_t nmid = P.has_mid ? (1u << P.outer.A) * ((1u << P.mid_logS) >> P.mid_logW) : 0;
  if (P.has_mid) {
    hipLaunchKernelGGL(k_mid<true>, dim3(nmid), dim3(256), 0, st, P.z61, P.z31, P.tw61, P.tw31, P.mid, P.mid_logS,
                       P.mid_logW);
    if ((e = mark("k_mid_fwd")) != hipSuccess) return e;
  }
  const uint32_t ncentre = (P.h >> P.centre.A) / 2 + 1;
  hipLaunchKernelGGL(k_centre, dim3(ncentre), dim3(256), 0, st, P.z61, P.z31, P.tw61, P.tw31, P.sq61, P.sq31,
                     P.centre);
  if ((e = mark("k_centre")) != hipSuccess) return e;
  if (P.has_mid) {
    hipLaunchKernelGGL(k_mid<false>, dim3(nmid), dim3(256), 0, st, P.z61, P.z31, P.tw61, P.tw31, P.mid, P.mid_logS,
                       P.mid_logW);
    if ((e = mark("k_mid_inv")) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_outer_inv, dim3(ngroups_outer), dim3(256), 0, st, P.z61, P.z31, P.dig, P.tw61, P.tw31, P.G,
                     P.outer, P.logS, P.logW, P.ln, P.mode);
  if ((e = mark("k_outer_inv")) != hipSuccess) return e;
  *count = k;
  return hipGetLastError();
}

}  // namespace m2
