// VALU instruction floor of the transform's arithmetic, for static counting
// (compiled, never launched; mersenne2gpu_amd/valu_floor.py builds it and
// counts the instructions of each kernel in the gfx950 assembly).
//
// Each probe kernel loads the operands of one primitive from global memory,
// runs the primitive exactly as the engine's kernels call it (radix-r
// butterflies with and without twiddle products, the squaring of a pair,
// the carry of one 8-digit segment) and stores the result; its skeleton
// does the same loads and stores with no arithmetic. VALU instructions of a
// probe minus those of its skeleton = the primitive's own instructions,
// with no addressing, LDS traffic or loop control of the real kernels.
#include "m2_kernels.hip"

namespace m2 {
namespace floor_probe {

__device__ __forceinline__ int tid() { return (int)(threadIdx.x + blockIdx.x * blockDim.x); }
template <int R, class E> __device__ void ld(const E* z, E* v) {
  for (int e = 0; e < R; ++e) v[e] = z[tid() * R + e];
}
template <int R, class E> __device__ void st(E* z, const E* v) {
  for (int e = 0; e < R; ++e) z[tid() * R + e] = v[e];
}

}  // namespace floor_probe

// PROBE(name, element type, points, with twiddles, body): v[R] the points, w[R-1] the twiddles
#define M2_FLOOR_PROBE(NAME, E, R, TW, ...)                                                   \
  extern "C" __global__ void __launch_bounds__(256) NAME(E* z, const E* tw) {                 \
    E v[R];                                                                                   \
    floor_probe::ld<R>(z, v);                                                                 \
    E w[R];                                                                                   \
    if (TW)                                                                                   \
      for (int e = 0; e < R - 1; ++e) w[e] = tw[floor_probe::tid() * R + e];                  \
    __VA_ARGS__;                                                                              \
    floor_probe::st<R>(z, v);                                                                 \
  }
#define M2_FLOOR_SKEL(NAME, E, R) \
  M2_FLOOR_PROBE(NAME, E, R, true, for (int e = 1; e < R; ++e) v[e].a ^= w[e - 1].a, v[e].b ^= w[e - 1].b)

M2_FLOOR_SKEL(skel8_61, e61, 8)
M2_FLOOR_SKEL(skel4_61, e61, 4)
M2_FLOOR_SKEL(skel2_61, e61, 2)
M2_FLOOR_PROBE(fwd8t_61, e61, 8, true, for (int e = 1; e < 8; ++e) v[e] = mul_tw(v[e], w[e - 1]); dft8_fwd_lazy(v))
M2_FLOOR_PROBE(inv8t_61, e61, 8, true, dft8_inv_lazy(v); for (int e = 1; e < 8; ++e) v[e] = mulconj_tw(v[e], w[e - 1]))
M2_FLOOR_PROBE(fwd8u_61, e61, 8, true, dft8_fwd_lazy(v))
M2_FLOOR_PROBE(inv8u_61, e61, 8, true, dft8_inv_lazy(v); for (int e = 1; e < 8; ++e) v[e] = {fold61(v[e].a), fold61(v[e].b)})
M2_FLOOR_PROBE(fwd4t_61, e61, 4, true, for (int e = 1; e < 4; ++e) v[e] = mul_tw(v[e], w[e - 1]); dft4_fwd_lazy(v))
M2_FLOOR_PROBE(inv4t_61, e61, 4, true, dft4_inv_lazy(v); for (int e = 1; e < 4; ++e) v[e] = mulconj_tw(v[e], w[e - 1]))
M2_FLOOR_PROBE(fwd2t_61, e61, 2, true, v[1] = mul_tw(v[1], w[0]); dft2_fwd_lazy(v[0], v[1]))
M2_FLOOR_PROBE(inv2t_61, e61, 2, true, dft2_inv_lazy(v[0], v[1]); v[1] = mulconj_tw(v[1], w[0]))
M2_FLOOR_PROBE(fwd2u_61, e61, 2, true, dft2_fwd_lazy(v[0], v[1]))
M2_FLOOR_PROBE(inv2u_61, e61, 2, true, dft2_inv_lazy(v[0], v[1]); for (int e = 0; e < 2; ++e) v[e] = {fold61(v[e].a), fold61(v[e].b)})
M2_FLOOR_PROBE(sq_61, e61, 2, true, square_pair(v[0], v[1], w[0]))

M2_FLOOR_SKEL(skel8_31, e31, 8)
M2_FLOOR_SKEL(skel4_31, e31, 4)
M2_FLOOR_SKEL(skel2_31, e31, 2)
M2_FLOOR_PROBE(fwd8t_31, e31, 8, true, for (int e = 1; e < 8; ++e) v[e] = mul_tw(v[e], w[e - 1]); dft8_fwd(v))
M2_FLOOR_PROBE(inv8t_31, e31, 8, true, dft8_inv(v); for (int e = 1; e < 8; ++e) v[e] = mulconj_tw(v[e], w[e - 1]))
M2_FLOOR_PROBE(fwd8u_31, e31, 8, true, dft8_fwd(v))
M2_FLOOR_PROBE(inv8u_31, e31, 8, true, dft8_inv(v))
M2_FLOOR_PROBE(fwd4t_31, e31, 4, true, for (int e = 1; e < 4; ++e) v[e] = mul_tw(v[e], w[e - 1]); dft4_fwd(v))
M2_FLOOR_PROBE(inv4t_31, e31, 4, true, dft4_inv(v); for (int e = 1; e < 4; ++e) v[e] = mulconj_tw(v[e], w[e - 1]))
M2_FLOOR_PROBE(fwd2t_31, e31, 2, true, v[1] = mul_tw(v[1], w[0]); dft2(v[0], v[1]))
M2_FLOOR_PROBE(inv2t_31, e31, 2, true, dft2(v[0], v[1]); v[1] = mulconj_tw(v[1], w[0]))
M2_FLOOR_PROBE(fwd2u_31, e31, 2, true, dft2(v[0], v[1]))
M2_FLOOR_PROBE(inv2u_31, e31, 2, true, dft2(v[0], v[1]))
M2_FLOOR_PROBE(sq_31, e31, 2, true, square_pair(v[0], v[1], w[0]))

// Carry of one segment (8 digits): spectrum in LDS, digits and digit
// metadata as in the outer passes. MODE 0: skeleton, 1: unweight + CRT +
// normalise, 2: ripple + weight, 3: both, 4: weight alone; every part takes
// its form for digit widths 16..32 (q = 26, as in both workloads), with bit 8
// of MODE the one for widths 32..44 (q = 40), with bit 16 the general one
// (q = 50).
template <int MODE>
__device__ void carry_probe(const TestDesc* tests, e61* zo) {
  __shared__ e61 L61[padded(PASS_MAX_ELEMS)];
  __shared__ e31 L31[padded(PASS_MAX_ELEMS)];
  const TestDesc td = tests[0];
  const int t = floor_probe::tid();
  const int idx0 = (int)threadIdx.x * SEG_ELEMS;
  for (int i = 0; i < SEG_ELEMS; ++i) {
    L61[PADX(idx0 + i)] = zo[t * 8 + i];
    L31[PADX(idx0 + i)] = e31{(uint32_t)zo[t * 8 + 4 + i].a, (uint32_t)zo[t * 8 + 4 + i].b};
  }
  SegDigits sd;
  seg_load(gptr(zo + t * 4 + 100000), sd);
  const SegMeta M = seg_meta(td.dgen, 2u * (uint32_t)t * SEG_ELEMS);
  const WidthMasks wm(td.dgen);
  const e61 cw = zo[t + 300000];
  i128 c = (i128)(((unsigned __int128)cw.b << 64) | cw.a);
  const uint32_t q = (MODE & 16) ? 50u : ((MODE & 8) ? 40u : 26u);
  if (MODE & 1) c = seg_normalise(L61, L31, idx0, M, wm, q, c, 0ull, sd);
  if (MODE & 2) seg_ripple(sd, M, wm, q, c);
  if (MODE & 6) seg_weight(sd, L61, L31, idx0, M, q);
  if (MODE != 3)
    for (int i = 0; i < 2 * SEG_ELEMS; ++i) sd.d[i] ^= M.m[i];
  seg_store(gptr(zo + t * 4 + 200000), sd);
  zo[t + 300000] = e61{(uint64_t)c, (uint64_t)(c >> 64)};
  for (int i = 0; i < SEG_ELEMS; ++i) {
    zo[t * 8 + i] = L61[PADX(idx0 + i)];
    zo[t * 8 + 4 + i] = e61{L31[PADX(idx0 + i)].a, L31[PADX(idx0 + i)].b};
  }
}
extern "C" __global__ void __launch_bounds__(256) skel_carry(const TestDesc* tests, e61* zo) { carry_probe<0>(tests, zo); }
extern "C" __global__ void __launch_bounds__(256) carry_norm(const TestDesc* tests, e61* zo) { carry_probe<1>(tests, zo); }
extern "C" __global__ void __launch_bounds__(256) carry_norm64(const TestDesc* tests, e61* zo) { carry_probe<9>(tests, zo); }
extern "C" __global__ void __launch_bounds__(256) carry_norm128(const TestDesc* tests, e61* zo) { carry_probe<17>(tests, zo); }
extern "C" __global__ void __launch_bounds__(256) carry_rw(const TestDesc* tests, e61* zo) { carry_probe<2>(tests, zo); }
extern "C" __global__ void __launch_bounds__(256) carry_all(const TestDesc* tests, e61* zo) { carry_probe<3>(tests, zo); }
extern "C" __global__ void __launch_bounds__(256) carry_w(const TestDesc* tests, e61* zo) { carry_probe<4>(tests, zo); }

}  // namespace m2
