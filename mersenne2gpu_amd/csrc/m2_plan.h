// Transform plan shared by the host engine and the HIP kernels.
//
// One squaring of x mod 2^p - 1 runs on h = n/2 elements (digit pairs packed as
// a + b i) in GF(M61^2) and GF(M31^2). The length-h transform is a sequence of
// radix-2 levels l = 0..L-1 (L = log2 h); level l splits each block j of size
// h/2^l with the twiddle tau = zeta_{l+1,2j} (the reference's w[2^l + j],
// mersenne2_gpu.cpp:3911-3917). Levels are grouped into radix-8 steps (one
// radix-2 or radix-4 step takes the remainder) and the steps into passes:
//   * small plan (h <= SMALL_H): one workgroup keeps the whole residue in LDS
//     and runs many iterations per launch;
//   * large plan: an outer pass over the first A levels (columns of stride
//     S = h/2^A, carry-fused at both ends), an optional middle pass, and a
//     centre pass over contiguous blocks of C = 2^c elements that does the
//     last forward levels, the real-data squaring and the first inverse levels.
#pragma once
#include <stdint.h>

namespace m2 {

constexpr int MAX_STEPS = 8;
constexpr int SMALL_LOG_H = 12;     // h <= 4096 elements (96 KB of data in LDS)
constexpr int CENTRE_LOG_C = 10;    // centre block: 1024 elements, two per workgroup
constexpr int PASS_MAX_ELEMS = 2048;  // R * W elements per strided-pass workgroup
constexpr int PASS_MAX_LOG_R = 9;     // strided passes: at most 2^9 rows (R - 1 twiddles)
// Threads per workgroup of the multi-pass kernels, half per field: 512 (8
// elements per thread of a 2048-element workgroup, radix-8 steps), or, for
// the centre pass, 1024 (4 elements per thread, radix-4 steps: twice the
// waves per SIMD, for a grid of at most one workgroup per CU, which would
// leave a SIMD two waves).
constexpr int MP_THREADS_NARROW = 512;
constexpr int MP_THREADS_WIDE = 1024;
// Centre workgroups: each thread keeps in registers the twiddles of its
// butterflies in the last forward step (the first inverse step uses them
// conjugated) and of its squaring pairs; only the other steps' twiddles are
// staged in LDS. Slots per thread, for nt threads per workgroup:
//   512: last step radix 8: 1 x 7 or radix 4: 2 x 3, then 2^c <= 4 x 256 pairs
//   1024: last step radix 4: 1 x 3, then 2^c <= 2 x 512 pairs
struct CentreSlots {
  int role_threads, step_slots, sq_slots;
};
constexpr CentreSlots centre_slots(int nt) {
  return nt == MP_THREADS_WIDE ? CentreSlots{512, 3, 2} : CentreSlots{256, 7, 4};
}
constexpr int CENTRE_EARLY_TW = 512;   // the other steps' twiddles, both blocks
constexpr int SEG_ELEMS = 4;        // carry segment: 4 elements = 8 digits

// One radix step inside a pass: global level `ell`, radix 2/4/8, and the
// offset of its twiddle block (tau^1..tau^(r-1) for each block j at level ell).
struct Step {
  int32_t ell;
  int32_t radix;
  uint32_t twoff;
};

struct Pass {
  int32_t ell0;   // first level of the pass
  int32_t A;      // number of levels (rows R = 2^A)
  int32_t nsteps;
  Step steps[MAX_STEPS];
};

// Digit metadata, one per digit (n of them): width in bits and the weight
// shifts in each field (host tables; the kernels derive the same values on
// the fly from DigitGen).
struct DigitInfo {
  uint8_t width, w61, w31, pad;
};

// Constants from which a kernel walks the digit metadata incrementally. With
// r_j = p j mod n (p odd) and s = p mod n:
//   width_j = q + [r_j + s > n] + [r_j = 0]
//   w61_j   = lr61 (n - r_j) mod 61 for r_j != 0, else 0   (same for 31)
// and from digit j to j+1: r += s (mod n), w61' = w61' - d61 + [wrap] (mod 61),
// where w61' = lr61 (n - r) mod 61 is tracked for every r (it is 1 at r = 0).
struct DigitGen {
  uint32_t p;
  uint32_t nmask;  // n - 1
  uint32_t s;      // p mod n
  uint32_t q;      // p >> ln
  uint32_t lr61, lr31;
  uint32_t d61, d31;  // lr * s mod 61 / 31
  uint32_t k61, k31;  // (ln + 1) mod 61 / 31: extra right shift of the unweight
};

struct Carry128 {
  uint64_t lo;
  int64_t hi;
};

}  // namespace m2
