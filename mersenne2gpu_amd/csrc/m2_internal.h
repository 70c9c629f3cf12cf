// Internal declarations shared by the engine (host) and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "m2_arith.h"
#include "m2_plan.h"

struct m2_ctx;
struct m2_test;

namespace m2 {

// One test as the kernels see it. Tests of the same transform length share a
// DevPlan (pass structure and twiddle tables) and run in the same launches,
// one descriptor per test (grid.y on the multi-pass plan, one workgroup per
// test on the single-workgroup plan).
struct TestDesc {
  e61* zd;       // multi-pass plan: digits as int64 pairs, strip-major (h slots)
  e61* z61;      // spectrum, M61 part (h elements); single-workgroup plan: the digits
  e31* z31;      // spectrum, M31 part
  Carry128* G;   // multi-pass plan: carries not yet applied (nseg)
  DigitGen dgen;
  int32_t mode;    // 0 = LL, 1 = PRP
  uint32_t iters;  // single-workgroup plan: iterations to run in this launch
  // Fused outer pass (single tests whose strips fit one workgroup per CU):
  // sync[c] counts strip c's launches, sync[NS] is set if a wait for a
  // neighbour timed out; cf holds the strip-row carries as tagged words (two
  // per strip row, tag = launch count mod 256).
  uint32_t* sync;
  uint64_t* cf;
};

// Plan and tables of one transform length (shared by every test of that length).
struct DevPlan {
  const e61* tw61 = nullptr;  // step twiddle tables
  const e31* tw31 = nullptr;
  const e61* sq61 = nullptr;  // centre twiddles, h/2 entries
  const e31* sq31 = nullptr;
  uint32_t h = 0;
  int ln = 0, L = 0;
  bool small = false;
  uint32_t cus = 0;  // compute units of the device the tables live on
  Pass outer{};  // small plan: the whole transform
  int logS = 0, logW = 0;
  bool has_mid = false;
  Pass mid{};
  int mid_logS = 0, mid_logW = 0;
  Pass centre{};
  uint32_t nseg = 0;  // number of carry segments (length of G)
  int centre_nt = MP_THREADS_NARROW;  // threads per centre workgroup (the strided passes use 512)
  // Centre twiddles regrouped per centre workgroup: record g (crec entries)
  // holds, for its blocks (B, B2), the twiddles of all but the last step in
  // the order the workgroup keeps them in LDS, then the register slots of
  // centre_slots(centre_nt), each of role_threads entries: slot s, entry t is
  // the twiddle thread t of either field keeps in register s
  // (centre_register_twiddles).
  const e61* ctw61 = nullptr;
  const e31* ctw31 = nullptr;
  uint32_t crec = 0;
};

// Blocks (B, B2) squared together by centre workgroup g: (0, 1) for g = 0,
// else the pair B + B2 = 3 2^a - 1 within octave a (the real-data pairing).
inline void centre_blocks(uint32_t g, uint32_t& B, uint32_t& B2) {
  if (g == 0) {
    B = 0;
    B2 = 1;
    return;
  }
  uint32_t a = 0;
  while ((2u << a) <= g) ++a;  // 2^a <= g < 2^(a+1)
  ++a;                         // blocks [2^a, 2^(a+1))
  B = (1u << a) + (g - (1u << (a - 1)));
  B2 = 3u * (1u << a) - 1u - B;
}

// Position of element e's digit pair between iterations: natural order on the
// small plan, strip-major (strip cg = column / W, then row, then column
// within the strip) on the multi-pass plan.
inline uint64_t digit_slot(const DevPlan& P, uint64_t e) {
  if (P.small) return e;
  const uint64_t S = uint64_t(1) << P.logS, W = uint64_t(1) << P.logW, R = uint64_t(1) << P.outer.A;
  const uint64_t r = e >> P.logS, c = e & (S - 1);
  return ((c >> P.logW) * R + r) * W + (c & (W - 1));
}
// First digit of the strip-row that carry G[g] enters (g = strip * R + row).
inline uint64_t carry_target_digit(const DevPlan& P, uint64_t g) {
  const uint64_t R = uint64_t(1) << P.outer.A, NS = (uint64_t(1) << P.logS) >> P.logW;
  uint64_t cg = g / R, r = g % R;
  if (++cg == NS) {
    cg = 0;
    if (++r == R) r = 0;
  }
  return 2 * ((r << P.logS) + (cg << P.logW));
}

// ---- launchers (m2_kernels.hip); `tests` is a device array of ntests descriptors ----
hipError_t launch_small(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st);
// v_mad_u64_u32 issue-rate probe: `blocks` x 256 threads into out[blocks * 256];
// *lane_ops receives the multiply-adds the launch issues.
hipError_t launch_mad_rate(uint64_t* out, uint32_t blocks, uint32_t iters, hipStream_t st, uint64_t* lane_ops);
hipError_t launch_iteration(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st);
// k iterations of one test with the outer passes of consecutive iterations
// fused: outer_fwd, (centre, outer_fused) x (k - 1), centre, outer_inv.
hipError_t launch_run_fused(const DevPlan& P, const TestDesc* tests, uint64_t k, hipStream_t st);
// The same with every launch's dispatch timestamps (at most max_events launches).
hipError_t launch_run_fused_events(const DevPlan& P, const TestDesc* tests, uint64_t k, hipStream_t st,
                                   hipEvent_t* starts, hipEvent_t* stops, int max_events, int* count,
                                   const char** names);
// Poll limit of the fused pass's wait for a neighbour (launches made after a
// change use the new value; captured graphs keep the value they were made with).
extern uint32_t g_fused_spin_limit;
// Number of outer-pass strips (workgroups per outer launch) of plan P.
inline uint32_t outer_strips(const DevPlan& P) { return (1u << P.logS) >> P.logW; }
// One iteration with each kernel's dispatch timestamps in starts[k]/stops[k].
hipError_t launch_iteration_events(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st,
                                   hipEvent_t* starts, hipEvent_t* stops, int max_events, int* count,
                                   const char** names);
hipError_t launch_small_events(const DevPlan& P, const TestDesc* tests, uint32_t ntests, hipStream_t st,
                               hipEvent_t start, hipEvent_t stop);

// ---- engine internals shared with m2_gerbicz.cpp and m2_noalloc.hip (m2_api.cpp) ----
// Device index of a context.
int ctx_device(const ::m2_ctx* ctx);
// Set-up of the no-alloc entry point on `device` (the current device): the
// scratch array's address, a stream, the single-workgroup tables (once per
// process). Called by m2_open; m2_noalloc.hip.
int noalloc_init(int device);
// Records `msg` as the calling thread's last error and returns `code`.
int test_error(int code, const std::string& msg);
// Context, exponent and mode of a test.
int test_params(const ::m2_test* t, ::m2_ctx** ctx, uint64_t* p, int* mode);

// ---- host side (m2_host.cpp) ----
int transform_log2(uint64_t p);

struct ClassTables {
  std::vector<e61> tw61;
  std::vector<e31> tw31;
  std::vector<e61> sq61;  // h/2
  std::vector<e31> sq31;
  std::vector<e61> ctw61;  // multi-pass plan: per-workgroup centre records
  std::vector<e31> ctw31;
};

// Plan choices for one transform length; zero fields mean automatic.
//   ls_cap      (-LS) the centre block is the largest power of two <= ls_cap
//               (at least 4) and the multi-pass plan is used;
//   carry_pairs (-B)  digit pairs per strip row of the outer pass (the carry
//               block), rounded down to a power of two and clamped to what
//               the outer pass's row count allows;
//   max_radix   8 (radix-8 steps) or 4 (radix-4 steps);
//   threads     (-W)  centre workgroup size: 512 (0: automatic) or 1024
//               (4 elements per thread, radix-4 centre steps).
struct PlanOpts {
  uint32_t ls_cap = 0, carry_pairs = 0, max_radix = 0;
  uint32_t threads = 0;
  bool operator<(const PlanOpts& o) const {
    if (ls_cap != o.ls_cap) return ls_cap < o.ls_cap;
    if (carry_pairs != o.carry_pairs) return carry_pairs < o.carry_pairs;
    if (max_radix != o.max_radix) return max_radix < o.max_radix;
    return threads < o.threads;
  }
};

// Plan and twiddle tables for transform length 2^ln.
bool build_class_plan(int ln, const PlanOpts& opts, DevPlan& P, ClassTables& T, std::string& err);
// Walk constants of exponent p on plan P. Checks the kernels' incremental
// walk against the closed forms (width, weight shifts) at every digit, and
// the carry bound of the plan.
bool make_digit_gen(uint32_t p, const DevPlan& P, DigitGen& D, std::string& err);
// Bit position of digit j of exponent p on a length-2^ln transform:
// ceil(p j / 2^ln) (digit j has width digit_bitpos(j + 1) - digit_bitpos(j)).
inline uint64_t digit_bitpos(uint32_t p, int ln, uint64_t j) {
  return j ? ((((uint64_t)p * j) - 1) >> ln) + 1 : 0;
}
std::string describe_plan(const DevPlan& P);

// Canonical residue (little-endian, ceil(p/8) bytes) of the state given by
// signed digits plus carries pending at given digit indices.
std::vector<uint8_t> residue_from_state(uint32_t p, int ln, const int64_t* digits, uint64_t ndigits,
                                        const Carry128* carries, const uint64_t* carry_digit, uint64_t ncarries);
// The 2^ln digits (int64) of a canonical residue.
void digits_from_residue(uint32_t p, int ln, const uint8_t* r, size_t len, int64_t* digits);

}  // namespace m2
