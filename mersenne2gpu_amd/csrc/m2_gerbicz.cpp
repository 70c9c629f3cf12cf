// Gerbicz error check for PRP runs (the reference has none; its driver,
// mersenne2_gpu.cpp:4811-4853, trusts every iteration).
//
// With x_j the residue after j blocks of L squarings since a verified state
// x_0, the product d_t = x_0 x_1 ... x_t satisfies d_t = x_0 d_{t-1}^(2^L):
// d_{t-1}^(2^L) = x_0^(2^L) ... x_{t-1}^(2^L) = x_1 ... x_t when every block
// squared correctly. An error in any block breaks the identity (except with
// negligible probability), so checking it every B blocks, at the price of L
// extra squarings, covers all B blocks; a failed check rolls the test back to
// the last verified state.
//
// Everything runs on the squaring engine: x^(2^k) is k iterations of a
// scratch PRP test of the same exponent, and the one general product d x is
// ((d + x)^2 - (d - x)^2) / 4, two single squarings plus additions, a
// complement and a rotation by 2 bits on the host (4^-1 = 2^(p-2) mod 2^p - 1).
#include <stdint.h>

#include <string>
#include <vector>

#include "m2_internal.h"
#include "m2_modn.h"
#include "mersenne2.h"

namespace m2 {
namespace {

using namespace modn;

// Test hook: corrupt the residue after this block of the next checked run
// (1-based; 0: never), once.
uint64_t g_fault_block = 0;

// Residue transfers and the squarings the check needs, on a scratch test s of
// the same exponent.
struct Checked {
  m2_test* s = nullptr;
  ModN m;
  size_t len;
  explicit Checked(uint32_t p) : m(p), len((p + 7) / 8) {}
  ~Checked() {
    if (s) m2_test_destroy(s);
  }
  int read(m2_test* u, Num& out) {
    std::vector<uint8_t> b(len);
    int rc = m2_test_residue(u, b.data(), b.size());
    if (rc) return rc;
    out = from_bytes(b, m);
    return M2_OK;
  }
  int write(m2_test* u, const Num& a, uint64_t iteration) {
    const std::vector<uint8_t> b = to_bytes(a, m);
    return m2_test_set_residue(u, b.data(), b.size(), iteration);
  }
  // a^(2^k)
  int pow2k(const Num& a, uint64_t k, Num& out) {
    int rc;
    if ((rc = write(s, a, 0)) || (rc = m2_test_run(s, k)) || (rc = read(s, out))) return rc;
    return M2_OK;
  }
  // a b = ((a + b)^2 - (a - b)^2) / 4
  int mul(const Num& a, const Num& b, Num& out) {
    Num u, v;
    int rc;
    if ((rc = pow2k(add(a, b, m), 1, u)) || (rc = pow2k(add(a, neg(b, m), m), 1, v))) return rc;
    out = quarter(add(u, neg(v, m), m), m);
    return M2_OK;
  }
};

}  // namespace
}  // namespace m2

using namespace m2;

extern "C" {

int m2_test_run_checked(m2_test* t, uint64_t iters, uint32_t block, uint32_t blocks_per_check, uint64_t* errors) {
  if (!t) return test_error(M2_E_INVAL, "null test");
  if (block == 0 || blocks_per_check == 0) return test_error(M2_E_INVAL, "block and blocks_per_check must be > 0");
  m2_ctx* ctx = nullptr;
  uint64_t p = 0;
  int mode = 0;
  int rc = test_params(t, &ctx, &p, &mode);
  if (rc) return rc;
  if (mode != M2_MODE_PRP) return test_error(M2_E_INVAL, "the Gerbicz check applies to PRP tests");
  Checked c((uint32_t)p);
  // The scratch test keeps its outer passes in separate launches: its
  // squarings overlap the test's own launches on another stream, and a fused
  // outer pass needs all its workgroups resident at once.
  m2_opts so{};
  so.no_fuse = 1;
  if ((rc = m2_test_create(ctx, p, M2_MODE_PRP, &so, &c.s))) return rc;
  uint64_t start = 0;
  if ((rc = m2_test_iteration(t, &start))) return rc;
  uint64_t nerr = 0;
  // last verified state
  uint64_t vk = start;
  Num vx;
  if ((rc = c.read(t, vx))) return rc;
  Num x0 = vx, d = vx, dprev;
  uint32_t nb = 0;  // blocks since the last verified state
  int fails = 0;
  const uint64_t full_end = start + iters - iters % block;
  uint64_t fault = g_fault_block, blocks_run = 0;
  g_fault_block = 0;
  auto roll_back = [&]() -> int {
    ++nerr;
    if (++fails >= 3) return test_error(M2_E_STATE, "Gerbicz check failed 3 times in a row");
    int r = c.write(t, vx, vk);
    x0 = vx;
    d = vx;
    nb = 0;
    return r;
  };
  // The next block is enqueued on the test before the product and the check
  // of the block just read, so that their host work and the scratch test's
  // squarings (its own stream) overlap the test's own GPU work; after a failed
  // check the block in flight is waited for and discarded by the roll-back.
  bool queued = false;
  for (uint64_t k = vk; k < full_end;) {
    Num x;
    if (!queued && (rc = m2_test_run(t, block))) return rc;
    rc = c.read(t, x);  // waits for the block
    queued = false;
    if (rc == M2_E_STATE) {
      // a fused outer pass timed out waiting for a neighbour: the block's
      // state is invalid, so it is treated as a failed check
      if ((rc = roll_back())) return rc;
      k = vk;
      continue;
    }
    if (rc) return rc;
    k += block;
    if (++blocks_run == fault) {  // planted error (test hook): one flipped bit
      x[0] ^= 1ull << 5;
      if ((rc = c.write(t, x, k))) return rc;
    }
    if (k < full_end) {
      if ((rc = m2_test_run(t, block))) return rc;
      queued = true;
    }
    dprev = d;
    if ((rc = c.mul(d, x, d))) return rc;
    ++nb;
    if (nb < blocks_per_check && k < full_end) continue;
    Num y, z;
    if ((rc = c.pow2k(dprev, block, y)) || (rc = c.mul(x0, y, z))) return rc;
    if (z == d) {
      vk = k;
      vx = x;
      x0 = x;
      d = x;
      nb = 0;
      fails = 0;
    } else {
      queued = false;  // roll_back's residue write waits for the block in flight first
      if ((rc = roll_back())) return rc;
      k = vk;
    }
  }
  // the last iters % block iterations: run twice from the verified state and compared
  const uint64_t tail = iters % block;
  while (tail) {
    Num a, b;
    if ((rc = m2_test_run(t, tail)) || (rc = c.read(t, a))) return rc;
    if ((rc = c.write(t, vx, vk)) || (rc = m2_test_run(t, tail)) || (rc = c.read(t, b))) return rc;
    if (a == b) break;
    if ((rc = roll_back())) return rc;
  }
  if (errors) *errors = nerr;
  return m2_test_sync(t);
}

int m2_debug_gerbicz_fault(uint64_t block) {
  g_fault_block = block;
  return M2_OK;
}

}  // extern "C"
