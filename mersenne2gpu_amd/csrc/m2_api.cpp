// C ABI of libm2.so (declared in include/mersenne2.h).
//
// Engine objects:
//   Class   - plan and twiddle tables of one transform length (cached per context);
//   TestObj - one resident test: digits, spectrum and carry buffers, digit constants;
//   Group   - tests of one class that run in the same launches, on one stream.
// A single test (m2_test) is a group of one; a batch (m2_batch) and the
// server (m2_server) keep one group per transform length, on separate
// streams, so different lengths overlap on the device.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "m2_internal.h"
#include "mersenne2.h"

using namespace m2;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                                                \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return fail(M2_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr uint32_t SMALL_CHUNK = 4096;  // max iterations per test in one single-workgroup launch
constexpr uint32_t GRAPH_ITERS = 16;    // iterations per captured graph (multi-pass plan)

uint64_t full_iters(uint32_t p, int mode) { return mode == M2_MODE_LL ? (p >= 2 ? p - 2 : 0) : p; }

struct Class {
  int ln = 0;
  uint32_t ls_cap = 0;
  DevPlan P;
  std::vector<void*> allocs;
  uint64_t bytes = 0;
};

struct TestObj {
  Class* cls = nullptr;
  uint32_t p = 0;
  int mode = 1;
  uint64_t iter = 0;    // iterations done
  uint64_t target = 0;  // iterations wanted (batch / server)
  DigitGen dgen{};
  TestTables tt;
  TestDesc desc{};
  std::vector<void*> allocs;
  uint64_t bytes = 0;
};

int dev_alloc(std::vector<void*>& allocs, uint64_t& bytes, void** dst, const void* src, size_t n) {
  void* d = nullptr;
  if (n == 0) n = 16;
  hipError_t e = hipMalloc(&d, n);
  if (e != hipSuccess) return fail(M2_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  allocs.push_back(d);
  bytes += n;
  if (src) {
    e = hipMemcpy(d, src, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(M2_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  *dst = d;
  return M2_OK;
}

void free_all(std::vector<void*>& allocs) {
  for (void* a : allocs) (void)hipFree(a);
  allocs.clear();
}

}  // namespace

struct m2_ctx {
  int device = 0;
  hipDeviceProp_t prop{};
  std::mutex mu;
  std::map<std::pair<int, uint32_t>, std::unique_ptr<Class>> classes;
};

namespace {

int get_class(m2_ctx* ctx, int ln, uint32_t ls_cap, Class** out) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto key = std::make_pair(ln, ls_cap);
  auto it = ctx->classes.find(key);
  if (it != ctx->classes.end()) {
    *out = it->second.get();
    return M2_OK;
  }
  std::unique_ptr<Class> c(new Class);
  c->ln = ln;
  c->ls_cap = ls_cap;
  ClassTables T;
  std::string err;
  if (!build_class_plan(ln, ls_cap, c->P, T, err)) return fail(M2_E_RANGE, err);
  void* d;
  int rc;
  auto bail = [&](int code) {
    free_all(c->allocs);
    return code;
  };
  if ((rc = dev_alloc(c->allocs, c->bytes, &d, T.tw61.data(), T.tw61.size() * sizeof(e61)))) return bail(rc);
  c->P.tw61 = (const e61*)d;
  if ((rc = dev_alloc(c->allocs, c->bytes, &d, T.tw31.data(), T.tw31.size() * sizeof(e31)))) return bail(rc);
  c->P.tw31 = (const e31*)d;
  if ((rc = dev_alloc(c->allocs, c->bytes, &d, T.sq61.data(), T.sq61.size() * sizeof(e61)))) return bail(rc);
  c->P.sq61 = (const e61*)d;
  if ((rc = dev_alloc(c->allocs, c->bytes, &d, T.sq31.data(), T.sq31.size() * sizeof(e31)))) return bail(rc);
  c->P.sq31 = (const e31*)d;
  *out = c.get();
  ctx->classes[key] = std::move(c);
  return M2_OK;
}

// digits: n digits in natural order
int upload_digits(TestObj* t, const int64_t* digits, hipStream_t st) {
  const DevPlan& P = t->cls->P;
  const uint64_t h = P.h;
  std::vector<int64_t> slots(2 * h);
  for (uint64_t e = 0; e < h; ++e) {
    const uint64_t s = digit_slot(P, e);
    slots[2 * s] = digits[2 * e];
    slots[2 * s + 1] = digits[2 * e + 1];
  }
  HIPCHK(hipMemcpyAsync(P.small ? t->desc.z61 : t->desc.zd, slots.data(), h * sizeof(e61), hipMemcpyHostToDevice,
                        st));
  if (!P.small) {
    std::vector<Carry128> zero(P.nseg, Carry128{0, 0});
    HIPCHK(hipMemcpyAsync(t->desc.G, zero.data(), zero.size() * sizeof(Carry128), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return M2_OK;
}

int test_obj_create(m2_ctx* ctx, uint64_t p, int mode, uint32_t ls_cap, hipStream_t st, TestObj** out) {
  if (mode != M2_MODE_LL && mode != M2_MODE_PRP) return fail(M2_E_INVAL, "mode must be LL (0) or PRP (1)");
  if (p < 3 || p > 0xFFFFFFFFull) return fail(M2_E_RANGE, "exponent out of range (3 <= p < 2^32)");
  const int ln = transform_log2(p);
  if (ln > 27) return fail(M2_E_RANGE, "exponent too large: transform length above 2^27");
  Class* cls = nullptr;
  int rc = get_class(ctx, ln, ls_cap, &cls);
  if (rc) return rc;
  std::unique_ptr<TestObj> t(new TestObj);
  t->cls = cls;
  t->p = (uint32_t)p;
  t->mode = mode;
  t->target = full_iters(t->p, mode);
  std::string err;
  if (!build_test_tables(t->p, cls->P, t->tt, t->dgen, err)) return fail(M2_E_RANGE, err);
  const DevPlan& P = cls->P;
  const uint64_t h = P.h;
  void* d;
  auto bail = [&](int code) {
    free_all(t->allocs);
    return code;
  };
  if ((rc = dev_alloc(t->allocs, t->bytes, &d, nullptr, h * sizeof(e61)))) return bail(rc);
  t->desc.z61 = (e61*)d;
  if ((rc = dev_alloc(t->allocs, t->bytes, &d, nullptr, h * sizeof(e31)))) return bail(rc);
  t->desc.z31 = (e31*)d;
  if (!P.small) {
    if ((rc = dev_alloc(t->allocs, t->bytes, &d, nullptr, h * sizeof(e61)))) return bail(rc);
    t->desc.zd = (e61*)d;
    if ((rc = dev_alloc(t->allocs, t->bytes, &d, nullptr, P.nseg * sizeof(Carry128)))) return bail(rc);
    t->desc.G = (Carry128*)d;
  }
  t->desc.dgen = t->dgen;
  t->desc.mode = mode;
  t->desc.iters = 0;
  // initial residue: 3 (PRP) or 4 (LL) in digit 0 (mersenne2_gpu.cpp:3908)
  std::vector<int64_t> digits(2 * h, 0);
  digits[0] = (mode == M2_MODE_LL) ? 4 : 3;
  if ((rc = upload_digits(t.get(), digits.data(), st))) return bail(rc);
  *out = t.release();
  return M2_OK;
}

void test_obj_destroy(TestObj* t) {
  if (!t) return;
  free_all(t->allocs);
  delete t;
}

int test_obj_residue(TestObj* t, std::vector<uint8_t>& r) {
  const DevPlan& P = t->cls->P;
  const uint64_t h = P.h, n = 2 * h;
  std::vector<int64_t> slots(n), digits(n);
  HIPCHK(hipMemcpy(slots.data(), P.small ? t->desc.z61 : t->desc.zd, h * sizeof(e61), hipMemcpyDeviceToHost));
  for (uint64_t e = 0; e < h; ++e) {
    const uint64_t s = digit_slot(P, e);
    digits[2 * e] = slots[2 * s];
    digits[2 * e + 1] = slots[2 * s + 1];
  }
  std::vector<Carry128> G;
  std::vector<uint64_t> at;
  if (!P.small) {
    G.resize(P.nseg);
    HIPCHK(hipMemcpy(G.data(), t->desc.G, G.size() * sizeof(Carry128), hipMemcpyDeviceToHost));
    at.resize(G.size());
    for (uint64_t g = 0; g < G.size(); ++g) at[g] = carry_target_digit(P, g);
  }
  r = residue_from_state(t->p, t->tt.bitpos, digits.data(), n, G.data(), at.data(), G.size());
  return M2_OK;
}

uint64_t res64_of(const std::vector<uint8_t>& r) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | (i < (int)r.size() ? r[i] : 0);
  return v;
}

// 1 if the residue is 0 (LL) or 9 mod M_p (PRP), the reference's verdicts.
int verdict_of(const TestObj* t, const std::vector<uint8_t>& r) {
  std::vector<uint8_t> want(r.size(), 0);
  if (t->mode == M2_MODE_PRP) {
    uint64_t nine = 9;
    if (t->p < 64) nine %= ((uint64_t(1) << t->p) - 1);
    for (size_t i = 0; i < want.size() && i < 8; ++i) want[i] = (uint8_t)(nine >> (8 * i));
  }
  return r == want ? 1 : 0;
}

// Tests of one class sharing launches on one stream. A launch covers a set
// of tests given by a device array of their descriptors. On the multi-pass
// plan the descriptor array and a captured graph of graph_iters iterations
// are kept per active set and reused while that set stays the same; on the
// single-workgroup plan each launch carries its own iteration counts.
struct Group {
  static constexpr uint32_t MAX_GRID_ROWS = 65535;  // tests per multi-pass launch (grid.y)
  Class* cls = nullptr;
  hipStream_t stream = nullptr;
  struct Round {
    TestDesc* d = nullptr;
    uint32_t n = 0;
    uint32_t G = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
  };
  std::map<std::pair<std::vector<TestObj*>, uint32_t>, Round> cache;
  std::vector<void*> cache_allocs;
  std::vector<void*> transient;  // per-launch descriptor arrays, freed once the stream is drained
  uint64_t bytes = 0;

  int create(Class* c) {
    cls = c;
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    return M2_OK;
  }
  void drop_cache() {
    for (auto& kv : cache) {
      if (kv.second.gexec) (void)hipGraphExecDestroy(kv.second.gexec);
      if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    cache.clear();
    free_all(cache_allocs);
  }
  void destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    free_all(transient);
    drop_cache();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
  int sync() {
    HIPCHK(hipStreamSynchronize(stream));
    free_all(transient);
    return M2_OK;
  }
  // Forget cached launches that name a test about to be destroyed (drains the stream).
  int forget(const TestObj* t) {
    HIPCHK(hipStreamSynchronize(stream));
    free_all(transient);
    bool hit = false;
    for (auto& kv : cache)
      for (TestObj* u : kv.first.first) hit |= (u == t);
    if (hit) drop_cache();
    return M2_OK;
  }
  // Descriptor array of `tests` (iteration counts for the single-workgroup plan).
  int upload_descs(const std::vector<TestObj*>& tests, const std::vector<uint32_t>* iters,
                   std::vector<void*>& owner, TestDesc** out) {
    std::vector<TestDesc> h(tests.size());
    for (size_t i = 0; i < tests.size(); ++i) {
      h[i] = tests[i]->desc;
      h[i].iters = iters ? (*iters)[i] : 0;
    }
    void* d = nullptr;
    int rc = dev_alloc(owner, bytes, &d, h.data(), h.size() * sizeof(TestDesc));
    if (rc) return rc;
    *out = (TestDesc*)d;
    return M2_OK;
  }
  // Cached multi-pass round of `tests`, with its graph when G > 1.
  int get_round(const std::vector<TestObj*>& tests, uint32_t G, Round** out) {
    auto key = std::make_pair(tests, G);
    auto it = cache.find(key);
    if (it != cache.end()) {
      *out = &it->second;
      return M2_OK;
    }
    Round r;
    r.n = (uint32_t)tests.size();
    r.G = G;
    int rc = upload_descs(tests, nullptr, cache_allocs, &r.d);
    if (rc) return rc;
    if (G > 1) {
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      for (uint32_t i = 0; i < G; ++i) {
        hipError_t e = launch_iteration(cls->P, r.d, r.n, stream);
        if (e != hipSuccess) {
          hipGraph_t g;
          (void)hipStreamEndCapture(stream, &g);
          return fail(M2_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        }
      }
      HIPCHK(hipStreamEndCapture(stream, &r.graph));
      HIPCHK(hipGraphInstantiate(&r.gexec, r.graph, nullptr, nullptr, 0));
    }
    *out = &(cache[key] = r);
    return M2_OK;
  }
  // Single-workgroup plan: one launch, test i runs iters[i] iterations.
  int run_small(const std::vector<TestObj*>& tests, const std::vector<uint32_t>& iters) {
    TestDesc* d = nullptr;
    int rc = upload_descs(tests, &iters, transient, &d);
    if (rc) return rc;
    HIPCHK(launch_small(cls->P, d, (uint32_t)tests.size(), stream));
    for (size_t i = 0; i < tests.size(); ++i) tests[i]->iter += iters[i];
    return M2_OK;
  }
  // Multi-pass plan: k iterations of every test, G per graph launch.
  int run_multi(const std::vector<TestObj*>& tests, uint64_t k, uint32_t G) {
    if (tests.empty() || k == 0) return M2_OK;
    if (k < G) G = 1;
    Round* r = nullptr;
    int rc = get_round(tests, G, &r);
    if (rc) return rc;
    uint64_t left = k;
    if (G > 1)
      for (; left >= G; left -= G) HIPCHK(hipGraphLaunch(r->gexec, stream));
    for (; left; --left) HIPCHK(launch_iteration(cls->P, r->d, r->n, stream));
    for (TestObj* t : tests) t->iter += k;
    return M2_OK;
  }
  // Build what advance() will launch for `tests` (keeps setup out of timings).
  int prepare(const std::vector<TestObj*>& tests, uint32_t graph_iters) {
    if (cls->P.small || tests.size() > MAX_GRID_ROWS) return M2_OK;
    Round* r = nullptr;
    return get_round(tests, graph_iters ? graph_iters : GRAPH_ITERS, &r);
  }
  // Advance every test towards its target, at most `limit` iterations each
  // (0: no limit), enqueuing launches without waiting.
  int advance(std::vector<TestObj*> tests, uint64_t limit, uint32_t graph_iters) {
    tests.erase(std::remove_if(tests.begin(), tests.end(), [](TestObj* t) { return t->iter >= t->target; }),
                tests.end());
    if (tests.empty()) return M2_OK;
    const uint32_t G = graph_iters ? graph_iters : GRAPH_ITERS;
    std::vector<uint64_t> stop(tests.size());
    for (size_t i = 0; i < tests.size(); ++i)
      stop[i] = limit ? std::min(tests[i]->target, tests[i]->iter + limit) : tests[i]->target;
    if (cls->P.small) {
      for (;;) {
        std::vector<TestObj*> act;
        std::vector<uint32_t> its;
        for (size_t i = 0; i < tests.size(); ++i)
          if (tests[i]->iter < stop[i]) {
            act.push_back(tests[i]);
            its.push_back((uint32_t)std::min<uint64_t>(SMALL_CHUNK, stop[i] - tests[i]->iter));
          }
        if (act.empty()) break;
        int rc = run_small(act, its);
        if (rc) return rc;
      }
      return M2_OK;
    }
    for (;;) {
      std::vector<TestObj*> act;
      uint64_t k = ~0ull;
      for (size_t i = 0; i < tests.size(); ++i)
        if (tests[i]->iter < stop[i]) {
          act.push_back(tests[i]);
          k = std::min(k, stop[i] - tests[i]->iter);
        }
      if (act.empty()) break;
      for (size_t c = 0; c < act.size(); c += MAX_GRID_ROWS) {
        std::vector<TestObj*> part(act.begin() + c, act.begin() + std::min(act.size(), c + MAX_GRID_ROWS));
        int rc = run_multi(part, k, G);
        if (rc) return rc;
      }
    }
    return M2_OK;
  }
};

}  // namespace

// A single test: a test object with its own group.
struct m2_test {
  m2_ctx* ctx = nullptr;
  TestObj* obj = nullptr;
  Group group;
  uint32_t graph_iters = GRAPH_ITERS;
};

struct m2_batch {
  m2_ctx* ctx = nullptr;
  std::vector<TestObj*> tests;
  std::map<Class*, Group> groups;
};

extern "C" {

const char* m2_last_error(void) { return g_err.c_str(); }

int m2_device_count(int* count) {
  if (!count) return fail(M2_E_INVAL, "null pointer");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return M2_OK;
}

int m2_open(int device, m2_ctx** out) {
  if (!out) return fail(M2_E_INVAL, "null pointer");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c == 0) return fail(M2_E_DEVICE, "No GPU device found");
  if (device < 0 || device >= c)
    return fail(M2_E_DEVICE, "Invalid GPU device index " + std::to_string(device) + " (available 0.." +
                                 std::to_string(c - 1) + ")");
  std::unique_ptr<m2_ctx> ctx(new m2_ctx);
  ctx->device = device;
  HIPCHK(hipGetDeviceProperties(&ctx->prop, device));
  HIPCHK(hipSetDevice(device));
  *out = ctx.release();
  return M2_OK;
}

int m2_close(m2_ctx* ctx) {
  if (!ctx) return M2_OK;
  (void)hipSetDevice(ctx->device);
  for (auto& kv : ctx->classes) free_all(kv.second->allocs);
  delete ctx;
  return M2_OK;
}

int m2_device_name(m2_ctx* ctx, char* buf, size_t len) {
  if (!ctx || !buf || !len) return fail(M2_E_INVAL, "null pointer");
  snprintf(buf, len, "%s", ctx->prop.gcnArchName);
  return M2_OK;
}

int m2_device_props(m2_ctx* ctx, uint32_t* cu, uint64_t* mem, uint32_t* max_wg, uint32_t* lds) {
  if (!ctx) return fail(M2_E_INVAL, "null pointer");
  if (cu) *cu = (uint32_t)ctx->prop.multiProcessorCount;
  if (mem) *mem = (uint64_t)ctx->prop.totalGlobalMem;
  if (max_wg) *max_wg = (uint32_t)ctx->prop.maxThreadsPerBlock;
  if (lds) *lds = (uint32_t)ctx->prop.sharedMemPerBlock;
  return M2_OK;
}

int m2_transform_log2(uint64_t p, int* ln) {
  if (!ln) return fail(M2_E_INVAL, "null pointer");
  if (p < 3 || p > 0xFFFFFFFFull) return fail(M2_E_RANGE, "exponent out of range");
  *ln = transform_log2(p);
  return M2_OK;
}

int m2_test_footprint(uint64_t p, uint64_t* bytes) {
  if (!bytes) return fail(M2_E_INVAL, "null pointer");
  if (p < 3 || p > 0xFFFFFFFFull) return fail(M2_E_RANGE, "exponent out of range");
  const uint64_t h = (uint64_t(1) << transform_log2(p)) / 2;
  // per test: spectrum 24 B/element, digits 16 B/element, pending carries
  // <= 16 B per 4 elements; shared per length: step twiddles < 24 B/element,
  // centre twiddles 12 B/element
  *bytes = 24 * h + 16 * h + 4 * h + 24 * h + 12 * h + 4096;
  return M2_OK;
}

int m2_test_create(m2_ctx* ctx, uint64_t p, int mode, const m2_opts* opts, m2_test** out) {
  if (!ctx || !out) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<m2_test> t(new m2_test);
  t->ctx = ctx;
  if (opts && opts->graph_iters) t->graph_iters = opts->graph_iters;
  const int ln = (p >= 3 && p <= 0xFFFFFFFFull) ? transform_log2(p) : 0;
  if (ln < 3 || ln > 27) return fail(M2_E_RANGE, "exponent out of range (3 <= p < 2^32, n <= 2^27)");
  Class* cls = nullptr;
  int rc = get_class(ctx, ln, opts ? opts->local_stage_cap : 0u, &cls);
  if (rc) return rc;
  if ((rc = t->group.create(cls))) return rc;
  if ((rc = test_obj_create(ctx, p, mode, opts ? opts->local_stage_cap : 0u, t->group.stream, &t->obj))) {
    t->group.destroy();
    return rc;
  }
  t->obj->target = ~0ull;  // a single test runs as many iterations as asked
  *out = t.release();
  return M2_OK;
}

int m2_test_destroy(m2_test* t) {
  if (!t) return M2_OK;
  (void)hipSetDevice(t->ctx->device);
  t->group.destroy();
  test_obj_destroy(t->obj);
  delete t;
  return M2_OK;
}

int m2_test_run(m2_test* t, uint64_t iters) {
  if (!t) return fail(M2_E_INVAL, "null test");
  if (iters == 0) return M2_OK;
  HIPCHK(hipSetDevice(t->ctx->device));
  return t->group.advance({t->obj}, iters, t->graph_iters);
}

int m2_test_sync(m2_test* t) {
  if (!t) return fail(M2_E_INVAL, "null test");
  return t->group.sync();
}

int m2_test_iteration(m2_test* t, uint64_t* iter) {
  if (!t || !iter) return fail(M2_E_INVAL, "null pointer");
  *iter = t->obj->iter;
  return M2_OK;
}

int m2_test_total_iters(m2_test* t, uint64_t* iters) {
  if (!t || !iters) return fail(M2_E_INVAL, "null pointer");
  *iters = full_iters(t->obj->p, t->obj->mode);
  return M2_OK;
}

int m2_test_residue(m2_test* t, uint8_t* buf, size_t len) {
  if (!t || !buf) return fail(M2_E_INVAL, "null pointer");
  const size_t need = (t->obj->p + 7) / 8;
  if (len < need) return fail(M2_E_INVAL, "residue buffer too small: need " + std::to_string(need) + " bytes");
  int rc = t->group.sync();
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t->obj, r))) return rc;
  memset(buf, 0, len);
  memcpy(buf, r.data(), r.size());
  return M2_OK;
}

int m2_test_res64(m2_test* t, uint64_t* res64) {
  if (!t || !res64) return fail(M2_E_INVAL, "null pointer");
  int rc = t->group.sync();
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t->obj, r))) return rc;
  *res64 = res64_of(r);
  return M2_OK;
}

int m2_test_is_prime(m2_test* t, int* verdict) {
  if (!t || !verdict) return fail(M2_E_INVAL, "null pointer");
  int rc = t->group.sync();
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t->obj, r))) return rc;
  *verdict = verdict_of(t->obj, r);
  return M2_OK;
}

int m2_test_set_residue(m2_test* t, const uint8_t* buf, size_t len, uint64_t iteration) {
  if (!t || !buf) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  int rc = t->group.sync();
  if (rc) return rc;
  std::vector<int64_t> digits(2ull * t->obj->cls->P.h);
  digits_from_residue(t->obj->p, t->obj->tt.bitpos, t->obj->tt.dig, buf, len, digits.data());
  if ((rc = upload_digits(t->obj, digits.data(), t->group.stream))) return rc;
  t->obj->iter = iteration;
  return M2_OK;
}

int m2_test_time(m2_test* t, uint64_t iters, double* ms) {
  if (!t || !ms) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  int rc = t->group.sync();
  if (rc) return rc;
  hipEvent_t a, b;
  if ((rc = t->group.prepare({t->obj}, t->graph_iters))) return rc;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, t->group.stream));
  if ((rc = t->group.advance({t->obj}, iters, t->graph_iters))) return rc;
  HIPCHK(hipEventRecord(b, t->group.stream));
  HIPCHK(hipEventSynchronize(b));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = f;
  return t->group.sync();
}

int m2_test_kernel_times(m2_test* t, uint64_t iters, double* ms, int max_kernels, int* count, char* names_buf,
                         size_t names_len) {
  if (!t || !ms || !count) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  int rc = t->group.sync();
  if (rc) return rc;
  constexpr int MAXE = 8;
  hipEvent_t starts[MAXE], stops[MAXE];
  for (int k = 0; k < MAXE; ++k) {
    HIPCHK(hipEventCreate(&starts[k]));
    HIPCHK(hipEventCreate(&stops[k]));
  }
  TestDesc* d = nullptr;
  std::vector<uint32_t> one(1, 1);
  if ((rc = t->group.upload_descs({t->obj}, &one, t->group.transient, &d))) return rc;
  const char* names[MAXE] = {nullptr};
  int nk = 0;
  std::vector<double> acc(MAXE, 0.0);
  for (uint64_t it = 0; it < iters; ++it) {
    if (t->obj->cls->P.small) {
      HIPCHK(launch_small_events(t->obj->cls->P, d, 1, t->group.stream, starts[0], stops[0]));
      nk = 1;
      names[0] = "k_small";
    } else {
      HIPCHK(launch_iteration_events(t->obj->cls->P, d, 1, t->group.stream, starts, stops, MAXE, &nk, names));
    }
    ++t->obj->iter;
    HIPCHK(hipEventSynchronize(stops[nk - 1]));
    for (int k = 0; k < nk; ++k) {
      float f = 0;
      HIPCHK(hipEventElapsedTime(&f, starts[k], stops[k]));
      acc[k] += f;
    }
  }
  for (int k = 0; k < MAXE; ++k) {
    (void)hipEventDestroy(starts[k]);
    (void)hipEventDestroy(stops[k]);
  }
  if ((rc = t->group.sync())) return rc;
  const int m = std::min(nk, max_kernels);
  std::string nm;
  for (int k = 0; k < m; ++k) {
    ms[k] = acc[k];
    nm += names[k] ? names[k] : "?";
    nm += "\n";
  }
  *count = m;
  if (names_buf && names_len) snprintf(names_buf, names_len, "%s", nm.c_str());
  return M2_OK;
}

int m2_test_plan(m2_test* t, char* buf, size_t len) {
  if (!t || !buf || !len) return fail(M2_E_INVAL, "null pointer");
  snprintf(buf, len, "%s", describe_plan(t->obj->cls->P).c_str());
  return M2_OK;
}

// ---- batches ----

int m2_batch_create(m2_ctx* ctx, size_t count, const uint64_t* ps, const int* modes, const uint64_t* iters,
                    const m2_opts* opts, m2_batch** out) {
  if (!ctx || !out || (count && (!ps || !modes))) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<m2_batch> b(new m2_batch);
  b->ctx = ctx;
  const uint32_t ls = opts ? opts->local_stage_cap : 0u;
  auto bail = [&](int code) {
    for (auto& kv : b->groups) kv.second.destroy();
    for (TestObj* t : b->tests) test_obj_destroy(t);
    return code;
  };
  for (size_t i = 0; i < count; ++i) {
    const int ln = (ps[i] >= 3 && ps[i] <= 0xFFFFFFFFull) ? transform_log2(ps[i]) : 0;
    if (ln < 3 || ln > 27) return bail(fail(M2_E_RANGE, "exponent out of range at index " + std::to_string(i)));
    Class* cls = nullptr;
    int rc = get_class(ctx, ln, ls, &cls);
    if (rc) return bail(rc);
    auto it = b->groups.find(cls);
    if (it == b->groups.end()) {
      Group g;
      if ((rc = g.create(cls))) return bail(rc);
      it = b->groups.emplace(cls, g).first;
    }
    TestObj* t = nullptr;
    if ((rc = test_obj_create(ctx, ps[i], modes[i], ls, it->second.stream, &t))) return bail(rc);
    if (iters && iters[i]) t->target = iters[i];
    b->tests.push_back(t);
  }
  *out = b.release();
  return M2_OK;
}

int m2_batch_destroy(m2_batch* b) {
  if (!b) return M2_OK;
  (void)hipSetDevice(b->ctx->device);
  for (auto& kv : b->groups) kv.second.destroy();
  for (TestObj* t : b->tests) test_obj_destroy(t);
  delete b;
  return M2_OK;
}

int m2_batch_run(m2_batch* b, uint64_t max_iters) {
  if (!b) return fail(M2_E_INVAL, "null batch");
  HIPCHK(hipSetDevice(b->ctx->device));
  std::map<Class*, std::vector<TestObj*>> by;
  for (TestObj* t : b->tests) by[t->cls].push_back(t);
  for (auto& kv : by) {
    int rc = b->groups[kv.first].advance(kv.second, max_iters, GRAPH_ITERS);
    if (rc) return rc;
  }
  return M2_OK;
}

int m2_batch_sync(m2_batch* b) {
  if (!b) return fail(M2_E_INVAL, "null batch");
  for (auto& kv : b->groups) {
    int rc = kv.second.sync();
    if (rc) return rc;
  }
  return M2_OK;
}

static int batch_test(m2_batch* b, size_t i, TestObj** t) {
  if (!b) return fail(M2_E_INVAL, "null batch");
  if (i >= b->tests.size()) return fail(M2_E_INVAL, "test index out of range");
  int rc = b->groups[b->tests[i]->cls].sync();
  if (rc) return rc;
  *t = b->tests[i];
  return M2_OK;
}

int m2_batch_iteration(m2_batch* b, size_t i, uint64_t* iter) {
  if (!iter) return fail(M2_E_INVAL, "null pointer");
  if (!b || i >= b->tests.size()) return fail(M2_E_INVAL, "test index out of range");
  *iter = b->tests[i]->iter;
  return M2_OK;
}

int m2_batch_res64(m2_batch* b, size_t i, uint64_t* res64) {
  TestObj* t = nullptr;
  if (!res64) return fail(M2_E_INVAL, "null pointer");
  int rc = batch_test(b, i, &t);
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t, r))) return rc;
  *res64 = res64_of(r);
  return M2_OK;
}

int m2_batch_residue(m2_batch* b, size_t i, uint8_t* buf, size_t len) {
  TestObj* t = nullptr;
  if (!buf) return fail(M2_E_INVAL, "null pointer");
  int rc = batch_test(b, i, &t);
  if (rc) return rc;
  const size_t need = (t->p + 7) / 8;
  if (len < need) return fail(M2_E_INVAL, "residue buffer too small: need " + std::to_string(need) + " bytes");
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t, r))) return rc;
  memset(buf, 0, len);
  memcpy(buf, r.data(), r.size());
  return M2_OK;
}

int m2_batch_is_prime(m2_batch* b, size_t i, int* verdict) {
  TestObj* t = nullptr;
  if (!verdict) return fail(M2_E_INVAL, "null pointer");
  int rc = batch_test(b, i, &t);
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t, r))) return rc;
  *verdict = verdict_of(t, r);
  return M2_OK;
}

}  // extern "C"

// ---- server ----

struct m2_server {
  m2_ctx* ctx = nullptr;
  uint64_t budget = 0, in_use = 0;
  uint32_t slice = 1000;
  uint32_t ls_cap = 0;
  std::mutex mu;  // guards `queue` and `done` (submit/poll may come from other threads)
  struct Req {
    uint64_t id, p, iters;
    int mode;
  };
  std::deque<Req> queue;
  struct Active {
    uint64_t id;
    TestObj* t;
    uint64_t bytes;
  };
  std::vector<Active> active;
  std::map<Class*, Group> groups;
  struct Done {
    uint64_t id, res64;
    int verdict;
  };
  std::deque<Done> done;
  uint64_t next_id = 1;
};

extern "C" {

int m2_server_create(m2_ctx* ctx, uint64_t hbm_budget_bytes, uint32_t slice_iters, const m2_opts* opts,
                     m2_server** out) {
  if (!ctx || !out) return fail(M2_E_INVAL, "null pointer");
  std::unique_ptr<m2_server> s(new m2_server);
  s->ctx = ctx;
  s->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(0.7 * (double)ctx->prop.totalGlobalMem);
  s->slice = slice_iters ? slice_iters : 1000;
  s->ls_cap = opts ? opts->local_stage_cap : 0u;
  *out = s.release();
  return M2_OK;
}

int m2_server_destroy(m2_server* s) {
  if (!s) return M2_OK;
  (void)hipSetDevice(s->ctx->device);
  for (auto& kv : s->groups) kv.second.destroy();
  for (auto& a : s->active) test_obj_destroy(a.t);
  delete s;
  return M2_OK;
}

int m2_server_submit(m2_server* s, uint64_t p, int mode, uint64_t iters, uint64_t* id) {
  if (!s) return fail(M2_E_INVAL, "null server");
  if (mode != M2_MODE_LL && mode != M2_MODE_PRP) return fail(M2_E_INVAL, "mode must be LL (0) or PRP (1)");
  if (p < 3 || p > 0xFFFFFFFFull || transform_log2(p) > 27) return fail(M2_E_RANGE, "exponent out of range");
  if ((p & 1u) == 0) return fail(M2_E_RANGE, "exponent must be odd");
  std::lock_guard<std::mutex> lk(s->mu);
  const uint64_t rid = s->next_id++;
  s->queue.push_back({rid, p, iters, mode});
  if (id) *id = rid;
  return M2_OK;
}

int m2_server_step(m2_server* s, int* busy) {
  if (!s) return fail(M2_E_INVAL, "null server");
  HIPCHK(hipSetDevice(s->ctx->device));
  // 1. admit queued requests in arrival order while their footprint fits
  for (;;) {
    m2_server::Req rq;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->queue.empty()) break;
      rq = s->queue.front();
    }
    uint64_t fp = 0;
    m2_test_footprint(rq.p, &fp);
    if (s->in_use + fp > s->budget && !s->active.empty()) break;
    Class* cls = nullptr;
    int rc = get_class(s->ctx, transform_log2(rq.p), s->ls_cap, &cls);
    if (rc) return rc;
    auto it = s->groups.find(cls);
    if (it == s->groups.end()) {
      Group g;
      if ((rc = g.create(cls))) return rc;
      it = s->groups.emplace(cls, g).first;
    }
    TestObj* t = nullptr;
    if ((rc = test_obj_create(s->ctx, rq.p, rq.mode, s->ls_cap, it->second.stream, &t))) return rc;
    if (rq.iters) t->target = std::min<uint64_t>(t->target, rq.iters);
    s->active.push_back({rq.id, t, fp});
    s->in_use += fp;
    std::lock_guard<std::mutex> lk(s->mu);
    s->queue.pop_front();
  }
  // 2. one time slice for every active test, each length on its own stream
  std::map<Class*, std::vector<TestObj*>> by;
  for (auto& a : s->active) by[a.t->cls].push_back(a.t);
  for (auto& kv : by) {
    int rc = s->groups[kv.first].advance(kv.second, s->slice, GRAPH_ITERS);
    if (rc) return rc;
  }
  for (auto& kv : by) {
    int rc = s->groups[kv.first].sync();
    if (rc) return rc;
  }
  // 3. retire finished tests
  std::vector<m2_server::Active> keep;
  for (auto& a : s->active) {
    if (a.t->iter >= a.t->target) {
      std::vector<uint8_t> r;
      int rc = test_obj_residue(a.t, r);
      if (rc) return rc;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->done.push_back({a.id, res64_of(r), verdict_of(a.t, r)});
      }
      s->in_use -= a.bytes;
      if ((rc = s->groups[a.t->cls].forget(a.t))) return rc;
      test_obj_destroy(a.t);
    } else {
      keep.push_back(a);
    }
  }
  s->active.swap(keep);
  if (busy) {
    std::lock_guard<std::mutex> lk(s->mu);
    *busy = (!s->active.empty() || !s->queue.empty()) ? 1 : 0;
  }
  return M2_OK;
}

int m2_server_poll(m2_server* s, uint64_t* ids, uint64_t* res64, int* verdicts, size_t max, size_t* count) {
  if (!s || !count) return fail(M2_E_INVAL, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  size_t k = 0;
  while (k < max && !s->done.empty()) {
    const auto& d = s->done.front();
    if (ids) ids[k] = d.id;
    if (res64) res64[k] = d.res64;
    if (verdicts) verdicts[k] = d.verdict;
    s->done.pop_front();
    ++k;
  }
  *count = k;
  return M2_OK;
}

int m2_server_stats(m2_server* s, uint64_t* queued, uint64_t* active, uint64_t* bytes_in_use) {
  if (!s) return fail(M2_E_INVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = s->queue.size();
  if (active) *active = s->active.size();
  if (bytes_in_use) *bytes_in_use = s->in_use;
  return M2_OK;
}

}  // extern "C"
