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
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
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

// HIP events destroyed on every return path of the function that owns them.
struct Events {
  std::vector<hipEvent_t> ev;
  Events() = default;
  Events(const Events&) = delete;
  Events& operator=(const Events&) = delete;
  ~Events() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  hipError_t create(size_t n) {
    ev.reserve(ev.size() + n);
    for (size_t i = 0; i < n; ++i) {
      hipEvent_t e = nullptr;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      ev.push_back(e);
    }
    return hipSuccess;
  }
  hipEvent_t operator[](size_t i) const { return ev[i]; }
};

constexpr uint32_t SMALL_CHUNK = 4096;  // max iterations per test in one single-workgroup launch
constexpr uint32_t GRAPH_ITERS = 16;    // iterations per captured graph (multi-pass plan)

uint64_t full_iters(uint32_t p, int mode) { return mode == M2_MODE_LL ? (p >= 2 ? p - 2 : 0) : p; }

// Plan for a single test on gfx950 when the caller leaves the plan fields at
// zero (the counterpart of the reference's per-device profile table,
// mersenne2_gpu.cpp:3515-3616): m2_autotune's choices for one test, measured
// with the fused outer pass (us/iteration, default plan -> chosen):
//   n = 2^13: 39.1 -> 30.0 (centre blocks 2^7, carry block 16)   2^14: 40.2 -> 32.3 (2^7, 16)
//   n = 2^15: 44.0 -> 34.0 (centre blocks 2^9, carry block 16)   2^16: 44.7 -> 36.4 (2^9, 16)
//   n = 2^17: 46.1 -> 40.0 (centre blocks 2^7)                   2^18: 50.0 -> 44.4 (2^9)
//   n = 2^19: 53.8 -> 48.7 (centre blocks 2^9)
//   n = 2^20: 58.0 -> 56.5 (1024-thread centre: two blocks of 2^10 on 256
//             workgroups, one per CU, leave a SIMD two 512-thread waves)
//   others: the default plan.
PlanOpts single_test_profile(int ln) {
  PlanOpts r;
  switch (ln) {
    case 13: case 14: r.ls_cap = 128; r.carry_pairs = 16; break;
    case 15: case 16: r.ls_cap = 512; r.carry_pairs = 16; break;
    case 17: r.ls_cap = 128; break;
    case 18: case 19: r.ls_cap = 512; break;
    case 20: r.threads = MP_THREADS_WIDE; break;
    default: break;
  }
  return r;
}

// Iterations per captured graph for a single test, by length: each graph
// starts with a forward and ends with an inverse outer pass instead of a fused
// one, and graph launches leave short gaps, so long graphs pay where an
// iteration is long (n = 2^20: 59.1 us/iteration at 16, 58.5 at 64, 58.3 at
// 256, measured on the MI355X).
uint32_t single_test_graph_iters(int ln) { return ln >= 17 ? 256u : (ln >= 13 ? 64u : GRAPH_ITERS); }

PlanOpts plan_opts(const m2_opts* o) {
  PlanOpts r;
  if (o) {
    r.ls_cap = o->local_stage_cap;
    r.carry_pairs = o->carry_pairs;
    r.max_radix = o->max_radix;
    // -W 1024 or 512 picks the multi-pass kernels' workgroup size; other
    // values leave it to the engine
    if (o->wg == (uint32_t)MP_THREADS_WIDE || o->wg == (uint32_t)MP_THREADS_NARROW) r.threads = o->wg;
  }
  return r;
}

struct Class {
  int ln = 0;
  PlanOpts popts;
  DevPlan P;
  bool fusable = false;  // single tests may use the fused outer pass (strips fit one workgroup per CU)
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
  std::map<std::pair<int, PlanOpts>, std::unique_ptr<Class>> classes;
  // autotuned choices per (transform length, batch-size bucket)
  std::map<std::pair<int, uint32_t>, m2_opts> tuned;
};

namespace {

int get_class(m2_ctx* ctx, int ln, const PlanOpts& po, Class** out) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto key = std::make_pair(ln, po);
  auto it = ctx->classes.find(key);
  if (it != ctx->classes.end()) {
    *out = it->second.get();
    return M2_OK;
  }
  std::unique_ptr<Class> c(new Class);
  c->ln = ln;
  c->popts = po;
  ClassTables T;
  std::string err;
  if (!build_class_plan(ln, po, c->P, T, err)) return fail(M2_E_RANGE, err);
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
  c->P.cus = (uint32_t)ctx->prop.multiProcessorCount;
  c->fusable = !c->P.small && !c->P.has_mid && outer_strips(c->P) <= c->P.cus;
  if (!c->P.small) {
    if ((rc = dev_alloc(c->allocs, c->bytes, &d, T.ctw61.data(), T.ctw61.size() * sizeof(e61)))) return bail(rc);
    c->P.ctw61 = (const e61*)d;
    if ((rc = dev_alloc(c->allocs, c->bytes, &d, T.ctw31.data(), T.ctw31.size() * sizeof(e31)))) return bail(rc);
    c->P.ctw31 = (const e31*)d;
  }
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
  if (t->desc.sync) {
    // a new state: clear the fused pass's time-out word (the per-strip launch
    // counters keep counting)
    const uint32_t zero = 0;
    HIPCHK(hipMemcpyAsync(t->desc.sync + outer_strips(P), &zero, sizeof zero, hipMemcpyHostToDevice, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return M2_OK;
}

// fused: the test may run on the fused outer pass (single tests), which needs
// the per-strip launch counters and the tagged carry words.
int test_obj_create(m2_ctx* ctx, uint64_t p, int mode, const PlanOpts& po, hipStream_t st, bool fused,
                    TestObj** out) {
  if (mode != M2_MODE_LL && mode != M2_MODE_PRP) return fail(M2_E_INVAL, "mode must be LL (0) or PRP (1)");
  if (p < 3 || p > 0xFFFFFFFFull) return fail(M2_E_RANGE, "exponent out of range (3 <= p < 2^32)");
  const int ln = transform_log2(p);
  if (ln > 27) return fail(M2_E_RANGE, "exponent too large: transform length above 2^27");
  Class* cls = nullptr;
  int rc = get_class(ctx, ln, po, &cls);
  if (rc) return rc;
  std::unique_ptr<TestObj> t(new TestObj);
  t->cls = cls;
  t->p = (uint32_t)p;
  t->mode = mode;
  t->target = full_iters(t->p, mode);
  std::string err;
  if (!make_digit_gen(t->p, cls->P, t->dgen, err)) return fail(M2_E_RANGE, err);
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
  if (fused && cls->fusable) {
    const size_t nsync = (outer_strips(P) + 1) * sizeof(uint32_t);
    if ((rc = dev_alloc(t->allocs, t->bytes, &d, nullptr, nsync))) return bail(rc);
    t->desc.sync = (uint32_t*)d;
    HIPCHK(hipMemsetAsync(d, 0, nsync, st));
    const size_t ncf = (size_t)P.nseg * 2 * sizeof(uint64_t);
    if ((rc = dev_alloc(t->allocs, t->bytes, &d, nullptr, ncf))) return bail(rc);
    t->desc.cf = (uint64_t*)d;
    HIPCHK(hipMemsetAsync(d, 0, ncf, st));
  }
  t->desc.dgen = t->dgen;
  t->desc.mode = mode;
  t->desc.iters = 0;
  // initial residue: 3 (PRP) or 4 (LL) in digit 0 (mersenne2_gpu.cpp:3908),
  // every other digit and pending carry 0: cleared on the device, and the
  // element holding digit 0 written
  {
    e61* z = P.small ? t->desc.z61 : t->desc.zd;
    const e61 first{(uint64_t)(mode == M2_MODE_LL ? 4 : 3), 0};
    HIPCHK(hipMemsetAsync(z, 0, h * sizeof(e61), st));
    HIPCHK(hipMemcpyAsync(z + digit_slot(P, 0), &first, sizeof first, hipMemcpyHostToDevice, st));
    if (!P.small) HIPCHK(hipMemsetAsync(t->desc.G, 0, P.nseg * sizeof(Carry128), st));
    HIPCHK(hipStreamSynchronize(st));
  }
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
  // staging kept per thread up to n = 2^22 (64 MiB): a server reads one
  // residue after another, and fresh pages for every read cost more than the
  // conversion itself at that length
  static thread_local std::vector<int64_t> keep_slots, keep_digits;
  std::vector<int64_t> own_slots, own_digits;
  const bool keep = n <= (uint64_t(1) << 22);
  std::vector<int64_t>& slots = keep ? keep_slots : own_slots;
  std::vector<int64_t>& digits = keep ? keep_digits : own_digits;
  if (slots.size() < n) slots.resize(n);
  if (digits.size() < n) digits.resize(n);
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
  r = residue_from_state(t->p, P.ln, digits.data(), n, G.data(), at.data(), G.size());
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
  uint32_t graph_iters = GRAPH_ITERS;  // used when advance() is given 0
  bool fused = false;                  // single test: outer passes of consecutive iterations fused
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
      hipError_t e = hipSuccess;
      if (use_fused(r.n)) e = launch_run_fused(cls->P, r.d, G, stream);
      else
        for (uint32_t i = 0; i < G && e == hipSuccess; ++i) e = launch_iteration(cls->P, r.d, r.n, stream);
      if (e != hipSuccess) {
        hipGraph_t g;
        (void)hipStreamEndCapture(stream, &g);
        return fail(M2_E_HIP, std::string("capture: ") + hipGetErrorString(e));
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
    if (left && use_fused(r->n)) HIPCHK(launch_run_fused(cls->P, r->d, left, stream));
    else
      for (; left; --left) HIPCHK(launch_iteration(cls->P, r->d, r->n, stream));
    for (TestObj* t : tests) t->iter += k;
    return M2_OK;
  }
  bool use_fused(uint32_t ntests) const { return fused && ntests == 1 && cls->fusable; }
  // Build what advance() will launch for `tests` (keeps setup out of timings).
  int prepare(const std::vector<TestObj*>& tests, uint32_t graph_iters) {
    if (cls->P.small || tests.size() > MAX_GRID_ROWS) return M2_OK;
    Round* r = nullptr;
    return get_round(tests, graph_iters ? graph_iters : this->graph_iters, &r);
  }
  // Advance every test towards its target, at most `limit` iterations each
  // (0: no limit), enqueuing launches without waiting.
  int advance(std::vector<TestObj*> tests, uint64_t limit, uint32_t graph_iters) {
    tests.erase(std::remove_if(tests.begin(), tests.end(), [](TestObj* t) { return t->iter >= t->target; }),
                tests.end());
    if (tests.empty()) return M2_OK;
    const uint32_t G = graph_iters ? graph_iters : this->graph_iters;
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

namespace m2 {
int test_error(int code, const std::string& msg) { return fail(code, msg); }
int ctx_device(const m2_ctx* ctx) { return ctx->device; }
int test_params(const m2_test* t, m2_ctx** ctx, uint64_t* p, int* mode) {
  if (!t) return fail(M2_E_INVAL, "null test");
  *ctx = t->ctx;
  *p = t->obj->p;
  *mode = t->obj->mode;
  return M2_OK;
}
}  // namespace m2

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
  // set-up of m2_square_noalloc on this device; a failure here is reported
  // by that entry point (which retries it), not by m2_open
  (void)noalloc_init(device);
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
  // <= 16 B per 4 elements; shared per length (counted in full, an upper
  // bound): step twiddles < 24 B/element, centre twiddles 12 B/element and
  // their per-workgroup records 3070 entries of 24 B per 2048 elements
  *bytes = 24 * h + 16 * h + 4 * h + 24 * h + 12 * h + 36 * h + 4096;
  return M2_OK;
}

int m2_test_create(m2_ctx* ctx, uint64_t p, int mode, const m2_opts* opts, m2_test** out) {
  if (!ctx || !out) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<m2_test> t(new m2_test);
  t->ctx = ctx;
  const int ln = (p >= 3 && p <= 0xFFFFFFFFull) ? transform_log2(p) : 0;
  if (ln < 3 || ln > 27) return fail(M2_E_RANGE, "exponent out of range (3 <= p < 2^32, n <= 2^27)");
  t->graph_iters = (opts && opts->graph_iters) ? opts->graph_iters : single_test_graph_iters(ln);
  PlanOpts po = plan_opts(opts);
  if (!po.ls_cap && !po.carry_pairs && !po.max_radix && !po.threads) po = single_test_profile(ln);
  // an exponent whose digits are too narrow for the profile's multi-pass plan keeps the default
  {
    const uint32_t q = (uint32_t)(p >> ln);
    if (po.ls_cap && q < 12) po = plan_opts(opts);
  }
  Class* cls = nullptr;
  int rc = get_class(ctx, ln, po, &cls);
  if (rc) return rc;
  if ((rc = t->group.create(cls))) return rc;
  if ((rc = test_obj_create(ctx, p, mode, po, t->group.stream, true, &t->obj))) {
    t->group.destroy();
    return rc;
  }
  t->obj->target = ~0ull;  // a single test runs as many iterations as asked
  t->group.fused = cls->fusable && !(opts && opts->no_fuse);
  *out = t.release();
  return M2_OK;
}

// Waits for the test's stream; on the fused plan also checks that no
// workgroup timed out waiting for a neighbour's carries.
static int test_sync(m2_test* t) {
  int rc = t->group.sync();
  if (rc || !t->obj->desc.sync) return rc;
  uint32_t fault = 0;
  HIPCHK(hipMemcpy(&fault, t->obj->desc.sync + outer_strips(t->obj->cls->P), sizeof fault, hipMemcpyDeviceToHost));
  if (fault)
    return fail(M2_E_STATE, "fused outer pass: a workgroup timed out waiting for its neighbour's carries "
                            "(another kernel held the GPU); the test's state is invalid");
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
  return test_sync(t);
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
  int rc = test_sync(t);
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t->obj, r))) return rc;
  memset(buf, 0, len);
  memcpy(buf, r.data(), r.size());
  return M2_OK;
}

int m2_test_res64(m2_test* t, uint64_t* res64) {
  if (!t || !res64) return fail(M2_E_INVAL, "null pointer");
  int rc = test_sync(t);
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t->obj, r))) return rc;
  *res64 = res64_of(r);
  return M2_OK;
}

int m2_test_is_prime(m2_test* t, int* verdict) {
  if (!t || !verdict) return fail(M2_E_INVAL, "null pointer");
  int rc = test_sync(t);
  if (rc) return rc;
  std::vector<uint8_t> r;
  if ((rc = test_obj_residue(t->obj, r))) return rc;
  *verdict = verdict_of(t->obj, r);
  return M2_OK;
}

int m2_test_set_residue(m2_test* t, const uint8_t* buf, size_t len, uint64_t iteration) {
  if (!t || !buf) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  // the state is replaced, so a fused-pass time-out in it does not matter
  // (upload_digits clears the time-out word)
  int rc = t->group.sync();
  if (rc) return rc;
  std::vector<int64_t> digits(2ull * t->obj->cls->P.h);
  digits_from_residue(t->obj->p, t->obj->cls->P.ln, buf, len, digits.data());
  if ((rc = upload_digits(t->obj, digits.data(), t->group.stream))) return rc;
  t->obj->iter = iteration;
  return M2_OK;
}

int m2_test_time(m2_test* t, uint64_t iters, double* ms) {
  if (!t || !ms) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  int rc = test_sync(t);
  if (rc) return rc;
  if ((rc = t->group.prepare({t->obj}, t->graph_iters))) return rc;
  Events ev;
  HIPCHK(ev.create(2));
  HIPCHK(hipEventRecord(ev[0], t->group.stream));
  if ((rc = t->group.advance({t->obj}, iters, t->graph_iters))) return rc;
  HIPCHK(hipEventRecord(ev[1], t->group.stream));
  HIPCHK(hipEventSynchronize(ev[1]));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, ev[0], ev[1]));
  *ms = f;
  return test_sync(t);
}

}  // extern "C"

namespace {

// Runs `iters` iterations with every launch between its own dispatch
// timestamps, in runs of up to RUN iterations (no graphs): per-kernel total
// device time and launch count, kernels in order of first launch.
struct KernelProfile {
  std::vector<std::string> order;
  std::map<std::string, std::pair<double, uint64_t>> acc;  // name -> (total ms, launches)
};

int kernel_profile(m2_test* t, uint64_t iters, KernelProfile& out) {
  HIPCHK(hipSetDevice(t->ctx->device));
  int rc = test_sync(t);
  if (rc) return rc;
  constexpr uint64_t RUN = 16;
  constexpr int MAXE = 2 * (int)RUN + 8;
  Events ev_start, ev_stop;
  HIPCHK(ev_start.create(MAXE));
  HIPCHK(ev_stop.create(MAXE));
  hipEvent_t* starts = ev_start.ev.data();
  hipEvent_t* stops = ev_stop.ev.data();
  TestDesc* d = nullptr;
  std::vector<uint32_t> one(1, 1);
  if ((rc = t->group.upload_descs({t->obj}, &one, t->group.transient, &d))) return rc;
  const DevPlan& P = t->obj->cls->P;
  const bool fused = t->group.use_fused(1);
  for (uint64_t done = 0; done < iters;) {
    const char* names[MAXE] = {nullptr};
    int nk = 0;
    uint64_t k = 1;
    if (P.small) {
      HIPCHK(launch_small_events(P, d, 1, t->group.stream, starts[0], stops[0]));
      nk = 1;
      names[0] = "k_small";
    } else if (fused) {
      k = std::min<uint64_t>(RUN, iters - done);
      HIPCHK(launch_run_fused_events(P, d, k, t->group.stream, starts, stops, MAXE, &nk, names));
    } else {
      HIPCHK(launch_iteration_events(P, d, 1, t->group.stream, starts, stops, MAXE, &nk, names));
    }
    done += k;
    t->obj->iter += k;
    HIPCHK(hipEventSynchronize(stops[nk - 1]));
    for (int j = 0; j < nk; ++j) {
      float f = 0;
      HIPCHK(hipEventElapsedTime(&f, starts[j], stops[j]));
      const std::string nm = names[j] ? names[j] : "?";
      if (!out.acc.count(nm)) out.order.push_back(nm);
      out.acc[nm].first += f;
      out.acc[nm].second += 1;
    }
  }
  return test_sync(t);
}

}  // namespace

extern "C" {

int m2_test_kernel_times(m2_test* t, uint64_t iters, double* ms, int max_kernels, int* count, char* names_buf,
                         size_t names_len) {
  if (!t || !ms || !count) return fail(M2_E_INVAL, "null pointer");
  KernelProfile kp;
  int rc = kernel_profile(t, iters, kp);
  if (rc) return rc;
  const int m = std::min((int)kp.order.size(), max_kernels);
  std::string nm;
  for (int k = 0; k < m; ++k) {
    const auto& a = kp.acc[kp.order[k]];
    ms[k] = a.first / (double)a.second;
    nm += kp.order[k] + "\n";
  }
  *count = m;
  if (names_buf && names_len) snprintf(names_buf, names_len, "%s", nm.c_str());
  return M2_OK;
}

int m2_test_kernel_profile(m2_test* t, uint64_t iters, double* total_ms, uint64_t* launches, int max_kernels,
                           int* count, char* names_buf, size_t names_len) {
  if (!t || !total_ms || !launches || !count) return fail(M2_E_INVAL, "null pointer");
  KernelProfile kp;
  int rc = kernel_profile(t, iters, kp);
  if (rc) return rc;
  const int m = std::min((int)kp.order.size(), max_kernels);
  std::string nm;
  for (int k = 0; k < m; ++k) {
    total_ms[k] = kp.acc[kp.order[k]].first;
    launches[k] = kp.acc[kp.order[k]].second;
    nm += kp.order[k] + "\n";
  }
  *count = m;
  if (names_buf && names_len) snprintf(names_buf, names_len, "%s", nm.c_str());
  return M2_OK;
}

int m2_debug_fused_spin_limit(uint32_t polls, uint32_t* previous) {
  if (previous) *previous = g_fused_spin_limit;
  g_fused_spin_limit = polls;
  return M2_OK;
}

int m2_debug_product_rate(m2_ctx* ctx, double* lane_ops_per_s) {
  if (!ctx || !lane_ops_per_s) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(ctx->device));
  // 16 workgroups of 256 per CU; a short launch first (clocks up, code loaded)
  const uint32_t blocks = 16u * (uint32_t)std::max(1, ctx->prop.multiProcessorCount);
  uint64_t* out = nullptr;
  HIPCHK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint64_t)));
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  uint64_t ops = 0;
  float ms = 0;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e == hipSuccess) e = launch_mad_rate(out, blocks, 256, st, &ops);
  if (e == hipSuccess) e = hipEventRecord(a, st);
  if (e == hipSuccess) e = launch_mad_rate(out, blocks, 8192, st, &ops);
  if (e == hipSuccess) e = hipEventRecord(b, st);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (b) (void)hipEventDestroy(b);
  if (a) (void)hipEventDestroy(a);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(M2_E_HIP, std::string("product rate probe: ") + hipGetErrorString(e));
  if (ms <= 0) return fail(M2_E_HIP, "product rate probe: no elapsed time");
  *lane_ops_per_s = (double)ops / (ms * 1e-3);
  return M2_OK;
}

// ---- checkpoints ----
// File layout (little-endian): magic "M2CKPT01", u32 p, u32 mode, u64
// iteration, u64 Res64, u64 n = ceil(p/8), n residue bytes (canonical R),
// u64 FNV-1a hash of every preceding byte.

static uint64_t fnv1a(const uint8_t* d, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= d[i];
    h *= 1099511628211ull;
  }
  return h;
}
static void put_u32(std::vector<uint8_t>& b, uint32_t v) {
  for (int i = 0; i < 4; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}
static void put_u64(std::vector<uint8_t>& b, uint64_t v) {
  for (int i = 0; i < 8; ++i) b.push_back((uint8_t)(v >> (8 * i)));
}
static uint64_t get_u64(const uint8_t* d) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | d[i];
  return v;
}
static uint32_t get_u32(const uint8_t* d) {
  uint32_t v = 0;
  for (int i = 3; i >= 0; --i) v = (v << 8) | d[i];
  return v;
}

struct CkptData {
  uint32_t p = 0;
  int mode = 0;
  uint64_t iter = 0, res64 = 0;
  std::vector<uint8_t> residue;
};

static int ckpt_read(const char* path, CkptData& c) {
  FILE* f = fopen(path, "rb");
  if (!f) return fail(M2_E_INVAL, std::string("cannot open checkpoint ") + path);
  std::vector<uint8_t> b;
  uint8_t tmp[65536];
  size_t k;
  while ((k = fread(tmp, 1, sizeof tmp, f)) > 0) b.insert(b.end(), tmp, tmp + k);
  fclose(f);
  const size_t head = 8 + 4 + 4 + 8 + 8 + 8;
  if (b.size() < head + 8 || memcmp(b.data(), "M2CKPT01", 8) != 0)
    return fail(M2_E_INVAL, std::string("not a checkpoint file: ") + path);
  c.p = get_u32(&b[8]);
  c.mode = (int)get_u32(&b[12]);
  c.iter = get_u64(&b[16]);
  c.res64 = get_u64(&b[24]);
  const uint64_t n = get_u64(&b[32]);
  if (c.p < 3 || (c.mode != M2_MODE_LL && c.mode != M2_MODE_PRP) || n != (c.p + 7u) / 8u ||
      b.size() != head + n + 8)
    return fail(M2_E_INVAL, std::string("malformed checkpoint: ") + path);
  if (fnv1a(b.data(), head + n) != get_u64(&b[head + n]))
    return fail(M2_E_INVAL, std::string("checkpoint checksum mismatch: ") + path);
  c.residue.assign(b.begin() + head, b.begin() + head + n);
  uint64_t r = 0;
  for (int i = 7; i >= 0; --i) r = (r << 8) | (i < (int)n ? c.residue[i] : 0);
  if (r != c.res64) return fail(M2_E_INVAL, std::string("checkpoint Res64 does not match its residue: ") + path);
  return M2_OK;
}

int m2_checkpoint_save(m2_test* t, const char* path) {
  if (!t || !path) return fail(M2_E_INVAL, "null pointer");
  std::vector<uint8_t> r;
  int rc = test_sync(t);
  if (rc) return rc;
  if ((rc = test_obj_residue(t->obj, r))) return rc;
  std::vector<uint8_t> b;
  b.insert(b.end(), "M2CKPT01", "M2CKPT01" + 8);
  put_u32(b, t->obj->p);
  put_u32(b, (uint32_t)t->obj->mode);
  put_u64(b, t->obj->iter);
  put_u64(b, res64_of(r));
  put_u64(b, r.size());
  b.insert(b.end(), r.begin(), r.end());
  put_u64(b, fnv1a(b.data(), b.size()));
  // write-then-rename, so an interrupted save leaves the previous checkpoint
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return fail(M2_E_INVAL, "cannot write " + tmp);
  const bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
  if (fclose(f) != 0 || !ok) return fail(M2_E_INVAL, "cannot write " + tmp);
  if (rename(tmp.c_str(), path) != 0) return fail(M2_E_INVAL, std::string("cannot rename to ") + path);
  return M2_OK;
}

int m2_checkpoint_load(m2_test* t, const char* path) {
  if (!t || !path) return fail(M2_E_INVAL, "null pointer");
  CkptData c;
  int rc = ckpt_read(path, c);
  if (rc) return rc;
  if (c.p != t->obj->p || c.mode != t->obj->mode)
    return fail(M2_E_INVAL, "checkpoint is for p=" + std::to_string(c.p) + (c.mode ? " prp" : " ll") +
                                ", not this test");
  return m2_test_set_residue(t, c.residue.data(), c.residue.size(), c.iter);
}

int m2_checkpoint_info(const char* path, uint64_t* p, int* mode, uint64_t* iteration, uint64_t* res64) {
  if (!path) return fail(M2_E_INVAL, "null pointer");
  CkptData c;
  int rc = ckpt_read(path, c);
  if (rc) return rc;
  if (p) *p = c.p;
  if (mode) *mode = c.mode;
  if (iteration) *iteration = c.iter;
  if (res64) *res64 = c.res64;
  return M2_OK;
}

int m2_test_plan(m2_test* t, char* buf, size_t len) {
  if (!t || !buf || !len) return fail(M2_E_INVAL, "null pointer");
  snprintf(buf, len, "%s", describe_plan(t->obj->cls->P).c_str());
  return M2_OK;
}

}  // extern "C"

// ---- autotuner ----
//
// For one transform length and batch size, times each candidate plan (single
// workgroup or multi-pass with several centre sizes, outer carry-block
// widths and radices, then the graph length of the best) on `ntests` copies
// of a test and keeps the fastest. Every candidate computes the same
// residues; only speed differs.

namespace {

uint32_t tune_bucket(uint32_t n) {
  uint32_t b = 1;
  while (b < n && b < 256) b <<= 1;
  return b;
}

std::string opts_str(const m2_opts& o) {
  char buf[128];
  const bool wide = o.wg == (uint32_t)MP_THREADS_WIDE;
  snprintf(buf, sizeof buf, "LS=%u B=%u radix=%u graph=%u%s", o.local_stage_cap, o.carry_pairs,
           wide ? 4u : (o.max_radix ? o.max_radix : 8u), o.graph_iters ? o.graph_iters : GRAPH_ITERS,
           wide ? " W=1024" : "");
  return buf;
}

// Device µs per iteration of `ntests` tests of exponent p under plan options o.
int time_candidate(m2_ctx* ctx, uint64_t p, uint32_t ntests, const m2_opts& o, double* us, bool* created) {
  const int ln = transform_log2(p);
  const PlanOpts po = plan_opts(&o);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    *created = ctx->classes.find(std::make_pair(ln, po)) == ctx->classes.end();
  }
  Class* cls = nullptr;
  int rc = get_class(ctx, ln, po, &cls);
  if (rc) return rc;
  Group g;
  if ((rc = g.create(cls))) return rc;
  g.graph_iters = o.graph_iters ? o.graph_iters : GRAPH_ITERS;
  g.fused = ntests == 1 && !o.no_fuse;  // a single test runs as m2_test would
  std::vector<TestObj*> ts;
  auto cleanup = [&](int code) {
    g.destroy();
    for (TestObj* t : ts) test_obj_destroy(t);
    return code;
  };
  for (uint32_t i = 0; i < ntests; ++i) {
    TestObj* t = nullptr;
    if ((rc = test_obj_create(ctx, p, M2_MODE_PRP, po, g.stream, g.fused, &t))) return cleanup(rc);
    t->target = ~0ull;
    ts.push_back(t);
  }
  Events ev;
  if (ev.create(2) != hipSuccess) return cleanup(fail(M2_E_HIP, "event"));
  hipEvent_t a = ev[0], b = ev[1];
  auto run = [&](uint64_t k, float* ms) -> int {
    int r = g.prepare(ts, 0);
    if (r) return r;
    HIPCHK(hipEventRecord(a, g.stream));
    if ((r = g.advance(ts, k, 0))) return r;
    HIPCHK(hipEventRecord(b, g.stream));
    HIPCHK(hipEventSynchronize(b));
    HIPCHK(hipEventElapsedTime(ms, a, b));
    return g.sync();
  };
  float ms = 0;
  uint64_t k = 32;
  rc = run(16, &ms);
  for (int rep = 0; !rc && rep < 3; ++rep) {
    if ((rc = run(k, &ms))) break;
    if (ms >= 4.0f) break;
    k = std::min<uint64_t>(8192, (uint64_t)(k * std::max(2.0, 8.0 / std::max(ms, 0.01f))));
  }
  if (!rc) *us = ms * 1e3 / (double)k;
  return cleanup(rc);
}

int autotune(m2_ctx* ctx, uint64_t p, uint32_t ntests, m2_opts* out, std::string* report) {
  const int ln = transform_log2(p);
  const int L = ln - 1;
  const uint32_t bucket = tune_bucket(ntests ? ntests : 1);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tuned.find(std::make_pair(ln, bucket));
    if (it != ctx->tuned.end()) {
      *out = it->second;
      if (report) *report += "ln=" + std::to_string(ln) + " (cached): " + opts_str(*out) + "\n";
      return M2_OK;
    }
  }
  std::vector<m2_opts> cand;
  auto add = [&](uint32_t ls, uint32_t cp, uint32_t mr) {
    m2_opts o{};
    o.local_stage_cap = ls;
    o.carry_pairs = cp;
    o.max_radix = mr;
    cand.push_back(o);
  };
  if (L <= SMALL_LOG_H) {
    add(0, 0, 8);
    add(0, 0, 4);
  }
  for (int cl : {6, 8, 10}) {
    if (L < cl + 2) continue;
    if (L <= SMALL_LOG_H && cl < L - 6) continue;  // small lengths: only the shallowest outer passes
    add(2u << cl, 0, 8);
    add(2u << cl, 0, 4);
    add(2u << cl, 16, 8);
  }
  // one test whose centre grid leaves CUs with one workgroup: the 1024-thread centre
  const uint32_t cus = (uint32_t)ctx->prop.multiProcessorCount;
  if (ntests <= 1 && L >= CENTRE_LOG_C + 2 && (1u << (L - CENTRE_LOG_C - 1)) <= cus) {
    add(2u << CENTRE_LOG_C, 0, 8);
    cand.back().wg = MP_THREADS_WIDE;
  }
  const uint32_t nt = std::min<uint32_t>(std::max<uint32_t>(ntests, 1), 256);
  std::vector<std::pair<int, PlanOpts>> made;
  double best = 1e300;
  m2_opts bo{};
  int rc = M2_OK;
  auto try_one = [&](const m2_opts& o) -> int {
    double us = 0;
    bool created = false;
    int r = time_candidate(ctx, p, nt, o, &us, &created);
    if (created) made.push_back(std::make_pair(ln, plan_opts(&o)));
    if (r == M2_E_RANGE) {  // plan not valid for this exponent: skip
      if (report) *report += "  " + opts_str(o) + ": not applicable\n";
      return M2_OK;
    }
    if (r) return r;
    if (report) {
      char buf[64];
      snprintf(buf, sizeof buf, ": %.3f us/iter\n", us);
      *report += "  " + opts_str(o) + buf;
    }
    if (us < best) {
      best = us;
      bo = o;
    }
    return M2_OK;
  };
  if (report) *report += "ln=" + std::to_string(ln) + " tests=" + std::to_string(nt) + ":\n";
  for (const m2_opts& o : cand)
    if ((rc = try_one(o))) break;
  const bool multi = !rc && best < 1e300 && !(L <= SMALL_LOG_H && bo.local_stage_cap == 0);
  if (multi) {
    for (uint32_t gi : {8u, 32u}) {
      m2_opts o = bo;
      o.graph_iters = gi;
      if ((rc = try_one(o))) break;
    }
  }
  // free the tables of candidates that lost (and were made only for tuning)
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto& key : made) {
      if (key.second < plan_opts(&bo) || plan_opts(&bo) < key.second) {
        auto it = ctx->classes.find(key);
        if (it != ctx->classes.end()) {
          free_all(it->second->allocs);
          ctx->classes.erase(it);
        }
      }
    }
  }
  if (rc) return rc;
  if (best >= 1e300) return fail(M2_E_RANGE, "no applicable plan");
  if (report) {
    char buf[64];
    snprintf(buf, sizeof buf, " (%.3f us/iter)\n", best);
    *report += "  best: " + opts_str(bo) + buf;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->tuned[std::make_pair(ln, bucket)] = bo;
  *out = bo;
  return M2_OK;
}

}  // namespace

extern "C" {

int m2_autotune(m2_ctx* ctx, uint64_t p, uint32_t ntests, m2_opts* out, char* report, size_t report_len) {
  if (!ctx || !out) return fail(M2_E_INVAL, "null pointer");
  if (p < 3 || p > 0xFFFFFFFFull || transform_log2(p) > 27) return fail(M2_E_RANGE, "exponent out of range");
  if ((p & 1u) == 0) return fail(M2_E_RANGE, "exponent must be odd");
  HIPCHK(hipSetDevice(ctx->device));
  std::string rep;
  int rc = autotune(ctx, p, ntests, out, report ? &rep : nullptr);
  if (report && report_len) snprintf(report, report_len, "%s", rep.c_str());
  return rc;
}

// ---- batches ----

int m2_batch_create(m2_ctx* ctx, size_t count, const uint64_t* ps, const int* modes, const uint64_t* iters,
                    const m2_opts* opts, m2_batch** out) {
  if (!ctx || !out || (count && (!ps || !modes))) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<m2_batch> b(new m2_batch);
  b->ctx = ctx;
  // plan options per transform length: the caller's, or tuned for this batch
  std::map<int, m2_opts> per_ln;
  std::map<int, uint32_t> count_ln;
  for (size_t i = 0; i < count; ++i)
    if (ps[i] >= 3 && ps[i] <= 0xFFFFFFFFull) ++count_ln[transform_log2(ps[i])];
  for (auto& kv : count_ln) {
    m2_opts o{};
    if (opts) o = *opts;
    if (opts && opts->autotune && kv.first <= 27) {
      uint64_t rep_p = 0;
      for (size_t i = 0; i < count && !rep_p; ++i)
        if (ps[i] >= 3 && ps[i] <= 0xFFFFFFFFull && transform_log2(ps[i]) == kv.first && (ps[i] & 1)) rep_p = ps[i];
      if (rep_p) {
        int rc = autotune(ctx, rep_p, kv.second, &o, nullptr);
        if (rc) return rc;
      }
    }
    per_ln[kv.first] = o;
  }
  auto bail = [&](int code) {
    for (auto& kv : b->groups) kv.second.destroy();
    for (TestObj* t : b->tests) test_obj_destroy(t);
    return code;
  };
  for (size_t i = 0; i < count; ++i) {
    const int ln = (ps[i] >= 3 && ps[i] <= 0xFFFFFFFFull) ? transform_log2(ps[i]) : 0;
    if (ln < 3 || ln > 27) return bail(fail(M2_E_RANGE, "exponent out of range at index " + std::to_string(i)));
    const m2_opts& o = per_ln[ln];
    Class* cls = nullptr;
    int rc = get_class(ctx, ln, plan_opts(&o), &cls);
    if (rc) return bail(rc);
    auto it = b->groups.find(cls);
    if (it == b->groups.end()) {
      Group g;
      if ((rc = g.create(cls))) return bail(rc);
      if (o.graph_iters) g.graph_iters = o.graph_iters;
      it = b->groups.emplace(cls, g).first;
    }
    TestObj* t = nullptr;
    if ((rc = test_obj_create(ctx, ps[i], modes[i], plan_opts(&o), it->second.stream, false, &t))) return bail(rc);
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
  // Enqueue in chunks, round-robin over the lengths, so that every length's
  // stream has work queued early instead of one stream taking the whole run.
  constexpr uint64_t CHUNK = 8192;
  uint64_t left = max_iters ? max_iters : ~0ull;
  while (left) {
    const uint64_t k = std::min(left, CHUNK);
    bool any = false;
    for (auto& kv : by) {
      for (TestObj* t : kv.second) any |= t->iter < t->target;
      int rc = b->groups[kv.first].advance(kv.second, k, 0);
      if (rc) return rc;
    }
    if (!any) break;
    left -= k;
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
//
// Continuous batching: each transform length has a lane (a group on its own
// stream) that runs one slice of its members at a time. A lane whose slice
// has finished retires its finished tests and starts the next slice with
// whatever members it has then, including tests admitted meanwhile, while
// the other lanes keep running.

struct m2_server {
  m2_ctx* ctx = nullptr;
  uint64_t budget = 0, in_use = 0;
  uint32_t slice = 1000;
  m2_opts opts{};
  hipStream_t setup = nullptr;  // uploads of new tests' initial state
  std::mutex mu;                // guards `queue` and `done` (submit/poll may come from other threads)
  struct Req {
    uint64_t id, p, iters;
    int mode;
  };
  std::deque<Req> queue;
  struct Member {
    uint64_t id;
    TestObj* t;
    uint64_t bytes;
  };
  struct Lane {
    Group g;
    std::vector<Member> members;
    hipEvent_t ev0 = nullptr, ev = nullptr;  // around the slice in flight
    bool inflight = false;
    uint64_t k = 0;          // iterations per member in the slice in flight
    double us_per_iter = 0;  // measured device time of one iteration of all members
  };
  std::map<Class*, Lane> lanes;
  uint64_t nactive = 0;
  struct Done {
    uint64_t id, res64;
    int verdict;
  };
  std::deque<Done> done;
  uint64_t next_id = 1;
};

namespace {

int server_admit(m2_server* s, bool* progressed) {
  for (;;) {
    m2_server::Req rq;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->queue.empty()) return M2_OK;
      rq = s->queue.front();
    }
    uint64_t fp = 0;
    m2_test_footprint(rq.p, &fp);
    if (s->in_use + fp > s->budget && s->nactive > 0) return M2_OK;
    m2_opts o = s->opts;
    if (s->opts.autotune) {
      // tuned for a lane of a few tests of this length
      int rc = autotune(s->ctx, rq.p, 8, &o, nullptr);
      if (rc) return rc;
    }
    Class* cls = nullptr;
    int rc = get_class(s->ctx, transform_log2(rq.p), plan_opts(&o), &cls);
    if (rc) return rc;
    auto it = s->lanes.find(cls);
    if (it == s->lanes.end()) {
      m2_server::Lane lane;
      if ((rc = lane.g.create(cls))) return rc;
      if (o.graph_iters) lane.g.graph_iters = o.graph_iters;
      if (hipEventCreate(&lane.ev0) != hipSuccess || hipEventCreate(&lane.ev) != hipSuccess) {
        if (lane.ev0) (void)hipEventDestroy(lane.ev0);
        lane.g.destroy();
        return fail(M2_E_HIP, "event");
      }
      it = s->lanes.emplace(cls, lane).first;
    }
    TestObj* t = nullptr;
    if ((rc = test_obj_create(s->ctx, rq.p, rq.mode, plan_opts(&o), s->setup, false, &t))) return rc;
    if (rq.iters) t->target = std::min<uint64_t>(t->target, rq.iters);
    it->second.members.push_back({rq.id, t, fp});
    s->in_use += fp;
    ++s->nactive;
    *progressed = true;
    std::lock_guard<std::mutex> lk(s->mu);
    s->queue.pop_front();
  }
}

// Lane whose slice is done (or idle): retire finished members, start the next slice.
// Device time a lane's slice may take: a slice is at most `slice` iterations
// and at most about this long, so that no lane holds a hardware queue for
// long while others (short tests) wait behind it.
constexpr double SLICE_TARGET_US = 4000.0;

int server_service_lane(m2_server* s, m2_server::Lane& lane, bool* progressed) {
  if (lane.inflight) {
    hipError_t q = hipEventQuery(lane.ev);
    if (q == hipErrorNotReady) return M2_OK;
    if (q != hipSuccess) return fail(M2_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
    lane.inflight = false;
    *progressed = true;
    float ms = 0;
    if (lane.k && hipEventElapsedTime(&ms, lane.ev0, lane.ev) == hipSuccess && ms > 0) {
      const double u = ms * 1e3 / (double)lane.k;
      lane.us_per_iter = lane.us_per_iter > 0 ? 0.5 * (lane.us_per_iter + u) : u;
    }
  }
  int rc = lane.g.sync();
  if (rc) return rc;
  std::vector<m2_server::Member> keep;
  for (auto& m : lane.members) {
    if (m.t->iter < m.t->target) {
      keep.push_back(m);
      continue;
    }
    std::vector<uint8_t> r;
    if ((rc = test_obj_residue(m.t, r))) return rc;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->done.push_back({m.id, res64_of(r), verdict_of(m.t, r)});
    }
    s->in_use -= m.bytes;
    --s->nactive;
    if ((rc = lane.g.forget(m.t))) return rc;
    test_obj_destroy(m.t);
    *progressed = true;
  }
  lane.members.swap(keep);
  if (lane.members.empty()) return M2_OK;
  std::vector<TestObj*> ts;
  for (auto& m : lane.members) ts.push_back(m.t);
  // first slice of a lane: a short probe; then sized by the measured time
  uint64_t k = s->slice;
  if (lane.us_per_iter <= 0) k = std::min<uint64_t>(k, 16);
  else k = std::max<uint64_t>(1, std::min<uint64_t>(k, (uint64_t)(SLICE_TARGET_US / lane.us_per_iter)));
  lane.k = k;
  HIPCHK(hipEventRecord(lane.ev0, lane.g.stream));
  if ((rc = lane.g.advance(ts, k, 0))) return rc;
  HIPCHK(hipEventRecord(lane.ev, lane.g.stream));
  lane.inflight = true;
  return M2_OK;
}

}  // namespace

extern "C" {

int m2_server_create(m2_ctx* ctx, uint64_t hbm_budget_bytes, uint32_t slice_iters, const m2_opts* opts,
                     m2_server** out) {
  if (!ctx || !out) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<m2_server> s(new m2_server);
  s->ctx = ctx;
  s->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(0.7 * (double)ctx->prop.totalGlobalMem);
  s->slice = slice_iters ? slice_iters : 1000;
  if (opts) s->opts = *opts;
  HIPCHK(hipStreamCreateWithFlags(&s->setup, hipStreamNonBlocking));
  *out = s.release();
  return M2_OK;
}

int m2_server_destroy(m2_server* s) {
  if (!s) return M2_OK;
  (void)hipSetDevice(s->ctx->device);
  for (auto& kv : s->lanes) {
    kv.second.g.destroy();
    for (auto& m : kv.second.members) test_obj_destroy(m.t);
    if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
    if (kv.second.ev0) (void)hipEventDestroy(kv.second.ev0);
  }
  if (s->setup) (void)hipStreamDestroy(s->setup);
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
  // Wait up to about a millisecond for something to happen, so that the
  // caller can submit arrivals in between.
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    bool progressed = false;
    int rc = server_admit(s, &progressed);
    if (rc) return rc;
    for (auto& kv : s->lanes)
      if ((rc = server_service_lane(s, kv.second, &progressed))) return rc;
    bool inflight = false;
    for (auto& kv : s->lanes) inflight |= kv.second.inflight;
    if (progressed || !inflight) break;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(1000)) break;
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  if (busy) {
    std::lock_guard<std::mutex> lk(s->mu);
    *busy = (s->nactive > 0 || !s->queue.empty()) ? 1 : 0;
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
  if (active) *active = s->nactive;
  if (bytes_in_use) *bytes_in_use = s->in_use;
  return M2_OK;
}

}  // extern "C"
