// C ABI of libm2.so (declared in include/mersenne2.h).
#include <stdio.h>
#include <string.h>

#include <algorithm>
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

#define HIPCHK(call)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return fail(M2_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr uint32_t SMALL_CHUNK = 2048;  // iterations per k_small launch
}  // namespace

struct m2_ctx {
  int device = 0;
  hipDeviceProp_t prop{};
};

struct m2_test {
  m2_ctx* ctx = nullptr;
  uint32_t p = 0;
  int mode = 1;
  uint64_t iter = 0;
  DevPlan P;
  HostTables T;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint32_t graph_iters = 0;
  uint64_t bytes = 0;
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
  auto* ctx = new m2_ctx;
  ctx->device = device;
  hipError_t e = hipGetDeviceProperties(&ctx->prop, device);
  if (e != hipSuccess) {
    delete ctx;
    return fail(M2_E_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  }
  *out = ctx;
  return M2_OK;
}

int m2_close(m2_ctx* ctx) {
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
  const uint64_t n = uint64_t(1) << transform_log2(p), h = n / 2;
  // spectrum 24 B/element, digits 16 B/element, step twiddles < h x 24 B,
  // centre twiddles h/2 x 24 B, pending carries <= 16 B per 4 elements
  *bytes = 24 * h + 16 * h + 24 * h + 12 * h + 4 * h + 4096;
  (void)n;
  return M2_OK;
}

static int alloc_copy(m2_test* t, void** dst, const void* src, size_t bytes) {
  void* d = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&d, bytes);
  if (e != hipSuccess) return fail(M2_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  t->allocs.push_back(d);
  t->bytes += bytes;
  if (src) {
    e = hipMemcpy(d, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(M2_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  *dst = d;
  return M2_OK;
}

int m2_test_destroy(m2_test* t) {
  if (!t) return M2_OK;
  (void)hipSetDevice(t->ctx->device);
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  if (t->gexec) (void)hipGraphExecDestroy(t->gexec);
  if (t->graph) (void)hipGraphDestroy(t->graph);
  for (void* a : t->allocs) (void)hipFree(a);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
  return M2_OK;
}

// digits: n digits in natural order
static int upload_state(m2_test* t, const int64_t* digits) {
  const uint64_t h = t->P.h;
  std::vector<int64_t> slots(2 * h);
  for (uint64_t e = 0; e < h; ++e) {
    const uint64_t s = digit_slot(t->P, e);
    slots[2 * s] = digits[2 * e];
    slots[2 * s + 1] = digits[2 * e + 1];
  }
  HIPCHK(hipMemcpyAsync(t->P.small ? t->P.z61 : t->P.zd, slots.data(), h * sizeof(e61), hipMemcpyHostToDevice,
                        t->stream));
  if (!t->P.small) {
    std::vector<Carry128> zero(t->P.nseg, Carry128{0, 0});
    HIPCHK(hipMemcpyAsync(t->P.G, zero.data(), zero.size() * sizeof(Carry128), hipMemcpyHostToDevice, t->stream));
    HIPCHK(hipStreamSynchronize(t->stream));
  }
  HIPCHK(hipStreamSynchronize(t->stream));
  return M2_OK;
}

int m2_test_create(m2_ctx* ctx, uint64_t p, int mode, const m2_opts* opts, m2_test** out) {
  if (!ctx || !out) return fail(M2_E_INVAL, "null pointer");
  if (mode != M2_MODE_LL && mode != M2_MODE_PRP) return fail(M2_E_INVAL, "mode must be LL (0) or PRP (1)");
  if (p < 3 || p > 0xFFFFFFFFull) return fail(M2_E_RANGE, "exponent out of range (3 <= p < 2^32)");
  HIPCHK(hipSetDevice(ctx->device));
  std::unique_ptr<m2_test> t(new m2_test);
  t->ctx = ctx;
  t->p = (uint32_t)p;
  t->mode = mode;
  std::string err;
  if (!build_plan((uint32_t)p, mode, opts ? opts->local_stage_cap : 0u, t->P, t->T, err)) return fail(M2_E_RANGE, err);
  t->graph_iters = (opts && opts->graph_iters) ? opts->graph_iters : 16;
  HIPCHK(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
  const uint64_t h = t->P.h, n = 2 * h;
  int rc;
  void* d;
  if ((rc = alloc_copy(t.get(), &d, nullptr, h * sizeof(e61)))) return rc;
  t->P.z61 = (e61*)d;
  if ((rc = alloc_copy(t.get(), &d, nullptr, h * sizeof(e31)))) return rc;
  t->P.z31 = (e31*)d;
  if ((rc = alloc_copy(t.get(), &d, t->T.tw61.data(), t->T.tw61.size() * sizeof(e61)))) return rc;
  t->P.tw61 = (const e61*)d;
  if ((rc = alloc_copy(t.get(), &d, t->T.tw31.data(), t->T.tw31.size() * sizeof(e31)))) return rc;
  t->P.tw31 = (const e31*)d;
  if ((rc = alloc_copy(t.get(), &d, t->T.sq61.data(), t->T.sq61.size() * sizeof(e61)))) return rc;
  t->P.sq61 = (const e61*)d;
  if ((rc = alloc_copy(t.get(), &d, t->T.sq31.data(), t->T.sq31.size() * sizeof(e31)))) return rc;
  t->P.sq31 = (const e31*)d;
  if (!t->P.small) {
    if ((rc = alloc_copy(t.get(), &d, nullptr, t->P.nseg * sizeof(Carry128)))) return rc;
    t->P.G = (Carry128*)d;
    if ((rc = alloc_copy(t.get(), &d, nullptr, h * sizeof(e61)))) return rc;
    t->P.zd = (e61*)d;
  }
  // the host copies of the large twiddle tables are no longer needed
  std::vector<e61>().swap(t->T.tw61);
  std::vector<e31>().swap(t->T.tw31);
  std::vector<e61>().swap(t->T.sq61);
  std::vector<e31>().swap(t->T.sq31);
  // initial residue: 3 (PRP) or 4 (LL) in digit 0 (mersenne2_gpu.cpp:3908)
  std::vector<int64_t> digits(n, 0);
  digits[0] = (mode == M2_MODE_LL) ? 4 : 3;
  if ((rc = upload_state(t.get(), digits.data()))) return rc;
  *out = t.release();
  return M2_OK;
}

static int enqueue(m2_test* t, uint64_t iters) {
  HIPCHK(hipSetDevice(t->ctx->device));
  if (t->P.small) {
    while (iters) {
      const uint32_t k = (uint32_t)std::min<uint64_t>(iters, SMALL_CHUNK);
      HIPCHK(launch_small(t->P, k, t->stream));
      iters -= k;
      t->iter += k;
    }
    return M2_OK;
  }
  const uint32_t G = t->graph_iters;
  if (iters >= G && G > 1) {
    if (!t->gexec) {
      HIPCHK(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
      for (uint32_t i = 0; i < G; ++i) {
        hipError_t e = launch_iteration(t->P, t->stream);
        if (e != hipSuccess) {
          hipGraph_t g;
          (void)hipStreamEndCapture(t->stream, &g);
          return fail(M2_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        }
      }
      HIPCHK(hipStreamEndCapture(t->stream, &t->graph));
      HIPCHK(hipGraphInstantiate(&t->gexec, t->graph, nullptr, nullptr, 0));
    }
    while (iters >= G) {
      HIPCHK(hipGraphLaunch(t->gexec, t->stream));
      iters -= G;
      t->iter += G;
    }
  }
  while (iters) {
    HIPCHK(launch_iteration(t->P, t->stream));
    --iters;
    ++t->iter;
  }
  return M2_OK;
}

int m2_test_run(m2_test* t, uint64_t iters) {
  if (!t) return fail(M2_E_INVAL, "null test");
  return enqueue(t, iters);
}

int m2_test_sync(m2_test* t) {
  if (!t) return fail(M2_E_INVAL, "null test");
  HIPCHK(hipStreamSynchronize(t->stream));
  return M2_OK;
}

int m2_test_iteration(m2_test* t, uint64_t* iter) {
  if (!t || !iter) return fail(M2_E_INVAL, "null pointer");
  *iter = t->iter;
  return M2_OK;
}

int m2_test_total_iters(m2_test* t, uint64_t* iters) {
  if (!t || !iters) return fail(M2_E_INVAL, "null pointer");
  *iters = (t->mode == M2_MODE_LL) ? (t->p >= 2 ? t->p - 2 : 0) : t->p;
  return M2_OK;
}

static int read_residue(m2_test* t, std::vector<uint8_t>& r) {
  HIPCHK(hipSetDevice(t->ctx->device));
  HIPCHK(hipStreamSynchronize(t->stream));
  const uint64_t h = t->P.h, n = 2 * h;
  std::vector<int64_t> slots(n), digits(n);
  HIPCHK(hipMemcpy(slots.data(), t->P.small ? t->P.z61 : t->P.zd, h * sizeof(e61), hipMemcpyDeviceToHost));
  for (uint64_t e = 0; e < h; ++e) {
    const uint64_t s = digit_slot(t->P, e);
    digits[2 * e] = slots[2 * s];
    digits[2 * e + 1] = slots[2 * s + 1];
  }
  std::vector<Carry128> G;
  std::vector<uint64_t> at;
  if (!t->P.small) {
    G.resize(t->P.nseg);
    HIPCHK(hipMemcpy(G.data(), t->P.G, G.size() * sizeof(Carry128), hipMemcpyDeviceToHost));
    at.resize(G.size());
    for (uint64_t g = 0; g < G.size(); ++g) at[g] = carry_target_digit(t->P, g);
  }
  r = residue_from_state(t->p, t->T.bitpos, digits.data(), n, G.data(), at.data(), G.size());
  return M2_OK;
}

int m2_test_residue(m2_test* t, uint8_t* buf, size_t len) {
  if (!t || !buf) return fail(M2_E_INVAL, "null pointer");
  const size_t need = (t->p + 7) / 8;
  if (len < need) return fail(M2_E_INVAL, "residue buffer too small: need " + std::to_string(need) + " bytes");
  std::vector<uint8_t> r;
  int rc = read_residue(t, r);
  if (rc) return rc;
  memset(buf, 0, len);
  memcpy(buf, r.data(), r.size());
  return M2_OK;
}

int m2_test_res64(m2_test* t, uint64_t* res64) {
  if (!t || !res64) return fail(M2_E_INVAL, "null pointer");
  std::vector<uint8_t> r;
  int rc = read_residue(t, r);
  if (rc) return rc;
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | (i < (int)r.size() ? r[i] : 0);
  *res64 = v;
  return M2_OK;
}

int m2_test_is_prime(m2_test* t, int* verdict) {
  if (!t || !verdict) return fail(M2_E_INVAL, "null pointer");
  std::vector<uint8_t> r;
  int rc = read_residue(t, r);
  if (rc) return rc;
  // expected value: 0 (LL) or 9 mod M_p (PRP)
  std::vector<uint8_t> want(r.size(), 0);
  if (t->mode == M2_MODE_PRP) {
    uint64_t nine = 9;
    if (t->p < 64) nine %= ((uint64_t(1) << t->p) - 1);
    for (size_t i = 0; i < want.size() && i < 8; ++i) want[i] = (uint8_t)(nine >> (8 * i));
  }
  *verdict = (r == want) ? 1 : 0;
  return M2_OK;
}

int m2_test_set_residue(m2_test* t, const uint8_t* buf, size_t len, uint64_t iteration) {
  if (!t || !buf) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  HIPCHK(hipStreamSynchronize(t->stream));
  std::vector<int64_t> digits(2ull * t->P.h);
  digits_from_residue(t->p, t->T.bitpos, t->T.dig, buf, len, digits.data());
  int rc = upload_state(t, digits.data());
  if (rc) return rc;
  t->iter = iteration;
  return M2_OK;
}

int m2_test_time(m2_test* t, uint64_t iters, double* ms) {
  if (!t || !ms) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, t->stream));
  int rc = enqueue(t, iters);
  if (rc) return rc;
  HIPCHK(hipEventRecord(b, t->stream));
  HIPCHK(hipEventSynchronize(b));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = f;
  return M2_OK;
}

int m2_test_kernel_times(m2_test* t, uint64_t iters, double* ms, int max_kernels, int* count, char* names_buf,
                         size_t names_len) {
  if (!t || !ms || !count) return fail(M2_E_INVAL, "null pointer");
  HIPCHK(hipSetDevice(t->ctx->device));
  constexpr int MAXE = 8;
  hipEvent_t starts[MAXE], stops[MAXE];
  for (int k = 0; k < MAXE; ++k) {
    HIPCHK(hipEventCreate(&starts[k]));
    HIPCHK(hipEventCreate(&stops[k]));
  }
  const char* names[MAXE] = {nullptr};
  int nk = 0;
  std::vector<double> acc(MAXE, 0.0);
  for (uint64_t it = 0; it < iters; ++it) {
    if (t->P.small) {
      HIPCHK(launch_small_events(t->P, 1, t->stream, starts[0], stops[0]));
      nk = 1;
      names[0] = "k_small";
    } else {
      HIPCHK(launch_iteration_events(t->P, t->stream, starts, stops, MAXE, &nk, names));
    }
    ++t->iter;
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
  snprintf(buf, len, "%s", describe_plan(t->P).c_str());
  return M2_OK;
}

}  // extern "C"
