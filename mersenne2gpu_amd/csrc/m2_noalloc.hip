// An entry point that allocates nothing: m2_square_noalloc advances a residue
// by a number of iterations with every device buffer it needs (the test
// descriptors, the transform's twiddle tables and the digits) inside one
// fixed file-scope device array, `scratch`. It runs the single-workgroup
// kernel (k_small, the same one m2_test and m2_batch use for these lengths),
// so it takes exponents whose transform length is at most 2^12.
//
// Set-up (m2_open, through noalloc_init): once per device, with the device
// current, the array's device address is taken with hipGetSymbolAddress and
// kept per device (in host code the array's name is not a device address);
// a stream is created; the twiddle tables of every length 2^3..2^12 are built
// on the host once per process. A call then only copies tables, descriptors
// and digits into the array, launches and copies the digits back: no
// hipMalloc, hipFree or other allocation. Calls on one device are serialised
// by a lock; the array keeps the last call's contents, and every call
// rewrites all of it that it reads.
#include <string.h>

#include <algorithm>
#include <mutex>

#include "m2_digits.h"
#include "m2_internal.h"
#include "mersenne2.h"

namespace m2 {

// Enough for the largest length (n = 2^12: 2 descriptors, 2047 step twiddles
// and 1024 centre twiddles in both fields, 2048 digit pairs: 106 KiB).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(M2_NOALLOC_SCRATCH_KIB) * 1024;
constexpr int NOALLOC_MAX_LN = 12;
static_assert(NOALLOC_MAX_LN <= SMALL_LOG_H + 1, "every length is a single-workgroup plan");
constexpr int NOALLOC_MAX_DEVICES = 16;
constexpr uint32_t NOALLOC_CHUNK = 4096;  // iterations per launch (as m2_test's single-workgroup plan)

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

struct SmallClass {
  bool ok = false;
  DevPlan P;  // pointers filled per call (offsets below)
  ClassTables T;
  size_t off_desc, off_tw61, off_tw31, off_sq61, off_sq31, off_z, end;
};

struct DeviceState {
  std::mutex mu;
  bool ready = false;
  unsigned char* ptr = nullptr;  // device address of `scratch` on this device
  hipStream_t stream = nullptr;
  int64_t digits[2u << NOALLOC_MAX_LN];
  uint64_t words[(1u << NOALLOC_MAX_LN) + 8];  // residue accumulator: p < 44 n bits
  TestDesc desc[2];
};

std::mutex g_init_mu;
SmallClass g_cls[NOALLOC_MAX_LN + 1];
bool g_cls_built = false;
DeviceState g_dev[NOALLOC_MAX_DEVICES];

size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

bool build_small_classes(std::string& err) {
  for (int ln = 3; ln <= NOALLOC_MAX_LN; ++ln) {
    SmallClass& c = g_cls[ln];
    if (!build_class_plan(ln, PlanOpts{}, c.P, c.T, err)) return false;
    if (!c.P.small) {
      err = "no-alloc entry point: length 2^" + std::to_string(ln) + " is not a single-workgroup plan";
      return false;
    }
    size_t o = 0;
    c.off_desc = o;
    o = up256(o + 2 * sizeof(TestDesc));
    c.off_tw61 = o;
    o = up256(o + c.T.tw61.size() * sizeof(e61));
    c.off_tw31 = o;
    o = up256(o + c.T.tw31.size() * sizeof(e31));
    c.off_sq61 = o;
    o = up256(o + c.T.sq61.size() * sizeof(e61));
    c.off_sq31 = o;
    o = up256(o + c.T.sq31.size() * sizeof(e31));
    c.off_z = o;
    o = up256(o + size_t(c.P.h) * sizeof(e61));
    c.end = o;
    if (c.end > NOALLOC_SCRATCH_BYTES) {
      err = "no-alloc entry point: length 2^" + std::to_string(ln) + " does not fit the scratch array";
      return false;
    }
    c.ok = true;
  }
  return true;
}

// The walk constants of exponent p at length 2^ln (the values
// build_test_tables derives, mersenne2_gpu.cpp:3921-3944).
DigitGen digit_gen(uint32_t p, int ln) {
  const uint64_t n = uint64_t(1) << ln;
  const uint64_t lr61 = ((uint64_t(1) << 60) >> ln) % 61u;
  const uint64_t lr31 = ((uint64_t(1) << 30) >> ln) % 31u;
  DigitGen D{};
  D.p = p;
  D.nmask = (uint32_t)(n - 1);
  D.s = (uint32_t)(p & (n - 1));
  D.q = p >> ln;
  D.lr61 = (uint32_t)lr61;
  D.lr31 = (uint32_t)lr31;
  D.d61 = (uint32_t)((lr61 * D.s) % 61u);
  D.d31 = (uint32_t)((lr31 * D.s) % 31u);
  D.k61 = (uint32_t)((ln + 1) % 61);
  D.k31 = (uint32_t)((ln + 1) % 31);
  return D;
}

uint64_t bitpos(uint32_t p, int ln, uint64_t j) { return j ? (((uint64_t)p * j - 1) >> ln) + 1 : 0; }

// Bits [b, b + w) of the little-endian residue r (w < 64; bits past len are 0).
uint64_t read_bits(const uint8_t* r, size_t len, uint64_t b, uint32_t w) {
  uint64_t v = 0;
  for (uint32_t k = 0; k < w;) {
    const uint64_t bb = b + k, byte = bb / 8;
    const uint32_t sh = (uint32_t)(bb % 8), take = std::min<uint32_t>(8 - sh, w - k);
    const uint64_t bits = byte < len ? (uint64_t)(r[byte] >> sh) & ((1u << take) - 1u) : 0;
    v |= bits << k;
    k += take;
  }
  return v;
}

// Canonical residue (sum d_j 2^B_j mod 2^p - 1, 2^p - 1 = 0) of n signed
// digits into out (ceil(p/8) bytes), with the accumulator `words` (at least
// p/64 + 4 words).
void digits_to_residue(uint32_t p, int ln, const int64_t* d, uint64_t* words, uint8_t* out) {
  const uint64_t n = uint64_t(1) << ln;
  const size_t nw = p / 64 + 4;
  memset(words, 0, nw * sizeof(uint64_t));
  // two's-complement sum over nw words (|sum| < 2^(p + 70))
  for (uint64_t j = 0; j < n; ++j) {
    if (!d[j]) continue;
    const uint64_t b = bitpos(p, ln, j);
    size_t i = b / 64;
    const unsigned sh = b % 64;
    const __int128 v = (__int128)d[j] << sh;  // |d| < 2^62: fits
    const uint64_t fill = v < 0 ? ~0ull : 0ull;
    const uint64_t parts[2] = {(uint64_t)v, (uint64_t)(v >> 64)};
    unsigned __int128 c = 0;
    for (size_t k = i; k < nw; ++k) {
      const uint64_t add = k - i < 2 ? parts[k - i] : fill;
      const unsigned __int128 s = (unsigned __int128)words[k] + add + c;
      words[k] = (uint64_t)s;
      c = s >> 64;
    }
  }
  // fold mod 2^p - 1: the value is lo + hi 2^p with hi a small signed number
  auto fold = [&]() {
    const unsigned pb = p % 64;
    const size_t pw = p / 64;
    // hi = value >> p (arithmetic), at most two words of magnitude
    __int128 hi = 0;
    {
      const uint64_t w0 = words[pw], w1 = words[pw + 1], w2 = words[pw + 2];
      const unsigned __int128 lo2 = ((unsigned __int128)w1 << 64) | w0;
      unsigned __int128 x = pb ? (lo2 >> pb) | ((unsigned __int128)w2 << (128 - pb)) : lo2;
      hi = (__int128)x;
    }
    for (size_t k = pw + 1; k < nw; ++k) words[k] = 0;
    if (pb) words[pw] &= (1ull << pb) - 1ull;
    else words[pw] = 0;
    // lo + hi
    const uint64_t fill = hi < 0 ? ~0ull : 0ull;
    unsigned __int128 c = 0;
    for (size_t k = 0; k < nw; ++k) {
      const uint64_t add = k == 0 ? (uint64_t)hi : (k == 1 ? (uint64_t)(hi >> 64) : fill);
      const unsigned __int128 s = (unsigned __int128)words[k] + add + c;
      words[k] = (uint64_t)s;
      c = s >> 64;
    }
  };
  for (int r = 0; r < 4; ++r) fold();
  // a negative value (top word set): add 2^p - 1 until it is not
  while ((int64_t)words[nw - 1] < 0) {
    unsigned __int128 c = 0;
    for (size_t k = 0; k < nw; ++k) {
      uint64_t add = 0;
      if (k < p / 64) add = ~0ull;
      else if (k == p / 64) add = (p % 64) ? ((1ull << (p % 64)) - 1ull) : 0ull;
      const unsigned __int128 s = (unsigned __int128)words[k] + add + c;
      words[k] = (uint64_t)s;
      c = s >> 64;
    }
    fold();
  }
  // 2^p - 1 itself is 0
  bool all_ones = true;
  for (uint32_t b = 0; b < p && all_ones; b += 64) {
    const uint32_t k = std::min<uint32_t>(64, p - b);
    const uint64_t m = k == 64 ? ~0ull : ((1ull << k) - 1ull);
    all_ones = (words[b / 64] & m) == m;
  }
  if (all_ones) memset(words, 0, nw * sizeof(uint64_t));
  const size_t nbytes = (p + 7) / 8;
  for (size_t i = 0; i < nbytes; ++i) out[i] = (uint8_t)(words[i / 8] >> (8 * (i % 8)));
}

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return M2_OK;  // the entry point refuses such devices
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (!g_cls_built) {
    std::string err;
    if (!build_small_classes(err)) return test_error(M2_E_STATE, err);
    g_cls_built = true;
  }
  DeviceState& ds = g_dev[device];
  if (ds.ready) return M2_OK;
  // with `device` current (m2_open sets it): the array's address on it
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return test_error(M2_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  e = hipStreamCreateWithFlags(&ds.stream, hipStreamNonBlocking);
  if (e != hipSuccess) return test_error(M2_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  ds.ptr = (unsigned char*)p;
  ds.ready = true;
  return M2_OK;
}

}  // namespace m2

using namespace m2;

#define NA_CHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return test_error(M2_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

extern "C" int m2_square_noalloc(m2_ctx* ctx, uint64_t p, int mode, const uint8_t* in, size_t in_len,
                                 uint64_t iters, uint8_t* out, size_t out_len) {
  if (!ctx || !in || !out) return test_error(M2_E_INVAL, "null pointer");
  if (mode != M2_MODE_LL && mode != M2_MODE_PRP) return test_error(M2_E_INVAL, "mode must be LL (0) or PRP (1)");
  if (p < 3 || p > 0xFFFFFFFFull || (p & 1u) == 0) return test_error(M2_E_RANGE, "exponent must be odd and >= 3");
  const int ln = transform_log2(p);
  if (ln > NOALLOC_MAX_LN)
    return test_error(M2_E_RANGE, "m2_square_noalloc: exponent " + std::to_string(p) + " needs n = 2^" +
                                      std::to_string(ln) + "; the scratch array (" +
                                      std::to_string(M2_NOALLOC_SCRATCH_KIB) + " KiB) holds n <= 2^" +
                                      std::to_string(NOALLOC_MAX_LN) + " (p <= 159739)");
  const size_t nbytes = (p + 7) / 8;
  if (in_len < nbytes || out_len < nbytes) return test_error(M2_E_INVAL, "residue buffers shorter than ceil(p/8)");
  const int device = ctx_device(ctx);
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return test_error(M2_E_DEVICE, "device index out of range");
  NA_CHK(hipSetDevice(device));
  DeviceState& ds = g_dev[device];
  if (!ds.ready) {
    // set-up did not complete in m2_open: once more, with the device current
    const int rc = noalloc_init(device);
    if (rc) return rc;
  }
  std::lock_guard<std::mutex> lk(ds.mu);
  const SmallClass& c = g_cls[ln];
  const uint64_t n = uint64_t(1) << ln, h = n / 2;
  const uint32_t pp = (uint32_t)p;
  // digits of the input residue (natural order: the single-workgroup plan's layout)
  for (uint64_t j = 0; j < n; ++j) {
    const uint64_t b = bitpos(pp, ln, j);
    ds.digits[j] = (int64_t)read_bits(in, in_len, b, (uint32_t)(bitpos(pp, ln, j + 1) - b));
  }
  unsigned char* base = ds.ptr;
  DevPlan P = c.P;
  P.tw61 = (const e61*)(base + c.off_tw61);
  P.tw31 = (const e31*)(base + c.off_tw31);
  P.sq61 = (const e61*)(base + c.off_sq61);
  P.sq31 = (const e31*)(base + c.off_sq31);
  e61* z = (e61*)(base + c.off_z);
  TestDesc* dd = (TestDesc*)(base + c.off_desc);
  const DigitGen D = digit_gen(pp, ln);
  for (int k = 0; k < 2; ++k) {
    ds.desc[k] = TestDesc{};
    ds.desc[k].z61 = z;
    ds.desc[k].dgen = D;
    ds.desc[k].mode = mode;
  }
  ds.desc[0].iters = NOALLOC_CHUNK;
  ds.desc[1].iters = (uint32_t)(iters % NOALLOC_CHUNK);
  hipStream_t st = ds.stream;
  NA_CHK(hipMemcpyAsync(dd, ds.desc, sizeof ds.desc, hipMemcpyHostToDevice, st));
  NA_CHK(hipMemcpyAsync((void*)P.tw61, c.T.tw61.data(), c.T.tw61.size() * sizeof(e61), hipMemcpyHostToDevice, st));
  NA_CHK(hipMemcpyAsync((void*)P.tw31, c.T.tw31.data(), c.T.tw31.size() * sizeof(e31), hipMemcpyHostToDevice, st));
  NA_CHK(hipMemcpyAsync((void*)P.sq61, c.T.sq61.data(), c.T.sq61.size() * sizeof(e61), hipMemcpyHostToDevice, st));
  NA_CHK(hipMemcpyAsync((void*)P.sq31, c.T.sq31.data(), c.T.sq31.size() * sizeof(e31), hipMemcpyHostToDevice, st));
  NA_CHK(hipMemcpyAsync(z, ds.digits, h * sizeof(e61), hipMemcpyHostToDevice, st));
  for (uint64_t left = iters; left > 0;) {
    const bool full = left >= NOALLOC_CHUNK;
    NA_CHK(launch_small(P, dd + (full ? 0 : 1), 1, st));
    left -= full ? NOALLOC_CHUNK : left;
  }
  NA_CHK(hipMemcpyAsync(ds.digits, z, h * sizeof(e61), hipMemcpyDeviceToHost, st));
  NA_CHK(hipStreamSynchronize(st));
  digits_to_residue(pp, ln, ds.digits, ds.words, out);
  return M2_OK;
}
