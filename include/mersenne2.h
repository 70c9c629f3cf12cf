/*
 * mersenne2 on MI355X: C ABI of the IBDWT squaring engine for Lucas-Lehmer and
 * base-3 PRP tests of 2^p - 1 (libm2.so).
 *
 * The reference (cherubrock-seb/mersenne2GPU) has no library interface: its
 * `mersenne2` program owns the whole test inside main() (mersenne2_gpu.cpp:
 * 3618-4960). Each entry point below replaces one part of that program, cited
 * per function; the `mersenne2` command line of this project is built on them.
 *
 * Conventions: every function returns 0 (M2_OK) or a negative M2_E* code and
 * never throws; m2_last_error() gives the message of the calling thread's last
 * failure. Plain pointers and sizes only. A test handle is used by one thread
 * at a time; a context may create and destroy tests from several threads.
 * Residues are canonical little-endian integers of ceil(p/8) bytes, 0 <= R <
 * 2^p - 1; Res64 is R mod 2^64. "Iteration k" is the state after k squarings
 * (PRP, from 3) or k Lucas-Lehmer steps (LL, from 4).
 */
#ifndef MERSENNE2_H
#define MERSENNE2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M2_OK 0
#define M2_E_INVAL (-1)     /* bad argument (usage error; the CLI exits 1) */
#define M2_E_DEVICE (-2)    /* no such device / no GPU */
#define M2_E_NOMEM (-3)     /* device or host allocation failed */
#define M2_E_HIP (-4)       /* a HIP runtime call failed */
#define M2_E_RANGE (-5)     /* exponent outside the supported range */
#define M2_E_STATE (-6)     /* call not valid in the object's current state */

#define M2_MODE_LL 0
#define M2_MODE_PRP 1

typedef struct m2_ctx m2_ctx;
typedef struct m2_test m2_test;
typedef struct m2_batch m2_batch;

/* Tuning options; zero fields mean "automatic". The reference's -W, -CW, -B
 * and -LS flags (mersenne2_gpu.cpp:3660-3681) map onto these. */
typedef struct m2_opts {
  uint32_t wg;              /* -W: workgroup size of the centre pass: 1024 (4 elements per
                               thread, radix-4 steps) or 512 (8 per thread); other values leave
                               the choice to the engine */
  uint32_t carry_wg;        /* -CW: carry work-group size hint (unused: carries run inside the passes) */
  uint32_t carry_pairs;     /* -B: digit pairs per carry block (a strip row of the outer pass),
                               rounded down to a power of two and clamped to the plan */
  uint32_t local_stage_cap; /* -LS: cap on the points of the LDS-resident centre block; a nonzero
                               value also selects the multi-pass plan on short transforms */
  uint32_t graph_iters;     /* iterations per captured hipGraph (0: 16) */
  uint32_t max_radix;       /* 8 (0: default) or 4: largest radix of the transform steps */
  uint32_t autotune;        /* batches and servers: choose the four fields above per transform
                               length by timing candidates (m2_autotune) */
  uint32_t no_fuse;         /* single tests: 1 = keep the outer passes of consecutive iterations
                               in separate launches (0: fuse them where the plan allows) */
} m2_opts;

/* Message of the calling thread's last failure ("" if none). */
const char *m2_last_error(void);

/* Number of visible HIP devices (the reference's GPU enumeration,
 * mersenne2_gpu.cpp:3735-3758). */
int m2_device_count(int *count);

/* Open device `device` (-d, mersenne2_gpu.cpp:3646-3651, :3753-3764). */
int m2_open(int device, m2_ctx **out);
int m2_close(m2_ctx *ctx);
/* Device name and properties (CL_DEVICE_NAME etc., mersenne2_gpu.cpp:3759-3770). */
int m2_device_name(m2_ctx *ctx, char *buf, size_t len);
int m2_device_props(m2_ctx *ctx, uint32_t *compute_units, uint64_t *global_mem, uint32_t *max_wg,
                    uint32_t *lds_bytes);

/* log2 of the transform length for p (transformsize, mersenne2_gpu.cpp:286-293). */
int m2_transform_log2(uint64_t p, int *ln);
/* Device bytes one test of exponent p holds while resident. */
int m2_test_footprint(uint64_t p, uint64_t *bytes);

/* Create a test at iteration 0 (tables, buffers and initial residue:
 * mersenne2_gpu.cpp:3902-4028). mode: M2_MODE_LL or M2_MODE_PRP. */
int m2_test_create(m2_ctx *ctx, uint64_t p, int mode, const m2_opts *opts, m2_test **out);
int m2_test_destroy(m2_test *t);

/* Enqueue `iters` iterations on the test's stream and return (the hot loop,
 * mersenne2_gpu.cpp:4811-4853). */
int m2_test_run(m2_test *t, uint64_t iters);
/* Wait for the test's stream (clFinish, mersenne2_gpu.cpp:4788, :4855). */
int m2_test_sync(m2_test *t);
int m2_test_iteration(m2_test *t, uint64_t *iter);
/* Iterations of a full test: p (PRP) or p - 2 (LL) (mersenne2_gpu.cpp:4780). */
int m2_test_total_iters(m2_test *t, uint64_t *iters);

/* Res64 and full residue of the current state (synchronises). The reference
 * prints neither; it reads the residue back (mersenne2_gpu.cpp:4862) only to
 * test it. */
int m2_test_res64(m2_test *t, uint64_t *res64);
int m2_test_residue(m2_test *t, uint8_t *buf, size_t len);
/* Replace the state by residue R at iteration `iteration` (resume). */
int m2_test_set_residue(m2_test *t, const uint8_t *buf, size_t len, uint64_t iteration);
/* Verdict of the current state (verify_is_zero / verify_is_Mp /
 * verify_prp_residue_9, mersenne2_gpu.cpp:3028-3076, :4865-4867):
 * 1 if LL residue is 0 or PRP residue is 9 (mod 2^p - 1), else 0. */
int m2_test_is_prime(m2_test *t, int *verdict);

/* Run `iters` iterations and return their wall time on the device in ms
 * (hipEvents around the launches). */
int m2_test_time(m2_test *t, uint64_t iters, double *ms);
/* Run `iters` iterations, recording every launch's dispatch start/stop
 * timestamps. names_buf receives the kernel names separated by '\n'; ms[i]
 * is kernel i's mean device time per launch, in milliseconds. */
int m2_test_kernel_times(m2_test *t, uint64_t iters, double *ms, int max_kernels, int *count, char *names_buf,
                         size_t names_len);
/* The same run, reporting each kernel's total device time (ms) and launch
 * count: the per-kernel event profile behind the CLI's --profile-kernels
 * (gf61_ntt_prp_opencl_inplace_nobitrev_fused_powonly_blockcarry.cpp:645-698,
 * which records OpenCL profiling events per launch and prints totals). The
 * test advances by `iters`; launches are not graphed while profiled. */
int m2_test_kernel_profile(m2_test *t, uint64_t iters, double *total_ms, uint64_t *launches, int max_kernels,
                           int *count, char *names_buf, size_t names_len);
/* Human-readable plan (the reference's --debug-plan, mersenne2_gpu.cpp:4767-4776). */
int m2_test_plan(m2_test *t, char *buf, size_t len);

/* Test hook: the number of polls the fused outer pass makes while waiting
 * for a neighbouring workgroup's carries before it gives up and reports
 * M2_E_STATE at the next sync (default 2^20, about a second). 0 reports the
 * time-out without waiting, so the reporting path can be exercised without a
 * real stall. Applies to launches made afterwards (a test's captured graphs
 * keep the value they were made with). */
int m2_debug_fused_spin_limit(uint32_t polls, uint32_t *previous);
/* Measurement hook (no counterpart in the reference): the device's issue rate
 * of v_mad_u64_u32, the 32x32->64 multiply-add the field products are made
 * of, in lane operations per second, timed with hipEvents over a short probe
 * kernel. It is the peak against which the integer roofline is priced. */
int m2_debug_product_rate(m2_ctx *ctx, double *lane_ops_per_s);

/* PRP run with a Gerbicz error check (the reference has none; its loop,
 * mersenne2_gpu.cpp:4811-4853, trusts every iteration). Advances a PRP test by
 * `iters` iterations in blocks of `block`: after every block the product d of
 * the block-end residues since the last verified state x_0 is updated, and
 * every `blocks_per_check` blocks (and after the last full block) the identity
 * d_t = x_0 d_{t-1}^(2^block) is checked with `block` extra squarings. A
 * failed check rolls the test back to the last verified state and repeats the
 * blocks; the iterations after the last full block are run twice from the
 * verified state and compared. *errors (optional) receives the number of
 * failed checks; three failures in a row return M2_E_STATE. LL tests are
 * refused (M2_E_INVAL). The results are those of m2_test_run. */
int m2_test_run_checked(m2_test *t, uint64_t iters, uint32_t block, uint32_t blocks_per_check, uint64_t *errors);
/* Test hook: flip one bit of the residue after block `block` (1-based; 0: never)
 * of the next m2_test_run_checked call, once, so that its check must catch it. */
int m2_debug_gerbicz_fault(uint64_t block);

/* Checkpoints: the canonical residue with p, mode, iteration and Res64 in a
 * small checksummed file (absent in the reference, which cannot stop and
 * resume a test). Save writes a temporary file and renames it over `path`.
 * Load needs a test of the same p and mode and continues from the saved
 * iteration; a damaged or foreign file is refused (M2_E_INVAL). Info reads a
 * file's header on the host, without a device. */
int m2_checkpoint_save(m2_test *t, const char *path);
int m2_checkpoint_load(m2_test *t, const char *path);
int m2_checkpoint_info(const char *path, uint64_t *p, int *mode, uint64_t *iteration, uint64_t *res64);

/* Autotuner: for the transform length of p and a batch of `ntests` tests,
 * time the candidate plans (single workgroup or multi-pass; centre size, carry
 * block width, radix, then graph length) and write the fastest into *out
 * (local_stage_cap, carry_pairs, max_radix, graph_iters). Results are cached
 * per context by length and batch-size bucket. Every candidate computes the
 * same residues. `report` (optional) receives one line per candidate with its
 * device time per iteration. The reference's counterpart is its device-name
 * profile table and planner (mersenne2_gpu.cpp:3515-3616, :4034-4110). */
int m2_autotune(m2_ctx *ctx, uint64_t p, uint32_t ntests, m2_opts *out, char *report, size_t report_len);

/*
 * Batches: many tests, of any mix of exponents, resident on one device and
 * advanced together. Tests of the same transform length share every launch
 * (one workgroup per test on single-workgroup lengths, one grid row per test
 * on multi-pass lengths); lengths run on separate streams. The reference runs
 * one exponent per process (mersenne2_gpu.cpp:3618-4960); a batch gives each
 * test the result that process would compute for it.
 */

/* Create `count` tests (ps[i], modes[i]); iters[i] is the number of
 * iterations test i stops at (0: its full test, p or p - 2). iters may be NULL. */
int m2_batch_create(m2_ctx *ctx, size_t count, const uint64_t *ps, const int *modes, const uint64_t *iters,
                    const m2_opts *opts, m2_batch **out);
int m2_batch_destroy(m2_batch *b);
/* Enqueue up to `max_iters` more iterations of every unfinished test (0: run
 * every test to its stop) and return. */
int m2_batch_run(m2_batch *b, uint64_t max_iters);
int m2_batch_sync(m2_batch *b);
/* Iterations test i has had enqueued. */
int m2_batch_iteration(m2_batch *b, size_t i, uint64_t *iter);
/* Res64, residue and verdict of test i (these wait for test i's stream). */
int m2_batch_res64(m2_batch *b, size_t i, uint64_t *res64);
int m2_batch_residue(m2_batch *b, size_t i, uint8_t *buf, size_t len);
int m2_batch_is_prime(m2_batch *b, size_t i, int *verdict);

/*
 * Server: a request queue over resident tests (continuous batching). A
 * request is admitted, in arrival order, once its device footprint
 * (m2_test_footprint) fits in the budget. Admitted tests of one transform
 * length form a lane that runs on its own stream, one slice at a time: a
 * slice advances every member of the lane by at most `slice_iters`
 * iterations and by at most about 4 ms of device time (measured per lane),
 * so long tests never hold the device while short ones wait. When a lane's
 * slice ends, m2_server_step retires its finished tests (their Res64 and
 * verdict go to the completion queue read by m2_server_poll; their memory is
 * freed) and starts the lane's next slice with the tests admitted meanwhile.
 * No counterpart in the reference, which tests one exponent per run.
 * m2_server_submit / _poll / _stats may be called from other threads while
 * one thread drives m2_server_step.
 */
typedef struct m2_server m2_server;
/* hbm_budget_bytes 0: 70% of the device memory; slice_iters 0: 1000. */
int m2_server_create(m2_ctx *ctx, uint64_t hbm_budget_bytes, uint32_t slice_iters, const m2_opts *opts,
                     m2_server **out);
int m2_server_destroy(m2_server *s);
/* Queue a test of 2^p - 1; iters 0 means the full test, else the test stops
 * after `iters` iterations (its verdict is then that of the partial state). */
int m2_server_submit(m2_server *s, uint64_t p, int mode, uint64_t iters, uint64_t *id);
/* Admit, retire and start slices on lanes whose slice has ended (waits up to
 * about 1 ms for one to end). *busy = 1 while requests remain queued or active. */
int m2_server_step(m2_server *s, int *busy);
/* Take up to `max` completions (any of ids/res64/verdicts may be NULL). */
int m2_server_poll(m2_server *s, uint64_t *ids, uint64_t *res64, int *verdicts, size_t max, size_t *count);
int m2_server_stats(m2_server *s, uint64_t *queued, uint64_t *active, uint64_t *bytes_in_use);

/*
 * No-alloc variant. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation). Advances the canonical residue
 * `in` (ceil(p/8) bytes, little-endian) of a test of 2^p - 1 by `iters` PRP
 * squarings (mode 1) or Lucas-Lehmer steps (mode 0) and writes the result to
 * `out` (ceil(p/8) bytes): the values m2_test_set_residue + m2_test_run +
 * m2_test_residue give. Every device buffer it uses (test descriptors, twiddle
 * tables, digits) lies in one fixed array of M2_NOALLOC_SCRATCH_KIB KiB per
 * device, set up by m2_open and held until the process ends; the call itself
 * makes no hipMalloc, hipFree or other allocation. Only exponents whose
 * transform fits one workgroup (n <= 2^12, p <= 159739) fit the array; larger
 * ones return M2_E_RANGE. Calls on one device share its array: they are
 * serialised by a lock, and calls that overlap on one device from several
 * processes would race. Synchronous: returns when `out` holds the result.
 */
#define M2_NOALLOC_SCRATCH_KIB 128
int m2_square_noalloc(m2_ctx *ctx, uint64_t p, int mode, const uint8_t *in, size_t in_len, uint64_t iters,
                      uint8_t *out, size_t out_len);

#ifdef __cplusplus
}
#endif
#endif
