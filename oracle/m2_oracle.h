/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, single-threaded CPU restatement of the reference's IBDWT squaring
 * iteration for Lucas-Lehmer and base-3 PRP tests of 2^p - 1 (mersenne2_gpu.cpp).
 * Only tests/, __graft_entry__.smoke() and bench.py (its cpu_baseline leg) may
 * load this library, and only to verify results or as the timed CPU baseline. The
 * product (mersenne2gpu_amd/) never links, loads or calls it.
 *
 * Every function cites the reference lines it restates. How each one is pinned
 * (closed forms, brute force, GMP) is listed in tests/test_oracle.py.
 */
#ifndef M2_ORACLE_H
#define M2_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One element of GF(M61^2) x GF(M31^2): same 24-byte layout as the reference's
 * device struct GF (mersenne2_gpu.cpp:299) and host GF61_31 (:241-261). */
typedef struct { uint64_t s0, s1; uint32_t t0, t1; } m2o_gf;

/* ---- scalar field arithmetic (host classes :78-169, device :336-391) ---- */
uint64_t m2o_add61(uint64_t a, uint64_t b);
uint64_t m2o_sub61(uint64_t a, uint64_t b);
uint64_t m2o_mul61(uint64_t a, uint64_t b);
uint64_t m2o_shl61(uint64_t a, unsigned s);   /* a * 2^s mod M61 */
uint32_t m2o_add31(uint32_t a, uint32_t b);
uint32_t m2o_sub31(uint32_t a, uint32_t b);
uint32_t m2o_mul31(uint32_t a, uint32_t b);
uint32_t m2o_shl31(uint32_t a, unsigned s);   /* a * 2^s mod M31 */

/* GF(p^2) product of both fields at once (Karatsuba form, device :400-416). */
void m2o_gf_mul(const m2o_gf *a, const m2o_gf *b, m2o_gf *out);

/* ---- transform size and IBDWT tables ---- */
int m2o_transformsize(uint32_t p);                                  /* :286-293 */
/* widths[n], w61[n], w31[n]; returns 0, or -1 on a width outside {q, q+1} (:3930-3933) */
int m2o_digit_tables(uint32_t p, int ln, uint8_t *width, uint8_t *w61, uint8_t *w31);
/* Twiddle table of n entries (index 0 and n/2 unused), :3911-3917. */
void m2o_twiddles(int ln, m2o_gf *w);
/* Root of unity of order 2^k in each field (root_nth, :202 and :237). */
void m2o_root(int k, m2o_gf *out);

/* ---- transform stages (z has h = n/2 elements) ---- */
void m2o_forward(int ln, m2o_gf *z, const m2o_gf *w);         /* :732, :571, plan :4327-4511 */
void m2o_square_half(int ln, m2o_gf *z, const m2o_gf *w);     /* :856-870 */
void m2o_inverse(int ln, m2o_gf *z, const m2o_gf *w);         /* :951, :580, plan :4534-4731 */
void m2o_weight(int ln, m2o_gf *z, const uint8_t *w61, const uint8_t *w31);   /* :517 */
void m2o_unweight(int ln, m2o_gf *z, const uint8_t *w61, const uint8_t *w31); /* :1975 */
/* Garner CRT of (x mod M61, x mod M31) into x < 2^92 (:2018-2030); out[0]=lo, out[1]=hi. */
void m2o_garner(uint64_t s, uint32_t t, uint64_t out[2]);

/* Host threads the independent loops are shared among (n > 0 sets it; 0
 * leaves it); returns the count in use (1 without OpenMP). */
int m2o_set_threads(int n);

/* ---- whole test ---- */
typedef struct m2o_test m2o_test;
/* mode: 0 = LL (start 4), 1 = PRP base 3 (start 3), :3908. carry_pairs: block size (:3946-3967). */
m2o_test *m2o_test_new(uint32_t p, int mode, uint32_t carry_pairs);
void m2o_test_free(m2o_test *t);
int m2o_test_ln(const m2o_test *t);
/* Runs `iters` iterations (:4811-4853 with :4754-4765 as the post plan). */
void m2o_test_run(m2o_test *t, uint64_t iters);
uint64_t m2o_test_iter(const m2o_test *t);
/* The n digits (s0/s1 of each element, in digit order). */
void m2o_test_digits(const m2o_test *t, uint64_t *out);
/* Canonical residue R = (sum d_j 2^ceil(pj/n)) mod (2^p-1), little-endian, len >= ceil(p/8). */
int m2o_test_residue(const m2o_test *t, uint8_t *buf, size_t len);
uint64_t m2o_test_res64(const m2o_test *t);
/* Reference verdict (:3028-3076, :4865-4867): 1 if prime / probable prime. */
int m2o_test_verdict(const m2o_test *t);

/* Canonical residue from an arbitrary digit vector (widths from p, n = 2^ln). */
int m2o_digits_residue(uint32_t p, int ln, const uint64_t *digits, uint8_t *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
