/*
 * ORACLE — test infrastructure only (see m2_oracle.h).
 *
 * Plain-C restatement of the reference's exact IBDWT squaring iteration
 * (cherubrock-seb/mersenne2GPU, mersenne2_gpu.cpp). The arithmetic follows the
 * reference step for step: the same radix-4 stage order with the same twiddle
 * table, the same real-data centre pass, the same shift weights, Garner CRT
 * and block carry. Citations are mersenne2_gpu.cpp line numbers.
 *
 * Deliberate difference: digit widths of 0 bits (only for p < 8, where
 * q = p >> ln is 0) are handled exactly; the reference shifts a 64-bit word
 * by 64 in that case (digit_adc, :1982-1993), which is undefined.
 *
 * Built with OpenMP, the loops whose iterations are independent (butterflies
 * of one stage, centre pairs, weights, the carry blocks' prepare step) are
 * shared among m2o_set_threads() host threads, as the reference's work-items
 * are among a CPU OpenCL device's cores; the results do not depend on the
 * thread count. The carry drain and the LL subtract stay serial, as in the
 * reference (single work-item kernels, :2967, :3006).
 */
#include "m2_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Work below this many elements per loop stays on one thread. */
#define PAR_MIN 4096

int m2o_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

typedef unsigned __int128 u128;

#define P61 ((((uint64_t)1) << 61) - 1)
#define P31 ((((uint32_t)1) << 31) - 1)

/* ------------------------------------------------------------------ */
/* Scalar fields: Z61 (:78-122, device :336-363), Z31 (:124-169, :368-391) */

uint64_t m2o_add61(uint64_t a, uint64_t b) {
    uint64_t t = a + b;
    return t >= P61 ? t - P61 : t;
}
uint64_t m2o_sub61(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P61 - b; }
uint64_t m2o_mul61(uint64_t a, uint64_t b) {
    u128 t = (u128)a * b;
    uint64_t r = ((uint64_t)t & P61) + (uint64_t)(t >> 61);
    r = (r & P61) + (r >> 61);
    return r >= P61 ? r - P61 : r;
}
/* 2^s has order 61 mod M61, so a*2^s is a rotation of the 61-bit word (:358-366). */
uint64_t m2o_shl61(uint64_t a, unsigned s) {
    s %= 61u;
    if (s == 0) return a;
    return m2o_add61((a << s) & P61, a >> (61u - s));
}
uint32_t m2o_add31(uint32_t a, uint32_t b) {
    uint32_t t = a + b;
    return t >= P31 ? t - P31 : t;
}
uint32_t m2o_sub31(uint32_t a, uint32_t b) { return a >= b ? a - b : a + P31 - b; }
uint32_t m2o_mul31(uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t r = (uint32_t)(t & P31) + (uint32_t)(t >> 31);
    return r >= P31 ? r - P31 : r;
}
uint32_t m2o_shl31(uint32_t a, unsigned s) {
    s %= 31u;
    if (s == 0) return a;
    return m2o_add31((a << s) & P31, a >> (31u - s));
}
static uint64_t shr61(uint64_t a, unsigned s) { return m2o_shl61(a, 61u - (s % 61u)); }
static uint32_t shr31(uint32_t a, unsigned s) { return m2o_shl31(a, 31u - (s % 31u)); }

/* ------------------------------------------------------------------ */
/* GF(p^2), i^2 = -1, both fields side by side (:171-261, device :393-454). */

static m2o_gf gf_add(m2o_gf a, m2o_gf b) {
    m2o_gf r = {m2o_add61(a.s0, b.s0), m2o_add61(a.s1, b.s1), m2o_add31(a.t0, b.t0), m2o_add31(a.t1, b.t1)};
    return r;
}
static m2o_gf gf_sub(m2o_gf a, m2o_gf b) {
    m2o_gf r = {m2o_sub61(a.s0, b.s0), m2o_sub61(a.s1, b.s1), m2o_sub31(a.t0, b.t0), m2o_sub31(a.t1, b.t1)};
    return r;
}
/* a + conj(b), a - conj(b), conj(a - b) (:396-398) */
static m2o_gf gf_addconj(m2o_gf a, m2o_gf b) {
    m2o_gf r = {m2o_add61(a.s0, b.s0), m2o_sub61(a.s1, b.s1), m2o_add31(a.t0, b.t0), m2o_sub31(a.t1, b.t1)};
    return r;
}
static m2o_gf gf_subconj(m2o_gf a, m2o_gf b) {
    m2o_gf r = {m2o_sub61(a.s0, b.s0), m2o_add61(a.s1, b.s1), m2o_sub31(a.t0, b.t0), m2o_add31(a.t1, b.t1)};
    return r;
}
static m2o_gf gf_sub_conj(m2o_gf a, m2o_gf b) {
    m2o_gf r = {m2o_sub61(a.s0, b.s0), m2o_sub61(b.s1, a.s1), m2o_sub31(a.t0, b.t0), m2o_sub31(b.t1, a.t1)};
    return r;
}
/* a + i*b, a - i*b (:399-400) */
static m2o_gf gf_addi(m2o_gf a, m2o_gf b) {
    m2o_gf r = {m2o_sub61(a.s0, b.s1), m2o_add61(a.s1, b.s0), m2o_sub31(a.t0, b.t1), m2o_add31(a.t1, b.t0)};
    return r;
}
static m2o_gf gf_subi(m2o_gf a, m2o_gf b) {
    m2o_gf r = {m2o_add61(a.s0, b.s1), m2o_sub61(a.s1, b.s0), m2o_add31(a.t0, b.t1), m2o_sub31(a.t1, b.t0)};
    return r;
}
/* Karatsuba product, 3 base multiplications per field (:401-416). */
static m2o_gf gf_mul(m2o_gf a, m2o_gf b) {
    uint64_t ac = m2o_mul61(a.s0, b.s0), bd = m2o_mul61(a.s1, b.s1);
    uint64_t k = m2o_mul61(m2o_add61(a.s0, a.s1), m2o_add61(b.s0, b.s1));
    uint32_t ac3 = m2o_mul31(a.t0, b.t0), bd3 = m2o_mul31(a.t1, b.t1);
    uint32_t k3 = m2o_mul31(m2o_add31(a.t0, a.t1), m2o_add31(b.t0, b.t1));
    m2o_gf r = {m2o_sub61(ac, bd), m2o_sub61(m2o_sub61(k, ac), bd), m2o_sub31(ac3, bd3),
                m2o_sub31(m2o_sub31(k3, ac3), bd3)};
    return r;
}
/* a * conj(b) (:417-433) */
static m2o_gf gf_mulconj(m2o_gf a, m2o_gf b) {
    uint64_t ac = m2o_mul61(a.s0, b.s0), bd = m2o_mul61(a.s1, b.s1);
    uint64_t k = m2o_mul61(m2o_add61(a.s0, a.s1), m2o_sub61(b.s0, b.s1));
    uint32_t ac3 = m2o_mul31(a.t0, b.t0), bd3 = m2o_mul31(a.t1, b.t1);
    uint32_t k3 = m2o_mul31(m2o_add31(a.t0, a.t1), m2o_sub31(b.t0, b.t1));
    m2o_gf r = {m2o_add61(ac, bd), m2o_add61(m2o_sub61(k, ac), bd), m2o_add31(ac3, bd3),
                m2o_add31(m2o_sub31(k3, ac3), bd3)};
    return r;
}
/* (a+ib)^2 = a^2-b^2 + 2ab i (:434-444) */
static m2o_gf gf_sqr(m2o_gf a) {
    uint64_t t = m2o_mul61(a.s0, a.s1);
    uint32_t t3 = m2o_mul31(a.t0, a.t1);
    m2o_gf r = {m2o_sub61(m2o_mul61(a.s0, a.s0), m2o_mul61(a.s1, a.s1)), m2o_add61(t, t),
                m2o_sub31(m2o_mul31(a.t0, a.t0), m2o_mul31(a.t1, a.t1)), m2o_add31(t3, t3)};
    return r;
}
void m2o_gf_mul(const m2o_gf *a, const m2o_gf *b, m2o_gf *out) { *out = gf_mul(*a, *b); }

static m2o_gf gf_pow(m2o_gf x, uint64_t e) {
    m2o_gf r = {1, 0, 1, 0};
    while (e) {
        if (e & 1) r = gf_mul(r, x);
        x = gf_sqr(x);
        e >>= 1;
    }
    return r;
}

/* Generators of the 2-Sylow subgroups: order 2^62 in GF(M61^2) (:174) and
 * order 2^32 in GF(M31^2) (:209); root_nth(2^k) = g^(ORDER / 2^k) (:202, :237). */
static const uint64_t G61_RE = 264036120304204ull, G61_IM = 4677669021635377ull;
static const uint32_t G31_RE = 7735u, G31_IM = 748621u;

void m2o_root(int k, m2o_gf *out) {
    m2o_gf g = {G61_RE, G61_IM, G31_RE, G31_IM};
    m2o_gf r61 = gf_pow(g, ((uint64_t)1) << (62 - k));
    m2o_gf r31 = gf_pow(g, ((uint64_t)1) << (32 - k));
    out->s0 = r61.s0;
    out->s1 = r61.s1;
    out->t0 = r31.t0;
    out->t1 = r31.t1;
}

/* ------------------------------------------------------------------ */
/* Transform size and IBDWT tables */

/* Smallest ln >= 3 with ln + 2*((p >> ln) + 1) < 92 (:286-293). */
int m2o_transformsize(uint32_t p) {
    int ln = 2;
    uint32_t w;
    do {
        ++ln;
        w = p >> ln;
    } while ((uint64_t)ln + 2u * ((uint64_t)w + 1u) >= 92u);
    return ln;
}

/* Digit j has width ceil(p(j+1)/n) - ceil(pj/n); its weight 2^((n - (pj mod n))/n)
 * is the shift lr*(n - (pj mod n)) where 2^lr is an n-th root of 2 in each field:
 * lr61 = (2^60/n) mod 61, lr31 = (2^30/n) mod 31 (:203, :238, :3921-3944). */
int m2o_digit_tables(uint32_t p, int ln, uint8_t *width, uint8_t *w61, uint8_t *w31) {
    const uint64_t n = (uint64_t)1 << ln;
    const uint32_t q = p >> ln;
    const uint64_t lr61 = (((uint64_t)1) << 60 >> ln) % 61u;
    const uint64_t lr31 = (((uint64_t)1) << 30 >> ln) % 31u;
    uint64_t prev = 0;
    for (uint64_t j = 1; j <= n; ++j) {
        uint64_t pj = (uint64_t)p * j;
        uint64_t c = ((pj - 1) >> ln) + 1;
        uint64_t wd = c - prev;
        if (wd != q && wd != q + 1u) return -1;
        width[j - 1] = (uint8_t)wd;
        prev = c;
        if (j < n) {
            uint64_t r = pj & (n - 1);
            w61[j] = (uint8_t)((lr61 * (n - r)) % 61u);
            w31[j] = (uint8_t)((lr31 * (n - r)) % 31u);
        }
    }
    w61[0] = 0;
    w31[0] = 0;
    return 0;
}

static uint64_t bitrev(uint64_t i, uint64_t s) {
    uint64_t r = 0;
    for (; s > 1; s >>= 1, i >>= 1) r = (r << 1) | (i & 1);
    return r;
}

/* w[s+j] = r_{2s}^bitrev(j,s) for s = 1..n/4, and the radix-4 companions
 * w[n/2+s+j] = w[s+j]*w[2(s+j)] (:3911-3917). Entries whose source lies outside
 * the filled range are left zero, as in the reference; no stage reads them. */
void m2o_twiddles(int ln, m2o_gf *w) {
    const uint64_t n = (uint64_t)1 << ln;
    memset(w, 0, sizeof(m2o_gf) * n);
    for (uint64_t s = 1; s <= n / 4; s <<= 1) {
        int k = 0;
        while ((((uint64_t)1) << k) < 2 * s) ++k;
        m2o_gf r;
        m2o_root(k, &r);
        for (uint64_t j = 0; j < s; ++j) w[s + j] = gf_pow(r, bitrev(j, s));
    }
    for (uint64_t s = 1; s <= n / 4; s <<= 1)
        for (uint64_t j = 0; j < s; ++j) w[n / 2 + s + j] = gf_mul(w[s + j], w[2 * (s + j)]);
}

/* ------------------------------------------------------------------ */
/* Transform stages over h = n/2 elements. */

/* One radix-4 stage with s blocks of 4m (forward4, :732-747): inputs 1,2,3 are
 * multiplied by w[2(s+j)], w[s+j], w[h+s+j], then the (+-1, +-i) butterfly. */
static void fwd4(m2o_gf *z, const m2o_gf *w, uint64_t h, uint64_t s, uint64_t m) {
#pragma omp parallel for collapse(2) if (s * m >= PAR_MIN)
    for (uint64_t j = 0; j < s; ++j) {
        for (uint64_t i = 0; i < m; ++i) {
            const m2o_gf w1 = w[2 * (s + j)], w2 = w[s + j], w3 = w[h + s + j];
            const uint64_t b = j * 4 * m;
            m2o_gf u0 = z[b + i];
            m2o_gf u1 = gf_mul(z[b + m + i], w1);
            m2o_gf u2 = gf_mul(z[b + 2 * m + i], w2);
            m2o_gf u3 = gf_mul(z[b + 3 * m + i], w3);
            m2o_gf v0 = gf_add(u0, u2), v1 = gf_add(u1, u3), v2 = gf_sub(u0, u2), v3 = gf_sub(u1, u3);
            z[b + i] = gf_add(v0, v1);
            z[b + m + i] = gf_sub(v0, v1);
            z[b + 2 * m + i] = gf_addi(v2, v3);
            z[b + 3 * m + i] = gf_subi(v2, v3);
        }
    }
}

/* Inverse radix-4 stage (backward4, :951-962). */
static void bwd4(m2o_gf *z, const m2o_gf *w, uint64_t h, uint64_t s, uint64_t m) {
#pragma omp parallel for collapse(2) if (s * m >= PAR_MIN)
    for (uint64_t j = 0; j < s; ++j) {
        for (uint64_t i = 0; i < m; ++i) {
            const m2o_gf w1 = w[2 * (s + j)], w2 = w[s + j], w3 = w[h + s + j];
            const uint64_t b = j * 4 * m;
            m2o_gf u0 = z[b + i], u1 = z[b + m + i], u2 = z[b + 2 * m + i], u3 = z[b + 3 * m + i];
            m2o_gf v0 = gf_add(u0, u1), v1 = gf_sub(u0, u1), v2 = gf_add(u2, u3), v3 = gf_sub(u3, u2);
            z[b + i] = gf_add(v0, v2);
            z[b + 2 * m + i] = gf_mulconj(gf_sub(v0, v2), w2);
            z[b + m + i] = gf_mulconj(gf_addi(v1, v3), w1);
            z[b + 3 * m + i] = gf_mulconj(gf_subi(v1, v3), w3);
        }
    }
}

/* Stage order of the reference plan (:4327-4511), mirroring its current_m
 * (:4328, :4365): radix-4 stages with m = h/4, h/16, ... while a full stage
 * fits, then radix-2 (forward2, :571) when log2 h is odd. Output is the
 * length-h DFT in bit-reversed order. */
static void forward_plan(int ln, m2o_gf *z, const m2o_gf *w) {
    const uint64_t n = ((uint64_t)1) << ln, h = n / 2;
    uint64_t current_m = n / 4, s = 1;
    for (; current_m > 1; current_m /= 4, s *= 4) fwd4(z, w, h, s, current_m / 2);
    if (current_m == 1) {
#pragma omp parallel for if (h / 2 >= PAR_MIN)
        for (uint64_t j = 0; j < h / 2; ++j) {
            m2o_gf u0 = z[2 * j], u1 = gf_mul(z[2 * j + 1], w[h / 2 + j]);
            z[2 * j] = gf_add(u0, u1);
            z[2 * j + 1] = gf_sub(u0, u1);
        }
    }
}

/* Inverse plan (:4534-4731): backward2 (:580) first when log2 h is odd, then
 * the inverse radix-4 stages from the smallest stride to the largest. */
static void inverse_plan(int ln, m2o_gf *z, const m2o_gf *w) {
    const uint64_t n = ((uint64_t)1) << ln, h = n / 2;
    uint64_t current_m = n / 4, s = 1;
    for (; current_m > 1; current_m /= 4, s *= 4) {
    }
    if (current_m == 1) {
#pragma omp parallel for if (h / 2 >= PAR_MIN)
        for (uint64_t j = 0; j < h / 2; ++j) {
            m2o_gf u0 = z[2 * j], u1 = z[2 * j + 1];
            z[2 * j] = gf_add(u0, u1);
            z[2 * j + 1] = gf_mulconj(gf_sub(u0, u1), w[h / 2 + j]);
        }
    }
    for (uint64_t back_m = (current_m == 1) ? 4 : 2, back_s = s / 4; back_s >= 1; back_m *= 4, back_s /= 4)
        bwd4(z, w, h, back_s, back_m / 2);
}

/* Real-data centre pass (square_half, :856-870). Position k = 2j pairs with
 * mk = 3*2^floor(log2 k) - k - 1 (its mirror in bit-reversed order); the
 * twiddle w[h/2+j] is the h-th root of unity of position k. */
void m2o_square_half(int ln, m2o_gf *z, const m2o_gf *w) {
    const uint64_t h = ((uint64_t)1) << (ln - 1);
    /* pairs (k, mk) are disjoint: k even, mk odd and >= 3 for k >= 2 */
#pragma omp parallel for if (h / 2 >= PAR_MIN)
    for (uint64_t j = 0; j < h / 2; ++j) {
        uint64_t k = 2 * j, mk = 0;
        if (k != 0) {
            uint64_t top = 1;
            while ((top << 1) <= k) top <<= 1;
            mk = 3 * top - k - 1;
        }
        m2o_gf zk = z[k], zmk = z[mk];
        m2o_gf u0 = gf_addconj(zk, zmk), u1 = gf_subconj(zk, zmk);
        m2o_gf v0 = gf_sub(gf_sqr(u0), gf_mul(gf_sqr(u1), w[h / 2 + j]));
        m2o_gf v1 = gf_mul(u0, gf_add(u1, u1));
        z[k] = gf_add(v0, v1);
        if (k == 0)
            z[1] = gf_sqr(gf_add(z[1], z[1]));
        else
            z[mk] = gf_sub_conj(v0, v1);
    }
}

void m2o_forward(int ln, m2o_gf *z, const m2o_gf *w) { forward_plan(ln, z, w); }
void m2o_inverse(int ln, m2o_gf *z, const m2o_gf *w) { inverse_plan(ln, z, w); }

/* z_k <- (d_2k * 2^w[2k], d_2k+1 * 2^w[2k+1]) in each field (weight, :517). */
void m2o_weight(int ln, m2o_gf *z, const uint8_t *w61, const uint8_t *w31) {
    const uint64_t h = ((uint64_t)1) << (ln - 1);
#pragma omp parallel for if (h >= PAR_MIN)
    for (uint64_t k = 0; k < h; ++k) {
        z[k].s0 = m2o_shl61(z[k].s0, w61[2 * k]);
        z[k].s1 = m2o_shl61(z[k].s1, w61[2 * k + 1]);
        z[k].t0 = m2o_shl31(z[k].t0, w31[2 * k]);
        z[k].t1 = m2o_shl31(z[k].t1, w31[2 * k + 1]);
    }
}

/* z_k <- z_k * 2^-(w + ln + 1): removes the weight, the 1/h of the inverse and
 * the factor 4 of the centre pass (unweight_norm, :1975-1980). */
void m2o_unweight(int ln, m2o_gf *z, const uint8_t *w61, const uint8_t *w31) {
    const uint64_t h = ((uint64_t)1) << (ln - 1);
    const unsigned l1 = (unsigned)ln + 1u;
#pragma omp parallel for if (h >= PAR_MIN)
    for (uint64_t k = 0; k < h; ++k) {
        z[k].s0 = shr61(z[k].s0, w61[2 * k] + l1);
        z[k].s1 = shr61(z[k].s1, w61[2 * k + 1] + l1);
        z[k].t0 = shr31(z[k].t0, w31[2 * k] + l1);
        z[k].t1 = shr31(z[k].t1, w31[2 * k + 1] + l1);
    }
}

/* x = t + M31 * ((s - t) * M31^-1 mod M61), with M31^-1 = 2^31 + 1 mod M61
 * (garner_reconstruct_from_u61_u31, :2018-2030). */
void m2o_garner(uint64_t s, uint32_t t, uint64_t out[2]) {
    uint64_t u = m2o_sub61(s, t);
    u = m2o_add61(u, m2o_shl61(u, 31));
    u128 x = ((u128)u << 31) - u + t;
    out[0] = (uint64_t)x;
    out[1] = (uint64_t)(x >> 64);
}

/* ------------------------------------------------------------------ */
/* Carry (digit_adc :1982-1993; block_prepare :2097-2123; init_block_carry
 * :2922-2930; block_carry_drain :2967-3004) and LL subtract (sub_kernel
 * :3006-3024). */

static uint64_t digit_adc(u128 lhs, unsigned wd, u128 *carry) {
    u128 s = lhs + *carry;
    if (wd == 0) {
        *carry = s;
        return 0;
    }
    *carry = s >> wd;
    return (uint64_t)s & ((((uint64_t)1) << wd) - 1);
}

struct m2o_test {
    uint32_t p;
    int mode, ln;
    uint64_t n, h, iter;
    uint32_t carry_pairs, nblocks;
    m2o_gf *z, *w;
    uint8_t *width, *w61, *w31;
    u128 *carry;
};

static void block_carry(m2o_test *t) {
    const uint32_t nb = t->nblocks, bp = t->carry_pairs;
    /* prepare: Garner + zero-carry-in normalisation of each block */
#pragma omp parallel for if (t->h >= PAR_MIN)
    for (uint32_t b = 0; b < nb; ++b) {
        u128 c = 0;
        uint64_t k0 = (uint64_t)b * bp, k1 = k0 + bp < t->h ? k0 + bp : t->h;
        for (uint64_t k = k0; k < k1; ++k) {
            uint64_t l0[2], l1[2];
            m2o_garner(t->z[k].s0, t->z[k].t0, l0);
            m2o_garner(t->z[k].s1, t->z[k].t1, l1);
            uint64_t d0 = digit_adc(((u128)l0[1] << 64) | l0[0], t->width[2 * k], &c);
            uint64_t d1 = digit_adc(((u128)l1[1] << 64) | l1[0], t->width[2 * k + 1], &c);
            t->z[k].s0 = d0;
            t->z[k].s1 = d1;
            t->z[k].t0 = (uint32_t)(d0 % P31);
            t->z[k].t1 = (uint32_t)(d1 % P31);
        }
        /* init: the carry-out of block b is the carry-in of block b+1 (mod nb) */
        t->carry[(b + 1 < nb) ? b + 1 : 0] = c;
    }
    /* drain: sweep the blocks until no carry remains; the top carry wraps to
     * block 0 because 2^p = 1 mod M_p */
    int any = 1;
    while (any) {
        any = 0;
        for (uint32_t b = 0; b < nb; ++b) {
            u128 c = t->carry[b];
            if (c == 0) continue;
            t->carry[b] = 0;
            any = 1;
            uint64_t k0 = (uint64_t)b * bp, k1 = k0 + bp < t->h ? k0 + bp : t->h;
            for (uint64_t k = k0; k < k1 && c != 0; ++k) {
                uint64_t d0 = digit_adc(t->z[k].s0, t->width[2 * k], &c);
                uint64_t d1 = digit_adc(t->z[k].s1, t->width[2 * k + 1], &c);
                t->z[k].s0 = d0;
                t->z[k].s1 = d1;
                t->z[k].t0 = (uint32_t)(d0 % P31);
                t->z[k].t1 = (uint32_t)(d1 % P31);
            }
            if (c != 0) t->carry[(b + 1 < nb) ? b + 1 : 0] += c;
        }
    }
}

/* Subtract a small a with borrow; a borrow out of the top digit re-enters at
 * digit 0 since 2^p = 1 mod M_p (sub_kernel :3006-3024, subtract_a :3051-3065). */
static void sub_small(uint64_t *d, const uint8_t *width, uint64_t nd, uint64_t a) {
    uint64_t borrow = a;
    while (borrow) {
        for (uint64_t j = 0; j < nd && borrow; ++j) {
            unsigned wd = width[j];
            __int128 acc = (__int128)d[j] - (__int128)borrow;
            if (wd == 0) {
                borrow = (uint64_t)(-acc);
                d[j] = 0;
                continue;
            }
            __int128 m = (__int128)1 << wd;
            __int128 r = acc % m;
            if (r < 0) r += m;
            borrow = (uint64_t)((r - acc) / m);
            d[j] = (uint64_t)r;
        }
    }
}

static void sub_ll(m2o_test *t) {
    uint64_t *d = (uint64_t *)malloc(sizeof(uint64_t) * t->n);
    for (uint64_t k = 0; k < t->h; ++k) {
        d[2 * k] = t->z[k].s0;
        d[2 * k + 1] = t->z[k].s1;
    }
    sub_small(d, t->width, t->n, 2);
    for (uint64_t k = 0; k < t->h; ++k) {
        t->z[k].s0 = d[2 * k];
        t->z[k].s1 = d[2 * k + 1];
        t->z[k].t0 = (uint32_t)(d[2 * k] % P31);
        t->z[k].t1 = (uint32_t)(d[2 * k + 1] % P31);
    }
    free(d);
}

m2o_test *m2o_test_new(uint32_t p, int mode, uint32_t carry_pairs) {
    m2o_test *t = (m2o_test *)calloc(1, sizeof(m2o_test));
    t->p = p;
    t->mode = mode;
    t->ln = m2o_transformsize(p);
    t->n = ((uint64_t)1) << t->ln;
    t->h = t->n / 2;
    t->z = (m2o_gf *)calloc(t->h, sizeof(m2o_gf));
    t->w = (m2o_gf *)calloc(t->n, sizeof(m2o_gf));
    t->width = (uint8_t *)malloc(t->n);
    t->w61 = (uint8_t *)malloc(t->n);
    t->w31 = (uint8_t *)malloc(t->n);
    if (m2o_digit_tables(p, t->ln, t->width, t->w61, t->w31) != 0) {
        m2o_test_free(t);
        return NULL;
    }
    m2o_twiddles(t->ln, t->w);
    /* clamp_carry_pairs (:3481-3487): power of two, 1 <= x <= h */
    uint32_t cp = carry_pairs ? carry_pairs : 64u, pw = 1;
    if (cp > t->h) cp = (uint32_t)t->h;
    while ((pw << 1) <= cp) pw <<= 1;
    t->carry_pairs = pw;
    t->nblocks = (uint32_t)((t->h + pw - 1) / pw);
    t->carry = (u128 *)calloc(t->nblocks, sizeof(u128));
    /* initial residue: digit 0 = 4 (LL) or 3 (PRP) in both fields (:3908) */
    uint32_t x0 = mode == 0 ? 4u : 3u;
    t->z[0].s0 = x0;
    t->z[0].t0 = x0;
    return t;
}

void m2o_test_free(m2o_test *t) {
    if (!t) return;
    free(t->z);
    free(t->w);
    free(t->width);
    free(t->w61);
    free(t->w31);
    free(t->carry);
    free(t);
}

int m2o_test_ln(const m2o_test *t) { return t->ln; }
uint64_t m2o_test_iter(const m2o_test *t) { return t->iter; }

void m2o_test_run(m2o_test *t, uint64_t iters) {
    for (uint64_t it = 0; it < iters; ++it) {
        m2o_weight(t->ln, t->z, t->w61, t->w31);
        forward_plan(t->ln, t->z, t->w);
        m2o_square_half(t->ln, t->z, t->w);
        inverse_plan(t->ln, t->z, t->w);
        m2o_unweight(t->ln, t->z, t->w61, t->w31);
        block_carry(t);
        if (t->mode == 0) sub_ll(t);
        t->iter++;
    }
}

void m2o_test_digits(const m2o_test *t, uint64_t *out) {
    for (uint64_t k = 0; k < t->h; ++k) {
        out[2 * k] = t->z[k].s0;
        out[2 * k + 1] = t->z[k].s1;
    }
}

/* R = sum_j d_j 2^ceil(pj/n) folded mod 2^p - 1 (reconstruct_value_from_digits
 * :3400-3411 and mersenne_mod :3413-3422). */
int m2o_digits_residue(uint32_t p, int ln, const uint64_t *digits, uint8_t *buf, size_t len) {
    const uint64_t n = ((uint64_t)1) << ln;
    const size_t nbytes = (p + 7u) / 8u;
    if (len < nbytes) return -1;
    const size_t nw = (p + 63u) / 64u + 3u;
    uint64_t *acc = (uint64_t *)calloc(nw, sizeof(uint64_t));
    uint64_t pos = 0;
    for (uint64_t j = 0; j < n; ++j) {
        uint64_t next = ((uint64_t)p * (j + 1) - 1) / n + 1;
        uint64_t d = digits[j];
        size_t wi = pos / 64;
        unsigned sh = pos % 64;
        u128 v = (u128)d << sh;
        u128 c = 0;
        for (size_t i = wi; (v != 0 || c != 0) && i < nw; ++i) {
            u128 sum = (u128)acc[i] + (uint64_t)v + c;
            acc[i] = (uint64_t)sum;
            c = sum >> 64;
            v >>= 64;
        }
        pos = next;
    }
    /* fold bits >= p back to bit 0 until the value is below 2^p */
    for (;;) {
        int high = 0;
        for (size_t i = p / 64; i < nw; ++i) {
            uint64_t word = acc[i];
            if (i == p / 64) word >>= (p % 64);
            if (word) high = 1;
        }
        if (!high) break;
        uint64_t *hi = (uint64_t *)calloc(nw, sizeof(uint64_t));
        for (size_t i = 0; i < nw; ++i) {
            uint64_t bit = (uint64_t)p + 64u * i;
            size_t wi = bit / 64;
            unsigned sh = bit % 64;
            if (wi >= nw) break;
            uint64_t lo = acc[wi] >> sh;
            uint64_t hiw = (sh && wi + 1 < nw) ? acc[wi + 1] << (64 - sh) : 0;
            hi[i] = lo | hiw;
        }
        acc[p / 64] &= (p % 64) ? ((((uint64_t)1) << (p % 64)) - 1) : 0;
        for (size_t i = p / 64 + 1; i < nw; ++i) acc[i] = 0;
        u128 c = 0;
        for (size_t i = 0; i < nw; ++i) {
            u128 sum = (u128)acc[i] + hi[i] + c;
            acc[i] = (uint64_t)sum;
            c = sum >> 64;
        }
        free(hi);
    }
    /* all ones (= M_p) is the other representation of 0 */
    int all_ones = 1;
    for (uint32_t b = 0; b < p && all_ones; ++b)
        if (!((acc[b / 64] >> (b % 64)) & 1)) all_ones = 0;
    if (all_ones) memset(acc, 0, nw * sizeof(uint64_t));
    memset(buf, 0, len);
    for (size_t i = 0; i < nbytes; ++i) buf[i] = (uint8_t)(acc[i / 8] >> (8 * (i % 8)));
    free(acc);
    return 0;
}

int m2o_test_residue(const m2o_test *t, uint8_t *buf, size_t len) {
    uint64_t *d = (uint64_t *)malloc(sizeof(uint64_t) * t->n);
    m2o_test_digits(t, d);
    int rc = m2o_digits_residue(t->p, t->ln, d, buf, len);
    free(d);
    return rc;
}

uint64_t m2o_test_res64(const m2o_test *t) {
    size_t len = (t->p + 7u) / 8u;
    if (len < 8) len = 8;
    uint8_t *buf = (uint8_t *)calloc(len, 1);
    m2o_test_residue(t, buf, len);
    uint64_t r = 0;
    for (int i = 7; i >= 0; --i) r = (r << 8) | buf[i];
    free(buf);
    return r;
}

/* verify_is_zero / verify_is_Mp / verify_prp_residue_9 (:3028-3076) */
int m2o_test_verdict(const m2o_test *t) {
    uint64_t *d = (uint64_t *)malloc(sizeof(uint64_t) * t->n);
    m2o_test_digits(t, d);
    if (t->mode == 1) sub_small(d, t->width, t->n, 9);
    int zero = 1, mp = 1;
    for (uint64_t j = 0; j < t->n; ++j) {
        if (d[j] != 0) zero = 0;
        uint64_t mask = t->width[j] ? ((((uint64_t)1) << t->width[j]) - 1) : 0;
        if (d[j] != mask) mp = 0;
    }
    free(d);
    return zero || mp;
}
