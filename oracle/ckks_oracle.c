/*
 * CPU oracle for the CKKS evaluation hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch plain-C restatement of the reference's algorithms
 * (TrustworthyComputing/aes-walsh-ckks).  Each function names the reference
 * file:line range whose behaviour it restates.  Used only as the checker by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; the HIP
 * product never links or calls it.
 *
 * Parity of this restatement is pinned by tests/test_oracle.py against:
 *   - golden vectors produced by the reference's own code (tests/golden/),
 *   - closed forms (brute-force negacyclic DFT, big-int CRT base extension,
 *     exact rounding division, automorphism maps),
 *   - invariants (inverse(forward(x)) == x, convolution theorem).
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "ckks_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

/* ===================================================================== */
/* Modular arithmetic: src/math/modarith/modarith.hpp:6-130, .cpp:3-40   */
/* ===================================================================== */

/* barrett_const_computation: mu = floor((2^128-1)/q)  (modarith.cpp:3-8) */
void oc_barrett_const(u64 q, u64 out[2]) {
    u128 r = (~(u128)0) / q;
    out[0] = (u64)r;
    out[1] = (u64)(r >> 64);
}

/* barrett_reduce(a*b) with one correction (modarith.hpp:16-33) */
static inline u64 barrett_mul_c(u64 a, u64 b, u64 q, u64 mu_lo, u64 mu_hi) {
    u128 z = (u128)a * b;
    u64 zl = (u64)z, zh = (u64)(z >> 64);
    u128 t = (u128)zh * mu_lo + (((u128)zl * mu_lo) >> 64);
    u64 carry = (u64)(t >> 64);
    u64 qh = (u64)((u128)zh * mu_hi + (((u128)zl * mu_hi) >> 64) + carry);
    u64 r = zl - qh * q;
    if (r >= q) r -= q;
    return r;
}

u64 oc_barrett_mul(u64 a, u64 b, u64 q) {
    u64 mu[2];
    oc_barrett_const(q, mu);
    return barrett_mul_c(a, b, q, mu[0], mu[1]);
}

/* barrett_reduce_u64 (modarith.hpp:53-58) */
static inline u64 barrett_u64_c(u64 a, u64 q, u64 mu_hi) {
    u64 qh = (u64)(((u128)a * mu_hi) >> 64);
    u64 r = a - qh * q;
    if (r >= q) r -= q;
    return r;
}

u64 oc_barrett_reduce_u64(u64 a, u64 q) {
    u64 mu[2];
    oc_barrett_const(q, mu);
    return barrett_u64_c(a, q, mu[1]);
}

/* compute_shoup: floor(w*2^64/q) (modarith.cpp:10-13) */
u64 oc_compute_shoup(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }

/* mul_mod_shoup with final correction (modarith.hpp:90-96) */
static inline u64 shoup_mul_c(u64 a, u64 w, u64 ws, u64 q) {
    u64 u = (u64)(((u128)a * ws) >> 64);
    u64 r = a * w - u * q;
    if (r >= q) r -= q;
    return r;
}

/* Shoup product without the final correction, result in [0, 2q)
 * (the lazy butterflies of ntt.cpp:176-188, 209-218). */
static inline u64 shoup_mul_lazy(u64 a, u64 w, u64 ws, u64 q) {
    u64 u = (u64)(((u128)a * ws) >> 64);
    return a * w - u * q;
}

u64 oc_shoup_mul(u64 a, u64 w, u64 q) { return shoup_mul_c(a, w, oc_compute_shoup(w, q), q); }

/* reduce_u64_mod_q_shoup (modarith.hpp:35-42) */
static inline u64 reduce_u64_shoup(u64 x, u64 q, u64 one_shoup) {
    u64 u = (u64)(((u128)x * one_shoup) >> 64);
    u64 r = x - u * q;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return r;
}

/* reduce_mod_4q (modarith.hpp:44-50) */
static inline u64 reduce_4q(u64 x, u64 q) {
    u64 tq = q << 1;
    if (x >= tq) x -= tq;
    if (x >= q) x -= q;
    return x;
}

/* add_mod_q / sub_mod_q (modarith.hpp:65-78) */
static inline u64 add_q(u64 a, u64 b, u64 q) {
    u64 r = a + b;
    if (r >= q) r -= q;
    return r;
}
static inline u64 sub_q(u64 a, u64 b, u64 q) {
    u64 r = a - b;
    u64 mask = (u64)0 - (u64)(a < b);
    return r + (mask & q);
}

/* reduce_i64_mod_q (modarith.hpp:80-87) */
static inline u64 reduce_i64(int64_t v, u64 q) {
    int64_t qi = (int64_t)q;
    int64_t r = v % qi;
    if (r < 0) r += qi;
    return (u64)r;
}

/* Montgomery, R = 2^64 (modarith.hpp:101-130) */
u64 oc_montgomery_neg_inverse(u64 q) {
    u64 inv = 1;
    for (int i = 0; i < 6; ++i) inv *= (u64)2 - q * inv;
    return (u64)0 - inv;
}
u64 oc_montgomery_radix_mod_q(u64 q) { return (u64)((((u128)1) << 64) % q); }

static inline u64 mont_mul_c(u64 normal, u64 mont, u64 q, u64 qninv) {
    u128 t = (u128)normal * mont;
    u64 m = (u64)t * qninv;
    u128 s = t + (u128)m * q;
    /* The 128-bit sum wraps only if t + m*q >= 2^128; the reference computes
     * (t + m*q) >> 64 in 128-bit arithmetic as well. */
    u64 u = (u64)(s >> 64);
    if (u >= q) u -= q;
    return u;
}
u64 oc_montgomery_mul(u64 normal, u64 mont, u64 q) {
    return mont_mul_c(normal, mont, q, oc_montgomery_neg_inverse(q));
}

/* powmod via naive 128-bit products (modarith.cpp:15-27) */
u64 oc_powmod(u64 a, u64 b, u64 m) {
    a %= m;
    u64 res = 1;
    while (b > 0) {
        if (b & 1) res = (u64)((u128)res * a % m);
        a = (u64)((u128)a * a % m);
        b >>= 1;
    }
    return res;
}

/* mod_inverse: extended Euclid (modarith.cpp:30-40).  For coprime inputs the
 * inverse in [0, m) is unique, so any exact method gives the reference value. */
u64 oc_mod_inverse(u64 a, u64 m) {
    if (m == 1) return 0;
    __int128 t = 0, newt = 1;
    __int128 r = m, newr = a % m;
    while (newr != 0) {
        __int128 qq = r / newr;
        __int128 tmp = t - qq * newt; t = newt; newt = tmp;
        tmp = r - qq * newr; r = newr; newr = tmp;
    }
    if (t < 0) t += m;
    return (u64)t;
}

/* ===================================================================== */
/* Prime search: src/util/random.cpp:8-131                                */
/* ===================================================================== */

static u64 mr_pow(u64 b, u64 e, u64 n) {
    u64 r = 1;
    b %= n;
    while (e) {
        if (e & 1) r = (u64)((u128)r * b % n);
        b = (u64)((u128)b * b % n);
        e >>= 1;
    }
    return r;
}

/* is_prime: Miller-Rabin with the 12 bases 2..37 (random.cpp:41-64) */
int oc_is_prime(u64 n) {
    if (n < 2) return 0;
    int s = 0;
    u64 d = n - 1;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    static const u64 bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (int i = 0; i < 12; ++i) {
        u64 a = bases[i];
        if (n == a) return 1;
        u64 x = mr_pow(a, d, n);
        if (x == 1 || x == n - 1) continue;
        int composite = 1;
        for (int r = 1; r < s; ++r) {
            x = (u64)((u128)x * x % n);
            if (x == n - 1) { composite = 0; break; }
        }
        if (composite) return 0;
    }
    return 1;
}

/* find_p_and_q_primes: P bit sizes are served first, then Q; one descending
 * cursor k per bit size starting at 2^bits/(2N); candidate q = 2kN+1; a prime
 * already taken is skipped (random.cpp:66-131). */
int oc_find_p_and_q_primes(const u64* p_bits, int np, const u64* q_bits, int nq, u64 N,
                           u64* p_out, u64* q_out) {
    u64 cursor_bits[128];
    u64 cursor_k[128];
    int ncur = 0;
    u64 taken[512];
    int ntaken = 0;
    for (int idx = 0; idx < np + nq; ++idx) {
        u64 bits = idx < np ? p_bits[idx] : q_bits[idx - np];
        if (bits < 2 || bits > 63) return -1;
        int c = -1;
        for (int j = 0; j < ncur; ++j)
            if (cursor_bits[j] == bits) c = j;
        if (c < 0) {
            if (ncur >= 128) return -1;
            c = ncur++;
            cursor_bits[c] = bits;
            cursor_k[c] = ((u64)1 << bits) / (2 * N);
        }
        const u64 start = ((u64)1 << (bits - 1)) / (2 * N);
        u64 found = 0;
        while (cursor_k[c]-- > start) {
            u64 cand = cursor_k[c] * 2 * N + 1;
            int dup = 0;
            for (int j = 0; j < ntaken; ++j)
                if (taken[j] == cand) dup = 1;
            if (dup) continue;
            if (oc_is_prime(cand)) { found = cand; break; }
        }
        if (!found || ntaken >= 512) return -2;
        taken[ntaken++] = found;
        if (idx < np) p_out[idx] = found; else q_out[idx - np] = found;
    }
    return 0;
}

/* ===================================================================== */
/* Twiddles and NTT: src/math/ntt/ntt.cpp:28-447                          */
/* ===================================================================== */

static u64 brev(u64 x, int bits) {
    u64 r = 0;
    for (int i = 0; i < bits; ++i) if (x & ((u64)1 << i)) r |= (u64)1 << (bits - 1 - i);
    return r;
}
static int ilog2u(u64 x) { int l = 0; while (((u64)1 << l) < x) ++l; return l; }

/* find_root_of_unity: psi = first a >= 2 with (a^((q-1)/2N))^N == -1 (ntt.cpp:28-42) */
u64 oc_find_root_of_unity(u64 N, u64 q) {
    if ((q - 1) % (2 * N) != 0) return 0;
    u64 e = (q - 1) / (2 * N);
    for (u64 a = 2; a <= q - 2; ++a) {
        u64 r = oc_powmod(a, e, q);
        if (oc_powmod(r, N, q) == q - 1) return r;
    }
    return 0;
}

/* generate_negacyclic_ntt_twiddles (ntt.cpp:54-148): forward table entry
 * [m+i] = omega^(brev_m(i)*N/2m) * psi^(N/2m); inverse entry [m+i] =
 * omega^-(brev_m(i)*N/2m) * psi^-(N/2m); Shoup companions; N^-1. */
int oc_ntt_tables(u64 N, u64 q, u64* fwd, u64* fwd_shoup, u64* inv, u64* inv_shoup, u64* n_inv) {
    u64 psi = oc_find_root_of_unity(N, q);
    if (!psi) return -1;
    u64 w = (u64)((u128)psi * psi % q);
    u64 half = N >> 1;
    u64* stage = (u64*)malloc(sizeof(u64) * (half ? half : 1));
    u64* psi_pow = (u64*)malloc(sizeof(u64) * N);
    u64* psi_ipow = (u64*)malloc(sizeof(u64) * N);
    psi_pow[0] = 1;
    for (u64 i = 1; i < N; ++i) psi_pow[i] = (u64)((u128)psi_pow[i - 1] * psi % q);
    u64 psi_inv = oc_powmod(psi, q - 2, q);
    psi_ipow[0] = 1;
    for (u64 i = 1; i < N; ++i) psi_ipow[i] = (u64)((u128)psi_ipow[i - 1] * psi_inv % q);

    for (u64 i = 0; i < N; ++i) { fwd[i] = 1; inv[i] = 1; }
    stage[0] = 1;
    for (u64 i = 1; i < half; ++i) stage[i] = (u64)((u128)stage[i - 1] * w % q);
    for (u64 m = 1; m < N; m <<= 1) {
        u64 step = N / (m << 1);
        int sb = ilog2u(m);
        for (u64 i = 0; i < m; ++i) fwd[m + i] = stage[brev(i, sb) * step];
    }
    for (u64 m = 1, k = N >> 1; m < N; m <<= 1, k >>= 1)
        for (u64 i = 0; i < m; ++i) fwd[m + i] = (u64)((u128)fwd[m + i] * psi_pow[k] % q);

    u64 w_inv = oc_powmod(w, q - 2, q);
    stage[0] = 1;
    for (u64 i = 1; i < half; ++i) stage[i] = (u64)((u128)stage[i - 1] * w_inv % q);
    for (u64 m = 1; m < N; m <<= 1) {
        u64 step = N / (m << 1);
        int sb = ilog2u(m);
        for (u64 i = 0; i < m; ++i) inv[m + i] = stage[brev(i, sb) * step];
    }
    for (u64 m = N >> 1, k = 1; m > 0; m >>= 1, k <<= 1)
        for (u64 i = 0; i < m; ++i) inv[m + i] = (u64)((u128)inv[m + i] * psi_ipow[k] % q);

    for (u64 i = 0; i < N; ++i) {
        fwd_shoup[i] = oc_compute_shoup(fwd[i], q);
        inv_shoup[i] = oc_compute_shoup(inv[i], q);
    }
    u64 ni = oc_powmod(N, q - 2, q);
    inv_shoup[N] = oc_compute_shoup(ni, q);
    *n_inv = ni;
    free(stage);
    free(psi_pow);
    free(psi_ipow);
    return 0;
}

/* ntt_forward_dit2_lazy / ntt_forward_dit2 (ntt.cpp:235-297): Harvey DIT,
 * natural -> bit-reversed, lazy output in [0,4q), else reduced to [0,q). */
void oc_ntt_forward(u64* a, u64 N, u64 q, const u64* tab, const u64* tabs, int lazy) {
    const u64 tq = q << 1;
    for (u64 m = 1, k = N >> 1; m < N; m <<= 1, k >>= 1) {
        for (u64 i = 0; i < m; ++i) {
            const u64 w = tab[m + i], ws = tabs[m + i];
            u64* x0 = a + 2 * i * k;
            u64* x1 = x0 + k;
            for (u64 j = 0; j < k; ++j) {
                u64 t = x0[j];
                if (t >= tq) t -= tq;
                u64 u = shoup_mul_lazy(x1[j], w, ws, q);
                x0[j] = t + u;
                x1[j] = t + tq - u;
            }
        }
    }
    if (!lazy)
        for (u64 i = 0; i < N; ++i) a[i] = reduce_4q(a[i], q);
}

/* ntt_inverse_dif2(_lazy) (ntt.cpp:331-386): Gentleman-Sande DIF,
 * bit-reversed -> natural, then x N^-1 (Shoup).  Both variants end in [0,q). */
void oc_ntt_inverse(u64* a, u64 N, u64 q, const u64* tab, const u64* tabs, u64 n_inv, int lazy) {
    const u64 tq = q << 1;
    for (u64 m = N >> 1, k = 1; m > 0; m >>= 1, k <<= 1) {
        for (u64 i = 0; i < m; ++i) {
            u64* x0 = a + 2 * i * k;
            u64* x1 = x0 + k;
            const u64 w = tab[m + i], ws = tabs[m + i];
            for (u64 j = 0; j < k; ++j) {
                const u64 t = x0[j], u = x1[j];
                u64 s = t + u;
                if (s >= tq) s -= tq;
                x0[j] = s;
                x1[j] = shoup_mul_lazy(t - u + tq, w, ws, q);
            }
        }
    }
    (void)lazy;
    for (u64 i = 0; i < N; ++i) a[i] = reduce_4q(shoup_mul_c(a[i], n_inv, tabs[N], q), q);
}

/* ===================================================================== */
/* Context                                                                */
/* ===================================================================== */

typedef struct {
    u64 q;
    u64 *fwd, *fwd_shoup, *inv, *inv_shoup;
    u64 n_inv;
    u64 mu_lo, mu_hi;
    u64 mont_ninv;
} oc_tables;

typedef struct {
    int ns, nd;
    u64 *src, *dst;
    u64 *qhat_inv, *qhat_inv_shoup;
    double* src_inv;
    u64 *qhat_mod_dst, *qhat_mod_dst_shoup; /* [dst][src] */
    u64* alpha_corr;                        /* [dst][alpha 0..ns] */
} oc_modup_t;

typedef struct {
    u64 *one_shoup, *inv_p, *inv_p_shoup, *half_ntt; /* d entries, half_ntt d*N */
} oc_rescale_t;

struct oc_ctx {
    u64 N;
    int logN;
    int nq, np, dnum, alpha;
    u64 *q, *p;
    oc_tables* tq;
    oc_tables* tp;
    oc_modup_t** modup; /* [level][part] */
    int* nparts;        /* [level] */
    u64** fold_pinv;    /* [level] (level+1)*np */
    u64** fold_pinv_shoup;
    oc_rescale_t* rc;   /* [d] for d >= 1 */
};

static void tables_init(oc_tables* t, u64 N, u64 q) {
    t->q = q;
    t->fwd = (u64*)malloc(8 * N);
    t->fwd_shoup = (u64*)malloc(8 * N);
    t->inv = (u64*)malloc(8 * N);
    t->inv_shoup = (u64*)malloc(8 * (N + 1));
    oc_ntt_tables(N, q, t->fwd, t->fwd_shoup, t->inv, t->inv_shoup, &t->n_inv);
    u64 mu[2];
    oc_barrett_const(q, mu);
    t->mu_lo = mu[0];
    t->mu_hi = mu[1];
    t->mont_ninv = oc_montgomery_neg_inverse(q);
}

static void tables_free(oc_tables* t) {
    free(t->fwd); free(t->fwd_shoup); free(t->inv); free(t->inv_shoup);
}

/* build_approx_crt_modup_precomp (relinearization_key_generation.cpp:45-129) */
static void modup_build(oc_modup_t* m, const u64* src, int ns, const u64* dst, int nd) {
    m->ns = ns;
    m->nd = nd;
    m->src = (u64*)malloc(8 * ns);
    m->dst = (u64*)malloc(8 * (nd ? nd : 1));
    memcpy(m->src, src, 8 * ns);
    if (nd) memcpy(m->dst, dst, 8 * nd);
    m->qhat_inv = (u64*)malloc(8 * ns);
    m->qhat_inv_shoup = (u64*)malloc(8 * ns);
    m->src_inv = (double*)malloc(sizeof(double) * ns);
    m->qhat_mod_dst = (u64*)malloc(8 * (nd * ns + 1));
    m->qhat_mod_dst_shoup = (u64*)malloc(8 * (nd * ns + 1));
    m->alpha_corr = (u64*)malloc(8 * (nd * (ns + 1) + 1));
    for (int i = 0; i < ns; ++i) {
        u64 qi = src[i];
        m->src_inv[i] = 1.0 / (double)qi;
        u64 qh = 1;
        for (int j = 0; j < ns; ++j)
            if (j != i) qh = oc_barrett_mul(qh, src[j] % qi, qi);
        m->qhat_inv[i] = oc_mod_inverse(qh, qi);
        m->qhat_inv_shoup[i] = oc_compute_shoup(m->qhat_inv[i], qi);
        for (int t = 0; t < nd; ++t) {
            u64 mm = dst[t];
            u64 v = 1;
            for (int j = 0; j < ns; ++j)
                if (j != i) v = oc_barrett_mul(v, src[j] % mm, mm);
            m->qhat_mod_dst[t * ns + i] = v;
            m->qhat_mod_dst_shoup[t * ns + i] = oc_compute_shoup(v, mm);
        }
    }
    for (int t = 0; t < nd; ++t) {
        u64 mm = dst[t];
        u64 prod = 1;
        for (int i = 0; i < ns; ++i) prod = oc_barrett_mul(prod, src[i] % mm, mm);
        u64 corr = 0;
        m->alpha_corr[t * (ns + 1)] = 0;
        for (int a = 1; a <= ns; ++a) {
            corr = add_q(corr, prod, mm);
            m->alpha_corr[t * (ns + 1) + a] = corr;
        }
    }
}

static void modup_free(oc_modup_t* m) {
    free(m->src); free(m->dst); free(m->qhat_inv); free(m->qhat_inv_shoup); free(m->src_inv);
    free(m->qhat_mod_dst); free(m->qhat_mod_dst_shoup); free(m->alpha_corr);
}

static int active_parts(int limbs, int alpha, int dnum) {
    int p = (limbs + alpha - 1) / alpha;
    return p < dnum ? p : dnum;
}

oc_ctx* oc_ctx_create(u64 N, const u64* q, int nq, const u64* p, int np, int dnum) {
    if (nq < 1 || np < 1 || dnum < 1 || dnum > nq) return NULL;
    oc_ctx* c = (oc_ctx*)calloc(1, sizeof(oc_ctx));
    c->N = N;
    c->logN = ilog2u(N);
    c->nq = nq;
    c->np = np;
    c->dnum = dnum;
    c->alpha = (nq + dnum - 1) / dnum; /* ckks_context.cpp:38-47 */
    c->q = (u64*)malloc(8 * nq);
    c->p = (u64*)malloc(8 * np);
    memcpy(c->q, q, 8 * nq);
    memcpy(c->p, p, 8 * np);
    c->tq = (oc_tables*)calloc(nq, sizeof(oc_tables));
    c->tp = (oc_tables*)calloc(np, sizeof(oc_tables));
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nq + np; ++i) {
        if (i < nq) tables_init(&c->tq[i], N, q[i]);
        else tables_init(&c->tp[i - nq], N, p[i - nq]);
    }

    /* build_modup_precompute_material (ckks_context.cpp:468-563) */
    c->modup = (oc_modup_t**)calloc(nq, sizeof(oc_modup_t*));
    c->nparts = (int*)calloc(nq, sizeof(int));
    u64* dst = (u64*)malloc(8 * (nq + np));
    for (int level = 0; level < nq; ++level) {
        int L = level + 1;
        int parts = active_parts(L, c->alpha, dnum);
        c->nparts[level] = parts;
        c->modup[level] = (oc_modup_t*)calloc(parts, sizeof(oc_modup_t));
        for (int part = 0; part < parts; ++part) {
            int s0 = part * c->alpha;
            int s1 = s0 + c->alpha < L ? s0 + c->alpha : L;
            int nd = 0;
            for (int i = 0; i < s0; ++i) dst[nd++] = q[i];
            for (int i = s1; i < L; ++i) dst[nd++] = q[i];
            for (int i = 0; i < np; ++i) dst[nd++] = p[i];
            modup_build(&c->modup[level][part], q + s0, s1 - s0, dst, nd);
        }
    }
    free(dst);

    /* folded P^-1 tables (ckks_context.cpp:537-560) */
    c->fold_pinv = (u64**)calloc(nq, sizeof(u64*));
    c->fold_pinv_shoup = (u64**)calloc(nq, sizeof(u64*));
    for (int level = 0; level < nq; ++level) {
        int L = level + 1;
        c->fold_pinv[level] = (u64*)malloc(8 * L * np);
        c->fold_pinv_shoup[level] = (u64*)malloc(8 * L * np);
        for (int qi = 0; qi < L; ++qi)
            for (int pi = 0; pi < np; ++pi) {
                u64 inv = oc_mod_inverse(p[pi] % q[qi], q[qi]);
                c->fold_pinv[level][qi * np + pi] = inv;
                c->fold_pinv_shoup[level][qi * np + pi] = oc_compute_shoup(inv, q[qi]);
            }
    }

    /* build_rescale_constants (ckks_context.cpp:565-599) */
    c->rc = (oc_rescale_t*)calloc(nq, sizeof(oc_rescale_t));
    for (int d = 1; d < nq; ++d) {
        u64 pp = q[d], half = pp >> 1;
        oc_rescale_t* r = &c->rc[d];
        r->one_shoup = (u64*)malloc(8 * d);
        r->inv_p = (u64*)malloc(8 * d);
        r->inv_p_shoup = (u64*)malloc(8 * d);
        r->half_ntt = (u64*)malloc(8 * d * N);
        for (int i = 0; i < d; ++i) {
            u64 qq = q[i];
            u64 os = oc_compute_shoup(1, qq);
            u64 pm = reduce_u64_shoup(pp, qq, os);
            u64 ip = oc_mod_inverse(pm, qq);
            u64 hm = reduce_u64_shoup(half, qq, os);
            u64* h = r->half_ntt + (u64)i * N;
            for (u64 j = 0; j < N; ++j) h[j] = hm;
            oc_ntt_forward(h, N, qq, c->tq[i].fwd, c->tq[i].fwd_shoup, 0);
            r->one_shoup[i] = os;
            r->inv_p[i] = ip;
            r->inv_p_shoup[i] = oc_compute_shoup(ip, qq);
        }
    }
    return c;
}

void oc_ctx_destroy(oc_ctx* c) {
    if (!c) return;
    for (int i = 0; i < c->nq; ++i) tables_free(&c->tq[i]);
    for (int i = 0; i < c->np; ++i) tables_free(&c->tp[i]);
    for (int l = 0; l < c->nq; ++l) {
        for (int part = 0; part < c->nparts[l]; ++part) modup_free(&c->modup[l][part]);
        free(c->modup[l]);
        free(c->fold_pinv[l]);
        free(c->fold_pinv_shoup[l]);
    }
    for (int d = 1; d < c->nq; ++d) {
        free(c->rc[d].one_shoup); free(c->rc[d].inv_p); free(c->rc[d].inv_p_shoup);
        free(c->rc[d].half_ntt);
    }
    free(c->rc);
    free(c->modup); free(c->nparts); free(c->fold_pinv); free(c->fold_pinv_shoup);
    free(c->tq); free(c->tp); free(c->q); free(c->p);
    free(c);
}

int oc_ctx_alpha(const oc_ctx* c) { return c->alpha; }

int oc_omp_threads(void) { return omp_get_max_threads(); }

static const oc_tables* basis_table(const oc_ctx* c, int idx) {
    return idx < c->nq ? &c->tq[idx] : &c->tp[idx - c->nq];
}

/* ntt_forward_rns_flat_inplace / ntt_inverse_rns_flat_inplace (ntt.cpp:416-447) */
void oc_ntt_fwd_q(const oc_ctx* c, u64* data, int limbs, int lazy) {
#pragma omp parallel for schedule(static) if (limbs > 1)
    for (int l = 0; l < limbs; ++l)
        oc_ntt_forward(data + (u64)l * c->N, c->N, c->q[l], c->tq[l].fwd, c->tq[l].fwd_shoup, lazy);
}
void oc_ntt_inv_q(const oc_ctx* c, u64* data, int limbs, int lazy) {
#pragma omp parallel for schedule(static) if (limbs > 1)
    for (int l = 0; l < limbs; ++l)
        oc_ntt_inverse(data + (u64)l * c->N, c->N, c->q[l], c->tq[l].inv, c->tq[l].inv_shoup,
                       c->tq[l].n_inv, lazy);
}
void oc_ntt_fwd_basis(const oc_ctx* c, u64* data, const int* basis, int limbs, int lazy) {
#pragma omp parallel for schedule(static) if (limbs > 1)
    for (int l = 0; l < limbs; ++l) {
        const oc_tables* t = basis_table(c, basis[l]);
        oc_ntt_forward(data + (u64)l * c->N, c->N, t->q, t->fwd, t->fwd_shoup, lazy);
    }
}
void oc_ntt_inv_basis(const oc_ctx* c, u64* data, const int* basis, int limbs, int lazy) {
#pragma omp parallel for schedule(static) if (limbs > 1)
    for (int l = 0; l < limbs; ++l) {
        const oc_tables* t = basis_table(c, basis[l]);
        oc_ntt_inverse(data + (u64)l * c->N, c->N, t->q, t->inv, t->inv_shoup, t->n_inv, lazy);
    }
}

/* ===================================================================== */
/* HPS base extension: ckks_keyswitch.cpp:278-332 (+ wrapper 375-391)     */
/* ===================================================================== */

static void hps_modup(const oc_modup_t* m, const u64* const* src, u64* const* dst, u64 N) {
    u64 y[64];
    for (u64 c = 0; c < N; ++c) {
        double est = 0.0;
        for (int i = 0; i < m->ns; ++i) {
            u64 qi = m->src[i];
            u64 x = reduce_4q(src[i][c], qi);
            u64 s = shoup_mul_c(x, m->qhat_inv[i], m->qhat_inv_shoup[i], qi);
            y[i] = s;
            est += (double)s * m->src_inv[i];
        }
        u64 al = (u64)(est + 0.5);
        if (al > (u64)m->ns) al = (u64)m->ns;
        for (int t = 0; t < m->nd; ++t) {
            u64 mm = m->dst[t];
            u64 acc = 0;
            const u64* row = m->qhat_mod_dst + t * m->ns;
            const u64* rows = m->qhat_mod_dst_shoup + t * m->ns;
            for (int i = 0; i < m->ns; ++i) acc = add_q(acc, shoup_mul_c(y[i], row[i], rows[i], mm), mm);
            dst[t][c] = sub_q(acc, m->alpha_corr[t * (m->ns + 1) + al], mm);
        }
    }
}

int oc_modup(const oc_ctx* c, int level, int part, const u64* src, u64* dst) {
    if (level < 0 || level >= c->nq || part < 0 || part >= c->nparts[level]) return -1;
    const oc_modup_t* m = &c->modup[level][part];
    const u64* sp[64];
    u64* dp[256];
    for (int i = 0; i < m->ns; ++i) sp[i] = src + (u64)i * c->N;
    for (int t = 0; t < m->nd; ++t) dp[t] = dst + (u64)t * c->N;
    hps_modup(m, sp, dp, c->N);
    return 0;
}

/* ===================================================================== */
/* Folded ModDown pair: ckks_keyswitch.cpp:521-607                        */
/* ===================================================================== */

static void folded_moddown_pair(const oc_ctx* c, int L, const u64* src_a_p, u64* dst_a,
                                const u64* src_b_p, u64* dst_b) {
    const u64 N = c->N;
    const int K = c->np;
    const u64* invs_all = c->fold_pinv[L - 1];
    const u64* invs_sh_all = c->fold_pinv_shoup[L - 1];
    double pinv[64];
    for (int j = 0; j < K; ++j) pinv[j] = 1.0 / (double)c->p[j];
#pragma omp parallel for schedule(static) if (L > 1)
    for (int qi = 0; qi < L; ++qi) {
        const u64 q = c->q[qi];
        const u64* invs = invs_all + qi * K;
        const u64* inv_sh = invs_sh_all + qi * K;
        u64* ca = (u64*)malloc(8 * N);
        u64* cb = (u64*)malloc(8 * N);
        for (u64 x = 0; x < N; ++x) {
            u64 sa = 0, sb = 0;
            double ea = 0.0, eb = 0.0;
            for (int j = 0; j < K; ++j) {
                u64 pj = c->p[j];
                u64 va = reduce_4q(src_a_p[(u64)j * N + x], pj);
                u64 vb = reduce_4q(src_b_p[(u64)j * N + x], pj);
                ea += (double)va * pinv[j];
                eb += (double)vb * pinv[j];
                u64 ta = shoup_mul_c(barrett_u64_c(va, q, c->tq[qi].mu_hi), invs[j], inv_sh[j], q);
                u64 tb = shoup_mul_c(barrett_u64_c(vb, q, c->tq[qi].mu_hi), invs[j], inv_sh[j], q);
                sa = add_q(sa, ta, q);
                sb = add_q(sb, tb, q);
            }
            u64 aa = (u64)(ea + 0.5), ab = (u64)(eb + 0.5);
            if (aa > (u64)K) aa = (u64)K;
            if (ab > (u64)K) ab = (u64)K;
            ca[x] = sub_q(sa, aa, q);
            cb[x] = sub_q(sb, ab, q);
        }
        oc_ntt_forward(ca, N, q, c->tq[qi].fwd, c->tq[qi].fwd_shoup, 0);
        oc_ntt_forward(cb, N, q, c->tq[qi].fwd, c->tq[qi].fwd_shoup, 0);
        u64* da = dst_a + (u64)qi * N;
        u64* db = dst_b + (u64)qi * N;
        for (u64 x = 0; x < N; ++x) {
            da[x] = sub_q(da[x], ca[x], q);
            db[x] = sub_q(db[x], cb[x], q);
        }
        free(ca);
        free(cb);
    }
}

/* ===================================================================== */
/* Hybrid key switch: ckks_keyswitch.cpp:1330-1641                        */
/* ===================================================================== */

/* Montgomery-key MAC (ckks_keyswitch.cpp:424-441, 445-461) or the normal-key
 * pair with input_mod_factor 4 (464-480, via ntt.cpp:520-536). */
static void mac_limb(const u64* digit, const u64* key, u64* acc, u64 N, const oc_tables* t,
                     int montgomery) {
    const u64 q = t->q;
    if (montgomery) {
        for (u64 i = 0; i < N; ++i) acc[i] = add_q(acc[i], mont_mul_c(digit[i], key[i], q, t->mont_ninv), q);
    } else {
        for (u64 i = 0; i < N; ++i) {
            u64 a = reduce_4q(digit[i], q), b = reduce_4q(key[i], q);
            acc[i] = add_q(acc[i], barrett_mul_c(a, b, q, t->mu_lo, t->mu_hi), q);
        }
    }
}

int oc_keyswitch(const oc_ctx* c, int level, u64* key_term, const u64* keys, u64* b_out,
                 u64* a_out, int b_mode, int a_mode, int montgomery) {
    if (level < 0 || level >= c->nq) return -1;
    const u64 N = c->N;
    const int L = level + 1, K = c->np, T = L + K;
    const int KL = c->nq + K; /* key limbs: Lmax+1+K */
    const int parts = c->nparts[level];
    const int alpha = c->alpha;

    u64* eval = (u64*)malloc(8 * N * L);
    memcpy(eval, key_term, 8 * N * L);
    oc_ntt_inv_q(c, key_term, L, 1); /* :1401 */

    u64* acc_b = (u64*)calloc(N * T, 8);
    u64* acc_a = (u64*)calloc(N * T, 8);
    u64* digit = (u64*)malloc(8 * N * T);
    for (int part = 0; part < parts; ++part) {
        const int s0 = part * alpha;
        const int s1 = s0 + alpha < L ? s0 + alpha : L;
        const u64* sp[64];
        u64* dp[256];
        int nd = 0;
        for (int i = 0; i < s1 - s0; ++i) sp[i] = key_term + (u64)(s0 + i) * N;
        for (int i = 0; i < s0; ++i) dp[nd++] = digit + (u64)i * N;
        for (int i = s1; i < L; ++i) dp[nd++] = digit + (u64)i * N;
        for (int i = 0; i < K; ++i) dp[nd++] = digit + (u64)(L + i) * N;
        if (nd) hps_modup(&c->modup[level][part], sp, dp, N); /* :1452 */

        const u64* kb = keys + (u64)part * 2 * KL * N;
        const u64* ka = kb + (u64)KL * N;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < L; ++i) {
            const u64* dptr = eval + (u64)i * N;
            if (!(i >= s0 && i < s1)) {
                u64* mptr = digit + (u64)i * N;
                oc_ntt_forward(mptr, N, c->q[i], c->tq[i].fwd, c->tq[i].fwd_shoup, 1);
                dptr = mptr;
            }
            mac_limb(dptr, ka + (u64)i * N, acc_a + (u64)i * N, N, &c->tq[i], montgomery);
            mac_limb(dptr, kb + (u64)i * N, acc_b + (u64)i * N, N, &c->tq[i], montgomery);
        }
#pragma omp parallel for schedule(static)
        for (int pi = 0; pi < K; ++pi) {
            u64* dptr = digit + (u64)(L + pi) * N;
            oc_ntt_forward(dptr, N, c->p[pi], c->tp[pi].fwd, c->tp[pi].fwd_shoup, 1);
            const u64 key_idx = (u64)c->nq + pi; /* key_q_limb_count + pi (:1510) */
            mac_limb(dptr, ka + key_idx * N, acc_a + (u64)(L + pi) * N, N, &c->tp[pi], montgomery);
            mac_limb(dptr, kb + key_idx * N, acc_b + (u64)(L + pi) * N, N, &c->tp[pi], montgomery);
        }
    }
    /* INTT of the P limbs (:1571-1579), then folded ModDown (:1580-1590) */
    for (int pi = 0; pi < K; ++pi) {
        const oc_tables* t = &c->tp[pi];
        oc_ntt_inverse(acc_a + (u64)(L + pi) * N, N, t->q, t->inv, t->inv_shoup, t->n_inv, 1);
        oc_ntt_inverse(acc_b + (u64)(L + pi) * N, N, t->q, t->inv, t->inv_shoup, t->n_inv, 1);
    }
    folded_moddown_pair(c, L, acc_a + (u64)L * N, acc_a, acc_b + (u64)L * N, acc_b);
    /* combine (:1624-1640) */
    for (int i = 0; i < L; ++i) {
        const u64 q = c->q[i];
        for (u64 x = 0; x < N; ++x) {
            u64 o = (u64)i * N + x;
            b_out[o] = b_mode == 0 ? add_q(b_out[o], acc_b[o], q) : acc_b[o];
            a_out[o] = a_mode == 0 ? add_q(a_out[o], acc_a[o], q) : acc_a[o];
        }
    }
    free(eval);
    free(acc_b);
    free(acc_a);
    free(digit);
    return 0;
}

/* ===================================================================== */
/* Automorphisms: rotation_key_generation.cpp:65-90; keyswitch.cpp:37-60  */
/* ===================================================================== */

uint64_t oc_rotation_galois_element(u64 N, int64_t rotation) {
    int64_t slots = (int64_t)(N >> 1);
    int64_t r = rotation % slots;
    if (r < 0) r += slots;
    return oc_powmod(5, (u64)r, N << 1);
}

/* build_ntt_automorphism_map: dst_to_src[brev(d)] = brev(((2d+1)g mod 2N - 1)/2) */
void oc_ntt_automorphism_map(u64 N, u64 g, u64* map) {
    int logN = ilog2u(N);
    u64 two_n = N << 1;
    for (u64 d = 0; d < N; ++d) {
        u64 e = (u64)(((u128)(2 * d + 1) * g) % two_n);
        u64 s = (e - 1) >> 1;
        map[brev(d, logN)] = brev(s, logN);
    }
}

void oc_apply_automorphism(u64* data, u64 N, int limbs, const u64* map) {
    u64* tmp = (u64*)malloc(8 * N);
    for (int l = 0; l < limbs; ++l) {
        u64* p = data + (u64)l * N;
        for (u64 i = 0; i < N; ++i) tmp[i] = p[map[i]];
        memcpy(p, tmp, 8 * N);
    }
    free(tmp);
}

/* ckks_rotate_inplace (ckks_rotation.cpp:41-96): switch, then permute. */
int oc_rotate(const oc_ctx* c, int level, u64* b, u64* a, int64_t rotation, const u64* keys,
              int montgomery) {
    int64_t slots = (int64_t)(c->N >> 1);
    int64_t r = rotation % slots;
    if (r < 0) r += slots;
    if (r == 0) return 0;
    int rc = oc_keyswitch(c, level, a, keys, b, a, 0, 1, montgomery);
    if (rc) return rc;
    u64* map = (u64*)malloc(8 * c->N);
    oc_ntt_automorphism_map(c->N, oc_powmod(5, (u64)r, c->N << 1), map);
    oc_apply_automorphism(b, c->N, level + 1, map);
    oc_apply_automorphism(a, c->N, level + 1, map);
    free(map);
    return 0;
}

/* ckks_conjugate_inplace (ckks_conjugate.cpp:16-61): permute, then switch. */
int oc_conjugate(const oc_ctx* c, int level, u64* b, u64* a, const u64* keys, int montgomery) {
    u64* map = (u64*)malloc(8 * c->N);
    oc_ntt_automorphism_map(c->N, 2 * c->N - 1, map);
    oc_apply_automorphism(b, c->N, level + 1, map);
    oc_apply_automorphism(a, c->N, level + 1, map);
    free(map);
    return oc_keyswitch(c, level, a, keys, b, a, 0, 1, montgomery);
}

/* ckks_relin_hybrid_inplace (ckks_relin.cpp:8-44): switch c2, add to (b, a). */
int oc_relin(const oc_ctx* c, int level, u64* b, u64* a, u64* c2, const u64* keys, int montgomery) {
    return oc_keyswitch(c, level, c2, keys, b, a, 0, 0, montgomery);
}

/* ===================================================================== */
/* Rescale: ckks_rescale.cpp:11-148                                       */
/* ===================================================================== */

int oc_rescale(const oc_ctx* c, int level, u64* polys, int npolys) {
    if (level < 1 || level >= c->nq) return -1;
    const u64 N = c->N;
    const int L = level + 1;
    const u64 p = c->q[level], half = p >> 1;
    const oc_rescale_t* rc = &c->rc[level];
    for (int k = 0; k < npolys; ++k) {
        u64* last = polys + (u64)k * L * N + (u64)level * N;
        oc_ntt_inverse(last, N, p, c->tq[level].inv, c->tq[level].inv_shoup, c->tq[level].n_inv, 0);
    }
#pragma omp parallel for schedule(static) if (level > 1)
    for (int i = 0; i < level; ++i) {
        const u64 q = c->q[i];
        u64* rem = (u64*)malloc(8 * N);
        const u64* half_ntt = rc->half_ntt + (u64)i * N;
        for (int k = 0; k < npolys; ++k) {
            const u64* last = polys + (u64)k * L * N + (u64)level * N;
            for (u64 j = 0; j < N; ++j) {
                u64 s = last[j] + half;
                if (s >= p) s -= p;
                rem[j] = reduce_u64_shoup(s, q, rc->one_shoup[i]);
            }
            oc_ntt_forward(rem, N, q, c->tq[i].fwd, c->tq[i].fwd_shoup, 0);
            u64* x = polys + (u64)k * L * N + (u64)i * N;
            for (u64 j = 0; j < N; ++j) {
                u64 num = add_q(x[j], half_ntt[j], q);
                u64 diff = sub_q(num, rem[j], q);
                x[j] = shoup_mul_c(diff, rc->inv_p[i], rc->inv_p_shoup[i], q);
            }
        }
        free(rem);
    }
    return 0;
}

/* ===================================================================== */
/* Multiplications and additions: ckks_mul.cpp:37-288; ckks_addition.cpp  */
/* ===================================================================== */

/* ckks_mul_inplace(ct, ct) tensor (ckks_mul.cpp:90-157) */
int oc_tensor(const oc_ctx* c, int level, u64* b1, u64* a1, const u64* b2, const u64* a2, u64* c2) {
    const u64 N = c->N;
    for (int i = 0; i <= level; ++i) {
        const oc_tables* t = &c->tq[i];
        const u64 q = t->q;
        for (u64 x = 0; x < N; ++x) {
            u64 o = (u64)i * N + x;
            u64 cc = barrett_mul_c(a1[o], a2[o], q, t->mu_lo, t->mu_hi);
            u64 tmp = barrett_mul_c(b1[o], a2[o], q, t->mu_lo, t->mu_hi);
            u64 bb = barrett_mul_c(b1[o], b2[o], q, t->mu_lo, t->mu_hi);
            u64 aa = barrett_mul_c(a1[o], b2[o], q, t->mu_lo, t->mu_hi);
            c2[o] = cc;
            b1[o] = bb;
            a1[o] = add_q(aa, tmp, q);
        }
    }
    return 0;
}

/* pointwise_multiply_scalar_inplace_dispatch(..., use_montgomery) (ntt.cpp:540-568) */
static u64 scalar_mul(u64 x, u64 s, const oc_tables* t, int montgomery) {
    if (!montgomery) return barrett_mul_c(x, s, t->q, t->mu_lo, t->mu_hi);
    u64 sm = (u64)(((u128)s * oc_montgomery_radix_mod_q(t->q)) % t->q);
    return mont_mul_c(x, sm, t->q, t->mont_ninv);
}

/* ckks_square_inplace (ckks_mul.cpp:37-88) */
int oc_square(const oc_ctx* c, int level, u64* b, u64* a, u64* c2, int montgomery) {
    const u64 N = c->N;
    for (int i = 0; i <= level; ++i) {
        const oc_tables* t = &c->tq[i];
        const u64 q = t->q;
        u64 sm = (u64)(((u128)2 * oc_montgomery_radix_mod_q(q)) % q);
        for (u64 x = 0; x < N; ++x) {
            u64 o = (u64)i * N + x;
            u64 cc = barrett_mul_c(a[o], a[o], q, t->mu_lo, t->mu_hi);
            u64 bb = barrett_mul_c(b[o], b[o], q, t->mu_lo, t->mu_hi);
            u64 ab = barrett_mul_c(a[o], b[o], q, t->mu_lo, t->mu_hi);
            ab = montgomery ? mont_mul_c(ab, sm, q, t->mont_ninv) : barrett_mul_c(ab, 2, q, t->mu_lo, t->mu_hi);
            c2[o] = cc;
            b[o] = bb;
            a[o] = ab;
        }
    }
    return 0;
}

/* ckks_mul_inplace(ct, pt) (ckks_mul.cpp:160-212) */
int oc_mul_plain(const oc_ctx* c, int level, u64* polys, int npolys, const u64* pt, int pt_mont) {
    const u64 N = c->N;
    const u64 L = (u64)level + 1;
    for (int k = 0; k < npolys; ++k)
        for (int i = 0; i <= level; ++i) {
            const oc_tables* t = &c->tq[i];
            u64* x = polys + (u64)k * L * N + (u64)i * N;
            const u64* y = pt + (u64)i * N;
            for (u64 j = 0; j < N; ++j)
                x[j] = pt_mont ? mont_mul_c(x[j], y[j], t->q, t->mont_ninv)
                               : barrett_mul_c(x[j], y[j], t->q, t->mu_lo, t->mu_hi);
        }
    return 0;
}

/* ckks_mul_inplace(ct, int64) (ckks_mul.cpp:214-272) */
int oc_mul_int(const oc_ctx* c, int level, u64* polys, int npolys, int64_t s, int montgomery) {
    const u64 N = c->N;
    const u64 L = (u64)level + 1;
    if (s == 1) return 0;
    if (s == 0) {
        memset(polys, 0, 8 * N * L * npolys);
        return 0;
    }
    for (int i = 0; i <= level; ++i) {
        const oc_tables* t = &c->tq[i];
        u64 sq = reduce_i64(s, t->q);
        for (int k = 0; k < npolys; ++k) {
            u64* x = polys + (u64)k * L * N + (u64)i * N;
            for (u64 j = 0; j < N; ++j) x[j] = scalar_mul(x[j], sq, t, montgomery);
        }
    }
    return 0;
}

/* ckks_add_inplace / ckks_sub_inplace (ckks_addition.cpp:62-139) */
int oc_add(const oc_ctx* c, int level, u64* x, const u64* y, int npolys) {
    const u64 N = c->N, L = (u64)level + 1;
    for (int k = 0; k < npolys; ++k)
        for (u64 i = 0; i < L; ++i)
            for (u64 j = 0; j < N; ++j) {
                u64 o = (u64)k * L * N + i * N + j;
                x[o] = add_q(x[o], y[o], c->q[i]);
            }
    return 0;
}
int oc_sub(const oc_ctx* c, int level, u64* x, const u64* y, int npolys) {
    const u64 N = c->N, L = (u64)level + 1;
    for (int k = 0; k < npolys; ++k)
        for (u64 i = 0; i < L; ++i)
            for (u64 j = 0; j < N; ++j) {
                u64 o = (u64)k * L * N + i * N + j;
                x[o] = sub_q(x[o], y[o], c->q[i]);
            }
    return 0;
}

/* ckks_add_inplace(ct, pt) (ckks_addition.cpp:35-60): Montgomery plaintexts are
 * converted back with REDC(1 * x) first (ckks_encoding.cpp:67-78). */
int oc_add_plain(const oc_ctx* c, int level, u64* b, const u64* pt, int pt_mont) {
    const u64 N = c->N;
    for (int i = 0; i <= level; ++i) {
        const oc_tables* t = &c->tq[i];
        for (u64 j = 0; j < N; ++j) {
            u64 o = (u64)i * N + j;
            u64 v = pt_mont ? mont_mul_c(1, pt[o], t->q, t->mont_ninv) : pt[o];
            b[o] = add_q(b[o], v, t->q);
        }
    }
    return 0;
}

/* ckks_add_integer_constant_inplace (ckks_addition.cpp:89-112): the per-limb
 * residue is rounded on the host in long double; here it is an input. */
int oc_add_scalar_residues(const oc_ctx* c, int level, u64* b, const u64* res) {
    const u64 N = c->N;
    for (int i = 0; i <= level; ++i)
        for (u64 j = 0; j < N; ++j) {
            u64 o = (u64)i * N + j;
            b[o] = add_q(b[o], res[i], c->q[i]);
        }
    return 0;
}

/* ===================================================================== */
/* Hoisting and QP-domain key switching: ckks_keyswitch.cpp:95-227,       */
/* 344-373, 675-1328                                                      */
/* ===================================================================== */

/* basis index of QP slot l at level L (Q limbs, then P limbs) */
static int qp_basis(const oc_ctx* c, int L, int l) { return l < L ? l : c->nq + (l - L); }

/* ckks_hoist_ciphertext_a_qp_eval (ckks_keyswitch.cpp:344-373): for every
 * digit, ModUp of a's digit limbs to the other Q limbs and P (coefficient
 * domain, then NTT); the digit's own limbs are a's evaluation-domain limbs.
 * digits: [parts][L+K][N], canonical. */
int oc_hoist(const oc_ctx* c, int level, const u64* a_eval, u64* digits) {
    if (level < 0 || level >= c->nq) return -1;
    const u64 N = c->N;
    const int L = level + 1, K = c->np, T = L + K, parts = c->nparts[level], alpha = c->alpha;
    u64* coeff = (u64*)malloc(8 * N * L);
    memcpy(coeff, a_eval, 8 * N * L);
    oc_ntt_inv_q(c, coeff, L, 1);
    for (int part = 0; part < parts; ++part) {
        const int s0 = part * alpha, s1 = s0 + alpha < L ? s0 + alpha : L;
        u64* D = digits + (u64)part * T * N;
        const u64* sp[64];
        u64* dp[256];
        int nd = 0;
        for (int i = 0; i < s1 - s0; ++i) sp[i] = coeff + (u64)(s0 + i) * N;
        for (int i = 0; i < s0; ++i) dp[nd++] = D + (u64)i * N;
        for (int i = s1; i < L; ++i) dp[nd++] = D + (u64)i * N;
        for (int i = 0; i < K; ++i) dp[nd++] = D + (u64)(L + i) * N;
        hps_modup(&c->modup[level][part], sp, dp, N);
        for (int i = s0; i < s1; ++i) memcpy(D + (u64)i * N, a_eval + (u64)i * N, 8 * N);
#pragma omp parallel for schedule(static)
        for (int l = 0; l < T; ++l) {
            if (l >= s0 && l < s1) continue;
            const oc_tables* t = basis_table(c, qp_basis(c, L, l));
            oc_ntt_forward(D + (u64)l * N, N, t->q, t->fwd, t->fwd_shoup, 0);
        }
    }
    free(coeff);
    return 0;
}

/* sum_d digits[d] * key[d] over the L+K QP slots into acc_b, acc_a (added). */
static void qp_key_mac(const oc_ctx* c, int level, const u64* digits, const u64* keys, u64* acc_b,
                       u64* acc_a, int montgomery) {
    const u64 N = c->N;
    const int L = level + 1, K = c->np, T = L + K, KL = c->nq + K, parts = c->nparts[level];
    for (int part = 0; part < parts; ++part) {
        const u64* kb = keys + (u64)part * 2 * KL * N;
        const u64* ka = kb + (u64)KL * N;
        const u64* D = digits + (u64)part * T * N;
#pragma omp parallel for schedule(static)
        for (int l = 0; l < T; ++l) {
            const int basis = qp_basis(c, L, l);
            const oc_tables* t = basis_table(c, basis);
            mac_limb(D + (u64)l * N, ka + (u64)basis * N, acc_a + (u64)l * N, N, t, montgomery);
            mac_limb(D + (u64)l * N, kb + (u64)basis * N, acc_b + (u64)l * N, N, t, montgomery);
        }
    }
}

/* ckks_hybrid_key_switch_hoisted_inplace (ckks_keyswitch.cpp:1072-1328):
 * MAC over the hoisted digits, INTT of the P limbs, folded ModDown pair,
 * combine into (b_out, a_out), then the automorphism (galois != 0). */
int oc_keyswitch_hoisted(const oc_ctx* c, int level, const u64* digits, const u64* keys, u64* b_out,
                         u64* a_out, int b_mode, int a_mode, u64 galois, int montgomery) {
    if (level < 0 || level >= c->nq) return -1;
    const u64 N = c->N;
    const int L = level + 1, K = c->np, T = L + K;
    u64* acc_b = (u64*)calloc(N * T, 8);
    u64* acc_a = (u64*)calloc(N * T, 8);
    qp_key_mac(c, level, digits, keys, acc_b, acc_a, montgomery);
    for (int pi = 0; pi < K; ++pi) {
        const oc_tables* t = &c->tp[pi];
        oc_ntt_inverse(acc_a + (u64)(L + pi) * N, N, t->q, t->inv, t->inv_shoup, t->n_inv, 1);
        oc_ntt_inverse(acc_b + (u64)(L + pi) * N, N, t->q, t->inv, t->inv_shoup, t->n_inv, 1);
    }
    folded_moddown_pair(c, L, acc_a + (u64)L * N, acc_a, acc_b + (u64)L * N, acc_b);
    for (int i = 0; i < L; ++i)
        for (u64 x = 0; x < N; ++x) {
            const u64 o = (u64)i * N + x;
            b_out[o] = b_mode == 0 ? add_q(b_out[o], acc_b[o], c->q[i]) : acc_b[o];
            a_out[o] = a_mode == 0 ? add_q(a_out[o], acc_a[o], c->q[i]) : acc_a[o];
        }
    if (galois) {
        u64* map = (u64*)malloc(8 * N);
        oc_ntt_automorphism_map(N, galois, map);
        oc_apply_automorphism(b_out, N, L, map);
        oc_apply_automorphism(a_out, N, L, map);
        free(map);
    }
    free(acc_b);
    free(acc_a);
    return 0;
}

/* ckks_hybrid_key_switch_hoisted_qp_rotated (ckks_keyswitch.cpp:675-852):
 * b_qp = [b_q ; 0] + sum_d digit*key_b, a_qp = sum_d digit*key_a, both in the
 * QP basis, then the automorphism over all L+K limbs (no ModDown). */
int oc_keyswitch_hoisted_qp_rotated(const oc_ctx* c, int level, const u64* digits, const u64* keys,
                                    const u64* b_q_eval, u64* b_qp, u64* a_qp, u64 galois,
                                    int montgomery) {
    if (level < 0 || level >= c->nq) return -1;
    const u64 N = c->N;
    const int L = level + 1, K = c->np, T = L + K;
    memset(b_qp, 0, 8 * N * T);
    memset(a_qp, 0, 8 * N * T);
    memcpy(b_qp, b_q_eval, 8 * N * L);
    qp_key_mac(c, level, digits, keys, b_qp, a_qp, montgomery);
    u64* map = (u64*)malloc(8 * N);
    oc_ntt_automorphism_map(N, galois, map);
    oc_apply_automorphism(b_qp, N, T, map);
    oc_apply_automorphism(a_qp, N, T, map);
    free(map);
    return 0;
}

/* ckks_hybrid_key_switch_qp_from_key_term (ckks_keyswitch.cpp:854-1070):
 * the key switch up to the inner product, result kept in the QP basis. */
int oc_keyswitch_qp_from_key_term(const oc_ctx* c, int level, const u64* key_term_eval,
                                  const u64* keys, u64* b_qp, u64* a_qp, int montgomery) {
    if (level < 0 || level >= c->nq) return -1;
    const u64 N = c->N;
    const int L = level + 1, K = c->np, T = L + K, parts = c->nparts[level];
    u64* digits = (u64*)malloc(8 * N * T * parts);
    oc_hoist(c, level, key_term_eval, digits);
    memset(b_qp, 0, 8 * N * T);
    memset(a_qp, 0, 8 * N * T);
    qp_key_mac(c, level, digits, keys, b_qp, a_qp, montgomery);
    free(digits);
    return 0;
}

/* ckks_apply_folded_approx_crt_moddown_eval_q (ckks_keyswitch.cpp:609-673)
 * as used by moddown_qp_poly_to_q_eval (ckks_linear_transform.cpp:581-605):
 * qp: L+K eval limbs in, the first L limbs hold the Q result. */
int oc_moddown_qp_to_q(const oc_ctx* c, int level, u64* qp) {
    if (level < 0 || level >= c->nq) return -1;
    const u64 N = c->N;
    const int L = level + 1, K = c->np;
    for (int pi = 0; pi < K; ++pi) {
        const oc_tables* t = &c->tp[pi];
        oc_ntt_inverse(qp + (u64)(L + pi) * N, N, t->q, t->inv, t->inv_shoup, t->n_inv, 1);
    }
    /* the single-poly version is the pair version applied to one poly */
    u64* dummy_src = (u64*)calloc(N * K, 8);
    u64* dummy_dst = (u64*)calloc(N * L, 8);
    folded_moddown_pair(c, L, qp + (u64)L * N, qp, dummy_src, dummy_dst);
    free(dummy_src);
    free(dummy_dst);
    return 0;
}

/* ===================================================================== */
/* BSGS linear transform, QP evaluator: ckks_linear_transform.cpp:1174-1407 */
/* ===================================================================== */

static const u64* find_key(const int64_t* amounts, const u64* const* keys, int nkeys, int64_t r) {
    for (int i = 0; i < nkeys; ++i)
        if (amounts[i] == r) return keys[i];
    return NULL;
}

/* diags: [nterms][L+K][N] QP plaintexts (Montgomery form if pt_mont);
 * diag_idx: diagonal index of each term; keys by rotation amount.
 * Output (a_out, b_out): L limbs each, before the optional rescale. */
int oc_lt_stage_qp(const oc_ctx* c, int level, const u64* a, const u64* b, int nterms,
                   const int64_t* diag_idx, const u64* diags, int pt_mont, int baby_steps,
                   int slots, const int64_t* rot_amounts, const u64* const* rot_keys, int nkeys,
                   int montgomery, u64* a_out, u64* b_out) {
    if (level < 0 || level >= c->nq || nterms < 1 || baby_steps < 1) return -1;
    const u64 N = c->N;
    const int L = level + 1, K = c->np, T = L + K, parts = c->nparts[level];
    const u64 W = (u64)T * N;
    const int giants = (slots + baby_steps - 1) / baby_steps;
    /* compile_linear_transform_stage_metadata (:818-859) */
    char* baby_needed = (char*)calloc(baby_steps, 1);
    for (int t = 0; t < nterms; ++t) baby_needed[diag_idx[t] % baby_steps] = 1;
    u64* baby_a = (u64*)calloc(W * baby_steps, 8);
    u64* baby_b = (u64*)calloc(W * baby_steps, 8);
    if (baby_needed[0]) { /* make_direct_input_qp (:487-532), folded keys */
        memcpy(baby_a, a, 8 * N * L);
        memcpy(baby_b, b, 8 * N * L);
    }
    u64* digits = (u64*)malloc(8 * W * parts);
    oc_hoist(c, level, a, digits);
    for (int r = 1; r < baby_steps; ++r) {
        if (!baby_needed[r]) continue;
        const u64* key = find_key(rot_amounts, rot_keys, nkeys, r);
        if (!key) { free(baby_needed); free(baby_a); free(baby_b); free(digits); return -2; }
        oc_keyswitch_hoisted_qp_rotated(c, level, digits, key, b, baby_b + W * r, baby_a + W * r,
                                        oc_rotation_galois_element(N, r), montgomery);
    }
    u64* out_a = (u64*)calloc(W, 8);
    u64* out_b = (u64*)calloc(W, 8);
    u64* in_a = (u64*)malloc(8 * W);
    u64* in_b = (u64*)malloc(8 * W);
    u64* sw_a = (u64*)malloc(8 * W);
    u64* sw_b = (u64*)malloc(8 * W);
    int have_out = 0;
    for (int g = 0; g < giants; ++g) {
        int have_in = 0;
        for (int t = 0; t < nterms; ++t) {
            if (diag_idx[t] / baby_steps != g) continue;
            const int bb = (int)(diag_idx[t] - (int64_t)g * baby_steps);
            const u64* pt = diags + (u64)t * W;
            for (int l = 0; l < T; ++l) {
                const oc_tables* tb = basis_table(c, qp_basis(c, L, l));
                const u64 q = tb->q;
                for (u64 x = 0; x < N; ++x) {
                    const u64 o = (u64)l * N + x;
                    u64 ta, tb_;
                    if (pt_mont) {
                        ta = mont_mul_c(baby_a[W * bb + o], pt[o], q, tb->mont_ninv);
                        tb_ = mont_mul_c(baby_b[W * bb + o], pt[o], q, tb->mont_ninv);
                    } else {
                        ta = barrett_mul_c(baby_a[W * bb + o], pt[o], q, tb->mu_lo, tb->mu_hi);
                        tb_ = barrett_mul_c(baby_b[W * bb + o], pt[o], q, tb->mu_lo, tb->mu_hi);
                    }
                    in_a[o] = have_in ? add_q(in_a[o], ta, q) : ta;
                    in_b[o] = have_in ? add_q(in_b[o], tb_, q) : tb_;
                }
            }
            have_in = 1;
        }
        if (!have_in) continue;
        if (g == 0) {
            for (u64 o = 0; o < W; ++o) {
                const u64 q = basis_table(c, qp_basis(c, L, (int)(o / N)))->q;
                out_a[o] = have_out ? add_q(out_a[o], in_a[o], q) : in_a[o];
                out_b[o] = have_out ? add_q(out_b[o], in_b[o], q) : in_b[o];
            }
            have_out = 1;
            continue;
        }
        const int64_t gr = (int64_t)g * baby_steps;
        const u64* key = find_key(rot_amounts, rot_keys, nkeys, gr);
        if (!key) { have_out = -1; break; }
        oc_moddown_qp_to_q(c, level, in_a); /* key term = ModDown(inner_a) */
        oc_keyswitch_qp_from_key_term(c, level, in_a, key, sw_b, sw_a, montgomery);
        /* automorphism_qp_pair_accumulate (:373-416) */
        u64* map = (u64*)malloc(8 * N);
        oc_ntt_automorphism_map(N, oc_rotation_galois_element(N, gr), map);
        for (int l = 0; l < T; ++l) {
            const u64 q = basis_table(c, qp_basis(c, L, l))->q;
            for (u64 x = 0; x < N; ++x) {
                const u64 s = (u64)l * N + map[x], o = (u64)l * N + x;
                const u64 add_a = sw_a[s];
                const u64 add_b = add_q(sw_b[s], in_b[s], q);
                out_a[o] = have_out ? add_q(out_a[o], add_a, q) : add_a;
                out_b[o] = have_out ? add_q(out_b[o], add_b, q) : add_b;
            }
        }
        free(map);
        have_out = 1;
    }
    int rc = 0;
    if (have_out != 1) {
        rc = -3;
    } else {
        /* moddown_qp_pair_to_q_eval (:641-718), folded */
        for (int pi = 0; pi < K; ++pi) {
            const oc_tables* t = &c->tp[pi];
            oc_ntt_inverse(out_a + (u64)(L + pi) * N, N, t->q, t->inv, t->inv_shoup, t->n_inv, 1);
            oc_ntt_inverse(out_b + (u64)(L + pi) * N, N, t->q, t->inv, t->inv_shoup, t->n_inv, 1);
        }
        folded_moddown_pair(c, L, out_a + (u64)L * N, out_a, out_b + (u64)L * N, out_b);
        memcpy(a_out, out_a, 8 * N * L);
        memcpy(b_out, out_b, 8 * N * L);
    }
    free(baby_needed); free(baby_a); free(baby_b); free(digits);
    free(out_a); free(out_b); free(in_a); free(in_b); free(sw_a); free(sw_b);
    return rc;
}

/* ===================================================================== */
/* ModRaise (ckks_modraise.cpp:31-75) and the q0p0 switch (1645-1816)     */
/* ===================================================================== */

/* poly_q0_eval: one q0 limb (evaluation domain); out: target+1 limbs. */
int oc_modraise(const oc_ctx* c, int target_level, const u64* poly_q0_eval, u64* out) {
    if (target_level < 1 || target_level >= c->nq) return -1;
    const u64 N = c->N;
    const u64 q0 = c->q[0];
    u64* coeff = (u64*)malloc(8 * N);
    memcpy(coeff, poly_q0_eval, 8 * N);
    oc_ntt_inverse(coeff, N, q0, c->tq[0].inv, c->tq[0].inv_shoup, c->tq[0].n_inv, 0);
    for (u64 x = 0; x < N; ++x) {
        const u64 v = coeff[x];
        const int64_t centred = v > (q0 >> 1) ? -(int64_t)(q0 - v) : (int64_t)v;
        for (int l = 0; l <= target_level; ++l) {
            const u64 q = c->q[l];
            u64 r;
            if (centred >= 0) r = (u64)centred % q;
            else {
                const u64 m = (u64)(-centred) % q;
                r = m == 0 ? 0 : q - m;
            }
            out[(u64)l * N + x] = r;
        }
    }
    oc_ntt_fwd_q(c, out, target_level + 1, 0);
    free(coeff);
    return 0;
}

/* hps_base_change_single_limb (ckks_keyswitch.cpp:1647-1659) */
static u64 single_limb_change(u64 v, u64 src, u64 dst, u64 dst_mu_hi) {
    const u64 r = reduce_4q(v, src);
    const u64 r_dst = barrett_u64_c(r, dst, dst_mu_hi);
    const u64 src_dst = barrett_u64_c(src, dst, dst_mu_hi);
    return r >= (src >> 1) + 1 ? sub_q(r_dst, src_dst, dst) : r_dst;
}

/* ckks_q0p0_key_switch_level0_inplace (ckks_keyswitch.cpp:1661-1816),
 * folded keys.  key: [b(q0, p0), a(q0, p0)] = 4N words. */
int oc_q0p0_keyswitch(const oc_ctx* c, u64* b, u64* a, const u64* key, int montgomery) {
    const u64 N = c->N;
    const oc_tables* tq0 = &c->tq[0];
    const oc_tables* tp0 = &c->tp[0];
    const u64 q0 = tq0->q, p0 = tp0->q;
    const u64* kb = key;
    const u64* ka = key + 2 * N;
    u64* dcoeff = (u64*)malloc(8 * N);
    u64* dp = (u64*)malloc(8 * N);
    u64* aq_a = (u64*)calloc(N, 8);
    u64* aq_b = (u64*)calloc(N, 8);
    u64* ap_a = (u64*)calloc(N, 8);
    u64* ap_b = (u64*)calloc(N, 8);
    memcpy(dcoeff, a, 8 * N);
    oc_ntt_inverse(dcoeff, N, q0, tq0->inv, tq0->inv_shoup, tq0->n_inv, 1);
    for (u64 x = 0; x < N; ++x) dp[x] = single_limb_change(dcoeff[x], q0, p0, tp0->mu_hi);
    oc_ntt_forward(dp, N, p0, tp0->fwd, tp0->fwd_shoup, 1);
    mac_limb(a, ka, aq_a, N, tq0, montgomery);
    mac_limb(a, kb, aq_b, N, tq0, montgomery);
    mac_limb(dp, ka + N, ap_a, N, tp0, montgomery);
    mac_limb(dp, kb + N, ap_b, N, tp0, montgomery);
    oc_ntt_inverse(ap_a, N, p0, tp0->inv, tp0->inv_shoup, tp0->n_inv, 1);
    oc_ntt_inverse(ap_b, N, p0, tp0->inv, tp0->inv_shoup, tp0->n_inv, 1);
    /* folded ModDown over the {q0} / {p0} bases with p0^-1 mod q0 */
    {
        const u64 inv = oc_mod_inverse(p0 % q0, q0);
        const u64 inv_s = oc_compute_shoup(inv, q0);
        const double pinv = 1.0 / (double)p0;
        for (u64 x = 0; x < N; ++x) {
            const u64 va = reduce_4q(ap_a[x], p0), vb = reduce_4q(ap_b[x], p0);
            u64 al_a = (u64)((double)va * pinv + 0.5), al_b = (u64)((double)vb * pinv + 0.5);
            if (al_a > 1) al_a = 1;
            if (al_b > 1) al_b = 1;
            const u64 sa = shoup_mul_c(barrett_u64_c(va, q0, tq0->mu_hi), inv, inv_s, q0);
            const u64 sb = shoup_mul_c(barrett_u64_c(vb, q0, tq0->mu_hi), inv, inv_s, q0);
            ap_a[x] = sub_q(sa, al_a, q0);
            ap_b[x] = sub_q(sb, al_b, q0);
        }
        oc_ntt_forward(ap_a, N, q0, tq0->fwd, tq0->fwd_shoup, 0);
        oc_ntt_forward(ap_b, N, q0, tq0->fwd, tq0->fwd_shoup, 0);
        for (u64 x = 0; x < N; ++x) {
            aq_a[x] = sub_q(aq_a[x], ap_a[x], q0);
            aq_b[x] = sub_q(aq_b[x], ap_b[x], q0);
        }
    }
    for (u64 x = 0; x < N; ++x) {
        b[x] = add_q(b[x], aq_b[x], q0);
        a[x] = aq_a[x];
    }
    free(dcoeff); free(dp); free(aq_a); free(aq_b); free(ap_a); free(ap_b);
    return 0;
}
