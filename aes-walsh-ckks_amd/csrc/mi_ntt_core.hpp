// Register-resident sub-transforms used by every NTT-shaped kernel.
//
// An N = 2^n negacyclic NTT (the reference's Harvey DIT forward / DIF inverse
// with merged psi twiddles, ntt.cpp:235-386) is split into two passes over
// HBM: n1 = n/2 "column" stages and n2 = n - n1 "row" stages.  Inside a pass
// each group of R = 2^LOGR elements (a column or a contiguous row block) is
// held by TG = R/16 threads, 16 words per thread in VGPRs.  Butterflies whose
// partner is in the same thread run in registers; one LDS exchange switches
// between the strided layout (position t + TG*j, coalesced in HBM) and the
// contiguous layout (position t*16 + j) that makes the remaining stages
// thread-local.  Twiddle indices are the reference's global table indices.
#pragma once
#include "mi_arith.hpp"

namespace mi {

constexpr int E = 16;
constexpr int LOGE = 4;
constexpr int NTT_THREADS = 256;
// LDS words per workgroup: NTT_THREADS*E data words, one pad per 16, and one
// more word per group (at most 128 groups) so that the group stride is odd.
constexpr int NTT_LDS_WORDS = NTT_THREADS * E + NTT_THREADS + 128;

__device__ __forceinline__ int lpad(int p) { return p + (p >> LOGE); }

// Modulus constants of one limb for the butterflies: q2 = 2q, and the
// negations negq = 2^64 - q (read from the basis table) and negq2 = 2^64 - 2q,
// so conditional subtractions become selects of a 64-bit add.
struct BfConst {
    u64 q2, negq, negq2;
};
__device__ __forceinline__ BfConst bf_const(u64 q, u64 negq) { return BfConst{q << 1, negq, negq << 1}; }

// Forward (DIT, Harvey lazy): x0,x1 in [0,4q) -> [0,4q).
__device__ __forceinline__ void bf_fwd(u64& x0, u64& x1, u64 w, u64 ws, const BfConst& k) {
    const u64 t = x0 >= k.q2 ? x0 + k.negq2 : x0;
    const u64 u = shoup_lazy_neg(x1, w, ws, k.negq);
    x0 = t + u;
    x1 = t + k.q2 - u;
}

// Inverse (DIF): x0,x1 in [0,2q) -> [0,2q).
__device__ __forceinline__ void bf_inv(u64& x0, u64& x1, u64 w, u64 ws, const BfConst& k) {
    const u64 t = x0, u = x1;
    const u64 s = t + u;
    x1 = shoup_lazy_neg(t - u + k.q2, w, ws, k.negq);
    x0 = s >= k.q2 ? s + k.negq2 : s;
}

template <int LOGR>
struct Grp {
    static constexpr int R = 1 << LOGR;
    static constexpr int TG = R / E;
    static constexpr int LOGTG = LOGR - LOGE;
    // Padded words per group.  The stride is odd: with R + R/E alone (272
    // words at R = 256, a multiple of 16) the same position of the 16 groups
    // that a column-pass wave spans falls in one LDS bank.
    static constexpr int LDS = R + R / E + 1;
    static_assert((NTT_THREADS / TG) * LDS <= NTT_LDS_WORDS, "groups of a workgroup exceed its LDS");
    static_assert(LOGR >= LOGE + 1 && LOGR <= 2 * LOGE, "group size 32..256");
};

// Exchange: strided (t + TG*j) <-> contiguous (t*E + j).  Caller guarantees a
// barrier separates it from any earlier reads of `lds`.
template <int LOGR>
__device__ __forceinline__ void to_contiguous(u64 (&x)[E], int t, u64* lds) {
    using G = Grp<LOGR>;
#pragma unroll
    for (int j = 0; j < E; ++j) lds[lpad(t + G::TG * j)] = x[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) x[j] = lds[lpad(t * E + j)];
}
template <int LOGR>
__device__ __forceinline__ void to_strided(u64 (&x)[E], int t, u64* lds) {
    using G = Grp<LOGR>;
#pragma unroll
    for (int j = 0; j < E; ++j) lds[lpad(t * E + j)] = x[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) x[j] = lds[lpad(t + G::TG * j)];
}

// Forward group transform.  In: strided layout; out: strided layout.
// idx(s, i) gives the global twiddle index of group sub-stage s, block i.
template <int LOGR, class IDX>
__device__ __forceinline__ void group_fwd(u64 (&x)[E], int t, u64* lds, const BfConst& kq,
                                          const u64* __restrict__ W, const u64* __restrict__ WS,
                                          IDX idx) {
    using G = Grp<LOGR>;
#pragma unroll
    for (int s = 0; s < LOGE; ++s) {
        const int dj = E >> (s + 1);
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & dj) continue;
            const int k = idx(s, j0 >> (LOGE - s));
            bf_fwd(x[j0], x[j0 + dj], W[k], WS[k], kq);
        }
    }
    __syncthreads();
    to_contiguous<LOGR>(x, t, lds);
#pragma unroll
    for (int s = LOGE; s < LOGR; ++s) {
        const int d = G::R >> (s + 1);
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & d) continue;
            const int k = idx(s, (t * E + j0) >> (LOGR - s));
            bf_fwd(x[j0], x[j0 + d], W[k], WS[k], kq);
        }
    }
    __syncthreads();
    to_strided<LOGR>(x, t, lds);
}

// Inverse group transform.  In: strided layout; out: strided layout.
// idx(u, i): global twiddle index for group sub-stage u (distance 2^u), block i.
// LAST = false leaves the last sub-stage (pairs j, j + E/2) to the caller.
template <int LOGR, bool LAST = true, class IDX>
__device__ __forceinline__ void group_inv(u64 (&x)[E], int t, u64* lds, const BfConst& kq,
                                          const u64* __restrict__ W, const u64* __restrict__ WS,
                                          IDX idx) {
    using G = Grp<LOGR>;
    __syncthreads();
    to_contiguous<LOGR>(x, t, lds);
#pragma unroll
    for (int u = 0; u < LOGE; ++u) {
        const int d = 1 << u;
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & d) continue;
            const int k = idx(u, (t * E + j0) >> (u + 1));
            bf_inv(x[j0], x[j0 + d], W[k], WS[k], kq);
        }
    }
    __syncthreads();
    to_strided<LOGR>(x, t, lds);
#pragma unroll
    for (int u = LOGE; u < (LAST ? LOGR : LOGR - 1); ++u) {
        const int dj = 1 << (u - G::LOGTG);
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & dj) continue;
            const int k = idx(u, j0 >> (u + 1 - G::LOGTG));
            bf_inv(x[j0], x[j0 + dj], W[k], WS[k], kq);
        }
    }
}

// The same exchanges through 32-bit LDS words, the low halves and then the
// high halves: half the LDS per workgroup (more workgroups per CU) for two
// more barriers per exchange.
template <int LOGR>
__device__ __forceinline__ void to_contiguous_h(u64 (&x)[E], int t, u32* lds) {
    using G = Grp<LOGR>;
    u32 lo[E];
#pragma unroll
    for (int j = 0; j < E; ++j) lds[lpad(t + G::TG * j)] = (u32)x[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) lo[j] = lds[lpad(t * E + j)];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) lds[lpad(t + G::TG * j)] = (u32)(x[j] >> 32);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) x[j] = ((u64)lds[lpad(t * E + j)] << 32) | lo[j];
}
template <int LOGR>
__device__ __forceinline__ void to_strided_h(u64 (&x)[E], int t, u32* lds) {
    using G = Grp<LOGR>;
    u32 lo[E];
#pragma unroll
    for (int j = 0; j < E; ++j) lds[lpad(t * E + j)] = (u32)x[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) lo[j] = lds[lpad(t + G::TG * j)];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) lds[lpad(t * E + j)] = (u32)(x[j] >> 32);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) x[j] = ((u64)lds[lpad(t + G::TG * j)] << 32) | lo[j];
}

// ---- several limbs per workgroup ----
// LPB limbs per workgroup: every thread holds E words of each limb at the same
// positions, the limbs' butterflies are interleaved (independent dependency
// chains for the scheduler) and their LDS exchanges run one after another
// through the one LDS region, so a workgroup keeps LPB limbs in flight for the
// LDS (and occupancy) of one.
template <int LOGR, int LPB, class LW>
__device__ __forceinline__ void to_contiguous_n(u64 (&x)[LPB][E], int t, LW* lds) {
#pragma unroll
    for (int i = 0; i < LPB; ++i) {
        if (i) __syncthreads();  // the previous limb's reads are done
        if constexpr (sizeof(LW) == 4) to_contiguous_h<LOGR>(x[i], t, lds);
        else to_contiguous<LOGR>(x[i], t, lds);
    }
}
template <int LOGR, int LPB, class LW>
__device__ __forceinline__ void to_strided_n(u64 (&x)[LPB][E], int t, LW* lds) {
#pragma unroll
    for (int i = 0; i < LPB; ++i) {
        if (i) __syncthreads();
        if constexpr (sizeof(LW) == 4) to_strided_h<LOGR>(x[i], t, lds);
        else to_strided<LOGR>(x[i], t, lds);
    }
}

// Twiddles of the first sub-stage after an LDS exchange, loaded before the
// exchange's barrier (loads cannot move across it), so their latency
// overlaps the exchange instead of stalling the first butterflies; one limb
// per workgroup only (the registers of two limbs' sets would cost occupancy).
template <int NT>
struct TwPre {
    u64 w[NT], ws[NT];
};

template <int LOGR, int LPB, class IDX, class LW>
__device__ __forceinline__ void group_fwd_n(u64 (&x)[LPB][E], int t, LW* lds, const BfConst (&kq)[LPB],
                                            const u64* const (&W)[LPB], const u64* const (&WS)[LPB], IDX idx) {
    using G = Grp<LOGR>;
#pragma unroll
    for (int s = 0; s < LOGE; ++s) {
        const int dj = E >> (s + 1);
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & dj) continue;
            const int k = idx(s, j0 >> (LOGE - s));
#pragma unroll
            for (int i = 0; i < LPB; ++i) bf_fwd(x[i][j0], x[i][j0 + dj], W[i][k], WS[i][k], kq[i]);
        }
    }
    // first contiguous sub-stage (s = LOGE): twiddle (t*E + j0) >> (LOGR - LOGE)
    constexpr int SH = LOGR - LOGE, NF = E >> SH;
    TwPre<NF> pf;
    if constexpr (LPB == 1) {
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            const int k = idx(LOGE, ((t * E) >> SH) + v);
            pf.w[v] = W[0][k];
            pf.ws[v] = WS[0][k];
        }
    }
    __syncthreads();
    to_contiguous_n<LOGR, LPB>(x, t, lds);
#pragma unroll
    for (int s = LOGE; s < LOGR; ++s) {
        const int d = G::R >> (s + 1);
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & d) continue;
            if (LPB == 1 && s == LOGE) {
                bf_fwd(x[0][j0], x[0][j0 + d], pf.w[j0 >> SH], pf.ws[j0 >> SH], kq[0]);
                continue;
            }
            const int k = idx(s, (t * E + j0) >> (LOGR - s));
#pragma unroll
            for (int i = 0; i < LPB; ++i) bf_fwd(x[i][j0], x[i][j0 + d], W[i][k], WS[i][k], kq[i]);
        }
    }
    __syncthreads();
    to_strided_n<LOGR, LPB>(x, t, lds);
}

template <int LOGR, bool LAST, int LPB, class IDX, class LW>
__device__ __forceinline__ void group_inv_n(u64 (&x)[LPB][E], int t, LW* lds, const BfConst (&kq)[LPB],
                                            const u64* const (&W)[LPB], const u64* const (&WS)[LPB], IDX idx) {
    using G = Grp<LOGR>;
    // first contiguous sub-stage (u = 0): twiddle (t*E + j0) >> 1
    constexpr int NC = E / 2;
    TwPre<NC> pc;
    if constexpr (LPB == 1) {
#pragma unroll
        for (int v = 0; v < NC; ++v) {
            const int k = idx(0, t * NC + v);
            pc.w[v] = W[0][k];
            pc.ws[v] = WS[0][k];
        }
    }
    __syncthreads();
    to_contiguous_n<LOGR, LPB>(x, t, lds);
#pragma unroll
    for (int u = 0; u < LOGE; ++u) {
        const int d = 1 << u;
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & d) continue;
            if (LPB == 1 && u == 0) {
                bf_inv(x[0][j0], x[0][j0 + d], pc.w[j0 >> 1], pc.ws[j0 >> 1], kq[0]);
                continue;
            }
            const int k = idx(u, (t * E + j0) >> (u + 1));
#pragma unroll
            for (int i = 0; i < LPB; ++i) bf_inv(x[i][j0], x[i][j0 + d], W[i][k], WS[i][k], kq[i]);
        }
    }
    // first strided sub-stage (u = LOGE): twiddle j0 >> (LOGE + 1 - LOGTG)
    constexpr bool HAS_S = (LAST ? LOGR : LOGR - 1) > LOGE;
    constexpr int SS = LOGE + 1 - G::LOGTG, NS = E >> SS;
    TwPre<NS> ps;
    if constexpr (LPB == 1 && HAS_S) {
#pragma unroll
        for (int v = 0; v < NS; ++v) {
            const int k = idx(LOGE, v);
            ps.w[v] = W[0][k];
            ps.ws[v] = WS[0][k];
        }
    }
    __syncthreads();
    to_strided_n<LOGR, LPB>(x, t, lds);
#pragma unroll
    for (int u = LOGE; u < (LAST ? LOGR : LOGR - 1); ++u) {
        const int dj = 1 << (u - G::LOGTG);
#pragma unroll
        for (int j0 = 0; j0 < E; ++j0) {
            if (j0 & dj) continue;
            if (LPB == 1 && u == LOGE) {
                bf_inv(x[0][j0], x[0][j0 + dj], ps.w[j0 >> SS], ps.ws[j0 >> SS], kq[0]);
                continue;
            }
            const int k = idx(u, j0 >> (u + 1 - G::LOGTG));
#pragma unroll
            for (int i = 0; i < LPB; ++i) bf_inv(x[i][j0], x[i][j0 + dj], W[i][k], WS[i][k], kq[i]);
        }
    }
}

// Geometry of one pass.  Column passes: group = column (stride R2 words
// between elements), thread (g = tid % G, t = tid / G).  Row passes: group =
// contiguous block of R2 words, thread (g = tid / TG, t = tid % TG).
struct PassGeom {
    int n, n1, n2;
    int threads;     // workgroup size
    int groups;      // groups per workgroup
    int wg_per_limb; // workgroups per limb
};

}  // namespace mi
