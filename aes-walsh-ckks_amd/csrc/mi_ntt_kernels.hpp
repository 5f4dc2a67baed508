// NTT pass kernels templated on a Job (see mi_ntt.hip).
//
// A Job is a trivially copyable struct with
//   __device__ View view(int k) const;   // limb k of the batch
//   __device__ bool skip(int k) const;   // whole limb skipped (uniform)
// and a View with
//   int basis;                            // basis index of the limb
//   __device__ u64 load(int i) const;     // coefficient / slot i
//   __device__ void store(int i, u64 v) const;
// A View may also offer `u64 pre(int i) const` and
// `void store_pre(int i, u64 v, u64 pre) const`: the pass then reads pre(i)
// for its output positions together with its input, before the butterflies,
// and stores through store_pre, so an output that combines the transform
// with another operand does not wait for that operand's loads at the end.
// Forward column pass: load = job input, store = intermediate ([0,4q)).
// Forward row pass: load = intermediate, store = canonical output.
// Inverse row pass: load = input (canonical), store = intermediate ([0,2q)).
// Inverse column pass: load = intermediate, store = canonical output.
#pragma once
#include <type_traits>
#include "mi_internal.hpp"
#include "mi_ntt_core.hpp"

namespace mi {

struct NttTabs {
    const u64* fwd;
    const u64* fwd_shoup;
    const u64* inv;
    const u64* inv_shoup;
    const BasisConst* bc;
    u64 N;
};

inline NttTabs tabs_of(const Context& c) {
    return NttTabs{c.d_fwd, c.d_fwd_shoup, c.d_inv, c.d_inv_shoup, c.d_bc, c.N};
}

// In-place NTT over a strided batch.  Workgroups walk the batch modulus-major
// (all polys' limb 0, then limb 1, ...): the workgroups resident on one XCD
// at a time then share a single modulus's twiddle tables in that XCD's L2.
struct PlainJob {
    LimbBatch b;
    u64 N;
    int npolys;        // polys per modulus: b.count / b.lpp
    FastDiv per_limb;  // npolys
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const {
        const int l = per_limb.div(k), poly = k - l * npolys;
        return View{b.basis0 + l, b.base + (long long)poly * b.stride + (long long)l * N};
    }
    __device__ bool skip(int) const { return false; }
};
inline PlainJob plain_job(const LimbBatch& b, u64 N) {
    return PlainJob{b, N, b.count / b.lpp, fast_div(b.count / b.lpp)};
}

// Pass geometry, fixed at compile time by log2 N: n1 = n/2 column stages and
// n2 = n - n1 row stages.  A pass whose groups have 2^logr elements and which
// has 2^logother groups per limb runs workgroups of pass_threads() threads
// holding pass_groups() groups, pass_wpl() workgroups per limb.
constexpr int pass_threads(int logr, int logother) {
    return ((1 << (logr - LOGE)) << logother) > NTT_THREADS ? NTT_THREADS : ((1 << (logr - LOGE)) << logother);
}
constexpr int pass_groups(int logr, int logother) { return pass_threads(logr, logother) >> (logr - LOGE); }
constexpr int pass_wpl(int logr, int logother) { return (1 << logother) / pass_groups(logr, logother); }

template <class V, class = void>
struct HasPre : std::false_type {};
template <class V>
struct HasPre<V, std::void_t<decltype(&V::pre)>> : std::true_type {};

// LPB limbs per workgroup (1 or 2): workgroup b serves limbs LPB*(b / wpl) +
// 0..LPB-1 of a launch over `count` limbs; a limb past `count` or skipped by
// the job is left alone (uniform over the workgroup).
template <int LPB, class Job, class V>
__device__ __forceinline__ bool limb_views(const Job& job, int first, int count, V (&v)[LPB], bool (&live)[LPB]) {
    bool any = false;
    int k0 = -1;
#pragma unroll
    for (int i = 0; i < LPB; ++i) {
        const int k = first + i;
        live[i] = k < count && !job.skip(k);
        if (live[i] && k0 < 0) k0 = k;
        any = any || live[i];
    }
    if (!any) return false;
#pragma unroll
    for (int i = 0; i < LPB; ++i) v[i] = job.view(live[i] ? first + i : k0);  // a dead limb borrows a live view
    return true;
}

template <int LOGN, bool FWD, int LPB, bool HALF, class Job>
__global__ void __launch_bounds__(NTT_THREADS)
k_ntt_cols(NttTabs tb, Job job, int count) {
    constexpr int n1 = LOGN / 2, n2 = LOGN - n1, LOGR = n1;
    constexpr int G = pass_groups(n1, n2), wpl = pass_wpl(n1, n2);
    using LW = std::conditional_t<HALF, u32, u64>;  // LDS exchange word
    __shared__ LW lds[NTT_LDS_WORDS];
    using Gr = Grp<LOGR>;
    using V = decltype(job.view(0));
    const int blk = blockIdx.x / wpl, w = blockIdx.x - blk * wpl;
    V v[LPB];
    bool live[LPB];
    if (!limb_views<LPB>(job, blk * LPB, count, v, live)) return;
    const int g = threadIdx.x % G, t = threadIdx.x / G;
    const int col = w * G + g;
    constexpr bool PRE = HasPre<V>::value;
    u64 x[LPB][E], pre[LPB][PRE ? E : 1];
    BasisConst bc[LPB];
    BfConst kq[LPB];
    const u64* W[LPB];
    const u64* WS[LPB];
#pragma unroll
    for (int i = 0; i < LPB; ++i) {
        if (live[i]) {
#pragma unroll
            for (int j = 0; j < E; ++j) x[i][j] = v[i].load(((t + Gr::TG * j) << n2) + col);
            if constexpr (PRE) {
#pragma unroll
                for (int j = 0; j < E; ++j) pre[i][j] = v[i].pre(((t + Gr::TG * j) << n2) + col);
            }
        }
        bc[i] = tb.bc[v[i].basis];
        kq[i] = bf_const(bc[i].q, bc[i].negq);
        W[i] = (FWD ? tb.fwd : tb.inv) + (long long)v[i].basis * tb.N;
        WS[i] = (FWD ? tb.fwd_shoup : tb.inv_shoup) + (long long)v[i].basis * tb.N;
    }
    LW* my = lds + g * Gr::LDS;
    if constexpr (FWD) {
        group_fwd_n<LOGR, LPB>(x, t, my, kq, W, WS, [](int s, int i) { return (1 << s) + i; });
    } else {
        group_inv_n<LOGR, false, LPB>(x, t, my, kq, W, WS, [](int u, int i) { return (1 << (n1 - 1 - u)) + i; });
        // last stage (distance N/2, twiddle index 1) with the N^-1 scaling:
        // x0 = (t + u) N^-1, x1 = (t - u) w_1 N^-1, canonical
#pragma unroll
        for (int i = 0; i < LPB; ++i) {
            const u64 q2 = bc[i].q << 1;
#pragma unroll
            for (int j = 0; j < E / 2; ++j) {
                const u64 a = x[i][j], b = x[i][j + E / 2];
                x[i][j] = shoup(a + b, bc[i].n_inv, bc[i].n_inv_shoup, bc[i].q);
                x[i][j + E / 2] = shoup(a - b + q2, bc[i].w1_ninv, bc[i].w1_ninv_shoup, bc[i].q);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < LPB; ++i) {
        if (!live[i]) continue;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            if constexpr (PRE) v[i].store_pre(((t + Gr::TG * j) << n2) + col, x[i][j], pre[i][j]);
            else v[i].store(((t + Gr::TG * j) << n2) + col, x[i][j]);
        }
    }
}

template <int LOGN, bool FWD, int LPB, bool HALF, class Job>
__device__ __forceinline__ void ntt_rows_body(const NttTabs& tb, const Job& job, int count) {
    constexpr int n1 = LOGN / 2, n2 = LOGN - n1, LOGR = n2;
    constexpr int G = pass_groups(n2, n1), wpl = pass_wpl(n2, n1);
    using LW = std::conditional_t<HALF, u32, u64>;  // LDS exchange word
    __shared__ LW lds[NTT_LDS_WORDS];
    using Gr = Grp<LOGR>;
    using V = decltype(job.view(0));
    const int blk = blockIdx.x / wpl, w = blockIdx.x - blk * wpl;
    V v[LPB];
    bool live[LPB];
    if (!limb_views<LPB>(job, blk * LPB, count, v, live)) return;
    const int g = threadIdx.x / Gr::TG, t = threadIdx.x % Gr::TG;
    const int row = w * G + g;
    const int off = row << n2;
    constexpr bool PRE = HasPre<V>::value;
    u64 x[LPB][E], pre[LPB][PRE ? E : 1];
    BasisConst bc[LPB];
    BfConst kq[LPB];
    const u64* W[LPB];
    const u64* WS[LPB];
#pragma unroll
    for (int i = 0; i < LPB; ++i) {
        if (live[i]) {
#pragma unroll
            for (int j = 0; j < E; ++j) x[i][j] = v[i].load(off + t + Gr::TG * j);
            if constexpr (PRE) {
#pragma unroll
                for (int j = 0; j < E; ++j) pre[i][j] = v[i].pre(off + t + Gr::TG * j);
            }
        }
        bc[i] = tb.bc[v[i].basis];
        kq[i] = bf_const(bc[i].q, bc[i].negq);
        W[i] = (FWD ? tb.fwd : tb.inv) + (long long)v[i].basis * tb.N;
        WS[i] = (FWD ? tb.fwd_shoup : tb.inv_shoup) + (long long)v[i].basis * tb.N;
    }
    LW* my = lds + g * Gr::LDS;
    if constexpr (FWD) {
        group_fwd_n<LOGR, LPB>(x, t, my, kq, W, WS,
                               [row](int s, int i) { return (1 << (n1 + s)) + (row << s) + i; });
#pragma unroll
        for (int i = 0; i < LPB; ++i)
#pragma unroll
            for (int j = 0; j < E; ++j) x[i][j] = reduce_4q(x[i][j], bc[i].q);
    } else {
        group_inv_n<LOGR, true, LPB>(x, t, my, kq, W, WS, [row](int u, int i) {
            return (1 << (LOGN - 1 - u)) + (row << (n2 - 1 - u)) + i;
        });
    }
#pragma unroll
    for (int i = 0; i < LPB; ++i) {
        if (!live[i]) continue;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            if constexpr (PRE) v[i].store_pre(off + t + Gr::TG * j, x[i][j], pre[i][j]);
            else v[i].store(off + t + Gr::TG * j, x[i][j]);
        }
    }
}

template <int LOGN, bool FWD, int LPB, bool HALF, class Job>
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_rows(NttTabs tb, Job job, int count) {
    ntt_rows_body<LOGN, FWD, LPB, HALF>(tb, job, count);
}
// The same pass held to 128 VGPRs (four waves per SIMD, the LDS limit), for
// jobs whose load or store does enough work of its own that the scheduler
// would otherwise trade occupancy for loads hoisted across the whole pass.
template <int LOGN, bool FWD, class Job>
__global__ void __launch_bounds__(NTT_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
k_ntt_rows_w4(NttTabs tb, Job job, int count) {
    ntt_rows_body<LOGN, FWD, 1, false>(tb, job, count);
}

#define MI_NTT_LOGN_CASES(K)                                                                          \
    switch (c.logN) {                                                                                 \
        case 10: K(10); break;                                                                        \
        case 11: K(11); break;                                                                        \
        case 12: K(12); break;                                                                        \
        case 13: K(13); break;                                                                        \
        case 14: K(14); break;                                                                        \
        case 15: K(15); break;                                                                        \
        case 16: K(16); break;                                                                        \
        default: throw Error(-1, "ntt: unsupported ring dimension");                                  \
    }

template <bool FWD, int LPB, bool HALF, class Job>
void launch_cols(const Context& c, const Job& job, int count, hipStream_t s) {
    if (count >= (1 << 20)) throw Error(-1, "ntt: more than 2^20 limbs in one launch");
    const NttTabs tb = tabs_of(c);
    const unsigned blocks = (unsigned)((count + LPB - 1) / LPB);
#define MI_COLS(LN)                                                                                   \
    hipLaunchKernelGGL((k_ntt_cols<LN, FWD, LPB, HALF, Job>), dim3(blocks * pass_wpl(LN / 2, LN - LN / 2)), \
                       dim3(pass_threads(LN / 2, LN - LN / 2)), 0, s, tb, job, count)
    MI_NTT_LOGN_CASES(MI_COLS)
#undef MI_COLS
    MI_HIP(hipGetLastError());
}

template <bool FWD, int LPB, bool HALF, class Job>
void launch_rows(const Context& c, const Job& job, int count, hipStream_t s) {
    if (count >= (1 << 20)) throw Error(-1, "ntt: more than 2^20 limbs in one launch");
    const NttTabs tb = tabs_of(c);
    const unsigned blocks = (unsigned)((count + LPB - 1) / LPB);
#define MI_ROWS(LN)                                                                                   \
    hipLaunchKernelGGL((k_ntt_rows<LN, FWD, LPB, HALF, Job>), dim3(blocks * pass_wpl(LN - LN / 2, LN / 2)), \
                       dim3(pass_threads(LN - LN / 2, LN / 2)), 0, s, tb, job, count)
    MI_NTT_LOGN_CASES(MI_ROWS)
#undef MI_ROWS
    MI_HIP(hipGetLastError());
}

template <bool FWD, class Job>
void launch_rows_w4(const Context& c, const Job& job, int count, hipStream_t s) {
    if (count >= (1 << 20)) throw Error(-1, "ntt: more than 2^20 limbs in one launch");
    const NttTabs tb = tabs_of(c);
#define MI_ROWS(LN)                                                                                   \
    hipLaunchKernelGGL((k_ntt_rows_w4<LN, FWD, Job>), dim3((unsigned)count * pass_wpl(LN - LN / 2, LN / 2)), \
                       dim3(pass_threads(LN - LN / 2, LN / 2)), 0, s, tb, job, count)
    MI_NTT_LOGN_CASES(MI_ROWS)
#undef MI_ROWS
    MI_HIP(hipGetLastError());
}

// One pass in the configuration `variant` of the context's NTT plan:
// 1 = one limb per workgroup, 2 = two limbs per workgroup, 3 = one limb
// with the LDS exchanges in 32-bit halves.
template <bool FWD, bool ROWS, class Job>
void launch_pass(const Context& c, const Job& job, int count, hipStream_t s) {
    switch (c.ntt_variant(count)) {
        case 2:
            if constexpr (ROWS) launch_rows<FWD, 2, false>(c, job, count, s);
            else launch_cols<FWD, 2, false>(c, job, count, s);
            break;
        case 3:
            if constexpr (ROWS) launch_rows<FWD, 1, true>(c, job, count, s);
            else launch_cols<FWD, 1, true>(c, job, count, s);
            break;
        default:
            if constexpr (ROWS) launch_rows<FWD, 1, false>(c, job, count, s);
            else launch_cols<FWD, 1, false>(c, job, count, s);
    }
}

// Forward NTT: first pass runs job `in` (load input, store intermediate),
// second pass runs job `out` (load intermediate, store output), each in the
// configuration the context's tuned plan gives for the launch size.
template <class JobIn, class JobOut>
void run_ntt_fwd(const Context& c, const JobIn& in, const JobOut& out, int count, hipStream_t s) {
    const long probe = c.probe_open(s, count);
    launch_pass<true, false>(c, in, count, s);
    launch_pass<true, true>(c, out, count, s);
    c.probe_close(probe, s);
}
template <class JobIn, class JobOut>
void run_ntt_inv(const Context& c, const JobIn& in, const JobOut& out, int count, hipStream_t s) {
    const long probe = c.probe_open(s, count);
    launch_pass<false, true>(c, in, count, s);
    launch_pass<false, false>(c, out, count, s);
    c.probe_close(probe, s);
}

// Block -> (item, slot, tile) with the blocks of one (slot, tile) for all
// items consecutive on one XCD: workgroups are dealt to the 8 XCDs round
// robin, so block b runs on XCD b % 8.  The key (or diagonal) tile that all
// items read is then fetched once into that XCD's L2.
struct TileMap {
    int item, slot, tile;
    bool valid;
};
__device__ __forceinline__ TileMap map_tile(int b, int items, int tiles_per_slot, int slots) {
    const int xcd = b & 7, local = b >> 3;
    const int it = local % items, tl = local / items;
    const int tile = tl * 8 + xcd;
    TileMap m;
    m.item = it;
    m.slot = tile / tiles_per_slot;
    m.tile = tile - m.slot * tiles_per_slot;
    m.valid = m.slot < slots;
    return m;
}
inline unsigned tile_grid(int items, int tiles_per_slot, int slots) {
    const long long tiles = (long long)tiles_per_slot * slots;
    return (unsigned)(((tiles + 7) / 8) * 8 * items);
}

}  // namespace mi
