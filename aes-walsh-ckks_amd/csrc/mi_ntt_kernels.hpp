// NTT pass kernels templated on a Job (see mi_ntt.hip).
//
// A Job is a trivially copyable struct with
//   __device__ View view(int k) const;   // limb k of the batch
//   __device__ bool skip(int k) const;   // whole limb skipped (uniform)
// and a View with
//   int basis;                            // basis index of the limb
//   __device__ u64 load(int i) const;     // coefficient / slot i
//   __device__ void store(int i, u64 v) const;
// Forward column pass: load = job input, store = intermediate ([0,4q)).
// Forward row pass: load = intermediate, store = canonical output.
// Inverse row pass: load = input (canonical), store = intermediate ([0,2q)).
// Inverse column pass: load = intermediate, store = canonical output.
#pragma once
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
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const {
        const int npolys = b.count / b.lpp;
        const int l = k / npolys, poly = k - l * npolys;
        return View{b.basis0 + l, b.base + (long long)poly * b.stride + (long long)l * N};
    }
    __device__ bool skip(int) const { return false; }
};

template <int LOGR, bool FWD, class Job>
__global__ void __launch_bounds__(NTT_THREADS)
k_ntt_cols(NttTabs tb, Job job, int n1, int n2, int G, int wpl) {
    __shared__ u64 lds[NTT_LDS_WORDS];
    using Gr = Grp<LOGR>;
    const int limb = blockIdx.x / wpl, w = blockIdx.x - limb * wpl;
    if (job.skip(limb)) return;
    const int g = threadIdx.x % G, t = threadIdx.x / G;
    const auto v = job.view(limb);
    const BasisConst bc = tb.bc[v.basis];
    const int col = w * G + g;
    u64 x[E];
#pragma unroll
    for (int j = 0; j < E; ++j) x[j] = v.load(((t + Gr::TG * j) << n2) + col);
    u64* my = lds + g * Gr::LDS;
    if constexpr (FWD) {
        const u64* W = tb.fwd + (long long)v.basis * tb.N;
        const u64* WS = tb.fwd_shoup + (long long)v.basis * tb.N;
        group_fwd<LOGR>(x, t, my, bf_const(bc.q, bc.negq), W, WS,
                        [](int s, int i) { return (1 << s) + i; });
    } else {
        const u64* W = tb.inv + (long long)v.basis * tb.N;
        const u64* WS = tb.inv_shoup + (long long)v.basis * tb.N;
        group_inv<LOGR>(x, t, my, bf_const(bc.q, bc.negq), W, WS,
                        [n1](int u, int i) { return (1 << (n1 - 1 - u)) + i; });
#pragma unroll
        for (int j = 0; j < E; ++j) x[j] = shoup(x[j], bc.n_inv, bc.n_inv_shoup, bc.q);
    }
#pragma unroll
    for (int j = 0; j < E; ++j) v.store(((t + Gr::TG * j) << n2) + col, x[j]);
}

template <int LOGR, bool FWD, class Job>
__global__ void __launch_bounds__(NTT_THREADS)
k_ntt_rows(NttTabs tb, Job job, int n1, int n2, int G, int wpl) {
    __shared__ u64 lds[NTT_LDS_WORDS];
    using Gr = Grp<LOGR>;
    const int limb = blockIdx.x / wpl, w = blockIdx.x - limb * wpl;
    if (job.skip(limb)) return;
    const int g = threadIdx.x / Gr::TG, t = threadIdx.x % Gr::TG;
    const auto v = job.view(limb);
    const BasisConst bc = tb.bc[v.basis];
    const int row = w * G + g;
    const int off = row << n2;
    u64 x[E];
#pragma unroll
    for (int j = 0; j < E; ++j) x[j] = v.load(off + t + Gr::TG * j);
    u64* my = lds + g * Gr::LDS;
    const int n = n1 + n2;
    if constexpr (FWD) {
        const u64* W = tb.fwd + (long long)v.basis * tb.N;
        const u64* WS = tb.fwd_shoup + (long long)v.basis * tb.N;
        group_fwd<LOGR>(x, t, my, bf_const(bc.q, bc.negq), W, WS,
                        [n1, row](int s, int i) { return (1 << (n1 + s)) + (row << s) + i; });
#pragma unroll
        for (int j = 0; j < E; ++j) x[j] = reduce_4q(x[j], bc.q);
    } else {
        const u64* W = tb.inv + (long long)v.basis * tb.N;
        const u64* WS = tb.inv_shoup + (long long)v.basis * tb.N;
        group_inv<LOGR>(x, t, my, bf_const(bc.q, bc.negq), W, WS, [n, n2, row](int u, int i) {
            return (1 << (n - 1 - u)) + (row << (n2 - 1 - u)) + i;
        });
    }
#pragma unroll
    for (int j = 0; j < E; ++j) v.store(off + t + Gr::TG * j, x[j]);
}

// Workgroup shape of a pass whose groups have 2^logr elements and which has
// 2^logother groups per limb.
inline void pass_shape(int logr, int logother, int& threads, int& G, int& wpl) {
    int tg = 1 << (logr - LOGE);
    threads = tg << logother;
    if (threads > NTT_THREADS) threads = NTT_THREADS;
    G = threads / tg;
    wpl = (1 << logother) / G;
}

template <bool FWD, class Job>
void launch_cols(const Context& c, const Job& job, int count, hipStream_t s) {
    int threads, G, wpl;
    pass_shape(c.n1, c.n2, threads, G, wpl);
    dim3 grid((unsigned)count * wpl), block(threads);
    NttTabs tb = tabs_of(c);
    switch (c.n1) {
        case 5: hipLaunchKernelGGL((k_ntt_cols<5, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        case 6: hipLaunchKernelGGL((k_ntt_cols<6, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        case 7: hipLaunchKernelGGL((k_ntt_cols<7, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        case 8: hipLaunchKernelGGL((k_ntt_cols<8, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        default: throw Error(-1, "ntt: unsupported ring dimension");
    }
    MI_HIP(hipGetLastError());
}

template <bool FWD, class Job>
void launch_rows(const Context& c, const Job& job, int count, hipStream_t s) {
    int threads, G, wpl;
    pass_shape(c.n2, c.n1, threads, G, wpl);
    dim3 grid((unsigned)count * wpl), block(threads);
    NttTabs tb = tabs_of(c);
    switch (c.n2) {
        case 5: hipLaunchKernelGGL((k_ntt_rows<5, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        case 6: hipLaunchKernelGGL((k_ntt_rows<6, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        case 7: hipLaunchKernelGGL((k_ntt_rows<7, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        case 8: hipLaunchKernelGGL((k_ntt_rows<8, FWD, Job>), grid, block, 0, s, tb, job, c.n1, c.n2, G, wpl); break;
        default: throw Error(-1, "ntt: unsupported ring dimension");
    }
    MI_HIP(hipGetLastError());
}

// Forward NTT: first pass runs job `in` (load input, store intermediate),
// second pass runs job `out` (load intermediate, store output).
template <class JobIn, class JobOut>
void run_ntt_fwd(const Context& c, const JobIn& in, const JobOut& out, int count, hipStream_t s) {
    launch_cols<true>(c, in, count, s);
    launch_rows<true>(c, out, count, s);
}
template <class JobIn, class JobOut>
void run_ntt_inv(const Context& c, const JobIn& in, const JobOut& out, int count, hipStream_t s) {
    launch_rows<false>(c, in, count, s);
    launch_cols<false>(c, out, count, s);
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
