// Hybrid key switching (reference ckks_hybrid_key_switch_inplace,
// ckks_keyswitch.cpp:1330-1641, fold_hybrid_keyswitch = true), hoisting
// (ckks_hoist_ciphertext_a_qp_eval, :344-373) and the QP-basis variants
// (:675-1328).
//
// Steps, each a device routine reused by the linear transform:
//   ks_modup     C = INTT(key term); D[d] = HPS ModUp of digit d's limbs of C
//                to the other Q limbs and P, the digit's own limbs copied from
//                the evaluation-domain key term (:278-332, :1452); NTT of D.
//   k_mac        acc = sum_d D[d] * key[d] per QP slot (Montgomery REDC or
//                Barrett), written either split (Q slots combined into the
//                outputs with Add/Assign, P slots kept for ModDown) or whole in
//                the QP basis, optionally through a Galois automorphism and
//                with a b term added (:675-852, :373-416).
//   ks_moddown   folded ModDown (:521-673): INTT of the P slots, contribution
//                sum_j x_j p_j^-1 - alpha per Q limb (the P-side HPS estimate
//                computed once per coefficient), NTT, subtracted from Q.
// Reordering "+ acc" before "- NTT(contribution)" leaves every canonical
// residue unchanged.
#include "mi_ntt_kernels.hpp"

#include <algorithm>
#include <vector>

namespace mi {

constexpr int KS_THREADS = 256;
constexpr int MAX_ALPHA = 16;   // digit width held in registers
constexpr int MODUP_DCH = 8;    // destination limbs per ModUp thread
constexpr int MODDOWN_TCH = 4;  // target limbs per ModDown-contribution thread

// ---- out-of-place inverse NTT of n key terms -> C (item-minor within a limb) ----
struct CopyListJob {
    const u64* src[KS_MAX_ITEMS];
    u64* dst;
    long long dstride;
    int n;
    u64 N;
    FastDiv per_limb;  // n
    struct View {
        int basis;
        const u64* s;
        u64* d;
        __device__ u64 load(int i) const { return s[i]; }
        __device__ void store(int i, u64 v) const { d[i] = v; }
    };
    __device__ View view(int k) const {
        const int l = per_limb.div(k), it = k - l * n;
        return View{l, src[it] + (long long)l * N, dst + it * dstride + (long long)l * N};
    }
    __device__ bool skip(int) const { return false; }
};

// ---- in-place NTT over a list of polys (limb-major: shared twiddles) ----
struct ListJob {
    u64* p[2 * KS_MAX_ITEMS];
    int npoly, basis0;
    u64 N;
    FastDiv per_limb;  // npoly
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const {
        const int l = per_limb.div(k), poly = k - l * npoly;
        return View{basis0 + l, p[poly] + (long long)l * N};
    }
    __device__ bool skip(int) const { return false; }
};

// ---- ModUp of all digits of n items ----
struct ModUpArgs {
    const u64* coeff;                   // C: item i's [L][N] at coeff + i*L*N
    const u64* eval[KS_MAX_ITEMS];      // key terms [L][N]
    u64* digits;                        // item i's D [parts][L+K][N] at digits + i*dstride
    long long dstride;
    int L, K, parts, chunks;
    int logN;
    ModUpTables t[16];
    const BasisConst* bc;
};

// One block: 256 coefficients of one digit of one item, up to MODUP_DCH
// destination limbs.  The digit's tables are staged in LDS first: every
// thread reads the same entries, so they are LDS broadcasts instead of a
// dependent global load per product.  NS = the digit's limb count.
template <int NS>
__global__ void __launch_bounds__(KS_THREADS) k_modup(ModUpArgs a) {
    __shared__ u64 s_row[MODUP_DCH][NS];       // qhat_i mod m_j
    __shared__ u64 s_corr[MODUP_DCH][NS + 1];  // alpha * Q_digit mod m_j
    __shared__ u64 s_dq[MODUP_DCH], s_ninv[MODUP_DCH], s_muhi[MODUP_DCH];
    __shared__ int s_slot[MODUP_DCH];
    __shared__ u64 s_sq[NS], s_qinv[NS], s_qinvs[NS];
    __shared__ double s_sinv[NS];
    const int d = blockIdx.y;
    const int it = blockIdx.z / a.chunks;
    const int chunk = blockIdx.z - it * a.chunks;
    const u64 N = (u64)1 << a.logN;
    const ModUpTables& t = a.t[d];
    const int j0 = chunk * MODUP_DCH;
    if (j0 >= t.nd) return;  // uniform over the block
    const int nj = j0 + MODUP_DCH < t.nd ? MODUP_DCH : t.nd - j0;
    for (int e = threadIdx.x; e < nj * NS; e += KS_THREADS)
        s_row[e / NS][e % NS] = t.qhat_mod_dst[(long long)(j0 + e / NS) * NS + e % NS];
    for (int e = threadIdx.x; e < nj * (NS + 1); e += KS_THREADS)
        s_corr[e / (NS + 1)][e % (NS + 1)] = t.alpha_corr[(long long)(j0 + e / (NS + 1)) * (NS + 1) + e % (NS + 1)];
    if (threadIdx.x < nj) {
        const int j = j0 + threadIdx.x;
        const BasisConst b = a.bc[t.dst_basis[j]];
        s_dq[threadIdx.x] = b.q;
        s_ninv[threadIdx.x] = b.qninv;
        s_muhi[threadIdx.x] = b.mu_hi;
        s_slot[threadIdx.x] =
            j < t.s0 ? j : (j < t.s0 + (a.L - t.s1) ? t.s1 + (j - t.s0) : a.L + (j - (a.L - NS)));
    }
    if (threadIdx.x >= 64 && threadIdx.x < 64 + NS) {
        const int i = threadIdx.x - 64;
        s_sq[i] = a.bc[t.s0 + i].q;
        s_qinv[i] = t.qhat_inv[i];
        s_qinvs[i] = t.qhat_inv_shoup[i];
        s_sinv[i] = t.src_inv[i];
    }
    __syncthreads();
    // N is a multiple of KS_THREADS: every thread has a coefficient
    const int c = blockIdx.x * KS_THREADS + threadIdx.x;
    const int slots = a.L + a.K;
    const u64* coeff = a.coeff + (long long)it * a.L * N;
    const u64* eval = a.eval[it];
    u64* D = a.digits + it * a.dstride + (long long)d * slots * N;
    u64 y[NS];
    double est = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const int limb = t.s0 + i;
        const u64 q = s_sq[i];
        const u64 x = reduce_4q(coeff[(long long)limb * N + c], q);
        y[i] = shoup(x, s_qinv[i], s_qinvs[i], q);
        est = __dadd_rn(est, __dmul_rn((double)y[i], s_sinv[i]));
        if (chunk == 0) D[(long long)limb * N + c] = eval[(long long)limb * N + c];
    }
    u64 al = (u64)__dadd_rn(est, 0.5);
    if (al > (u64)NS) al = (u64)NS;
    for (int jj = 0; jj < nj; ++jj) {
        // sum of NS products < 16 * 2^124 fits in 128 bits; the rows are in
        // Montgomery form, so one REDC gives sum y_i qhat_i mod m
        U128 acc{0, 0};
#pragma unroll
        for (int i = 0; i < NS; ++i) mac128(acc, y[i], s_row[jj][i]);
        const u64 q = s_dq[jj];
        u64 v = redc128_any(acc.lo, acc.hi, q, s_ninv[jj], s_muhi[jj]);
        v = sub_mod(v, s_corr[jj][al], q);
        D[(long long)s_slot[jj] * N + c] = v;
    }
}

#define MI_NS_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

// ---- forward NTT of the ModUp'd slots ----
struct DigitJob {
    u64* digits;
    long long dstride;
    int L, K, nq, alpha, parts, n;
    u64 N;
    FastDiv per_slot, per_item;  // parts * n, parts
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    // slot-major: every item's and digit's slot 0, then slot 1, ... (shared twiddles)
    __device__ void split(int k, int& l, int& it, int& d) const {
        l = per_slot.div(k);
        const int r = k - l * (parts * n);
        it = per_item.div(r);
        d = r - it * parts;
    }
    __device__ View view(int k) const {
        int l, it, d;
        split(k, l, it, d);
        return View{l < L ? l : nq + (l - L), digits + it * dstride + ((long long)d * (L + K) + l) * N};
    }
    __device__ bool skip(int k) const {
        int l, it, d;
        split(k, l, it, d);
        const int s0 = d * alpha;
        return l >= s0 && l < s0 + alpha && l < L;
    }
};

void ks_modup_n(Context& c, int level, int n, const u64* const* key_terms, u64* D, long long dstride,
                u64* C, hipStream_t s, bool cols_only) {
    const int L = level + 1, K = c.np, parts = c.parts(level);
    const u64 N = c.N;
    if (n < 1 || n > KS_MAX_ITEMS) throw Error(-1, "ModUp batch size out of range");
    if (parts > 16) throw Error(-1, "too many digits");
    for (int d = 0; d < parts; ++d)
        if (c.modup[level][d].ns > MAX_ALPHA) throw Error(-1, "digit wider than 16 limbs");
    {
        CopyListJob in{};
        for (int i = 0; i < n; ++i) in.src[i] = key_terms[i];
        in.dst = C;
        in.dstride = (long long)L * (long long)N;
        in.n = n;
        in.N = N;
        in.per_limb = fast_div(n);
        const PlainJob out = plain_job(LimbBatch{C, (long long)L * (long long)N, L, 0, L * n}, N);
        run_ntt_inv(c, in, out, L * n, s);
    }
    int maxnd = 0;
    {
        ModUpArgs a{};
        a.coeff = C;
        for (int i = 0; i < n; ++i) a.eval[i] = key_terms[i];
        a.digits = D;
        a.dstride = dstride;
        a.L = L;
        a.K = K;
        a.parts = parts;
        a.logN = c.logN;
        a.bc = c.d_bc;
        for (int d = 0; d < parts; ++d) {
            a.t[d] = c.modup[level][d];
            maxnd = a.t[d].nd > maxnd ? a.t[d].nd : maxnd;
        }
        a.chunks = (maxnd + MODUP_DCH - 1) / MODUP_DCH;
        if (a.chunks < 1) a.chunks = 1;
        // digits of one width go in one launch (all but the last have alpha limbs)
        int d0 = 0;
        while (d0 < parts) {
            int d1 = d0 + 1;
            while (d1 < parts && a.t[d1].ns == a.t[d0].ns) ++d1;
            ModUpArgs b = a;
            for (int d = d0; d < d1; ++d) b.t[d - d0] = a.t[d];
            // the launch's digit index starts at 0: shift the digit outputs
            b.digits = a.digits + (long long)d0 * (L + K) * (long long)N;
            dim3 grid((unsigned)(N / KS_THREADS), d1 - d0, a.chunks * n);
            switch (a.t[d0].ns) {
#define MI_MODUP_CASE(NS) \
                case NS: hipLaunchKernelGGL(k_modup<NS>, grid, dim3(KS_THREADS), 0, s, b); break;
                MI_NS_CASES(MI_MODUP_CASE)
#undef MI_MODUP_CASE
                default: throw Error(-1, "digit wider than 16 limbs");
            }
            MI_HIP(hipGetLastError());
            d0 = d1;
        }
    }
    DigitJob j{D, dstride, L, K, c.nq, c.alpha, parts, n, N, fast_div(parts * n), fast_div(parts)};
    if (cols_only) launch_pass<true, false>(c, j, parts * (L + K) * n, s);
    else run_ntt_fwd(c, j, j, parts * (L + K) * n, s);
}

void ks_modup(Context& c, int level, const u64* key_term, u64* D, u64* C, hipStream_t s) {
    const u64* kt[1] = {key_term};
    ks_modup_n(c, level, 1, kt, D, 0, C, s);
}

// ---- key inner product ----
// Evaluation-domain source index of output index i under Galois element g:
// dst[brev(d)] = src[brev(((2d+1) g mod 2N - 1) / 2)].
__device__ __forceinline__ u32 galois_src(u32 i, u64 g, int logN) {
    const u32 mask2n = (2u << logN) - 1;
    const u32 d = __brev(i) >> (32 - logN);
    const u32 e = (u32)(((u64)(2 * d + 1) * g) & mask2n);
    return __brev((e - 1) >> 1) >> (32 - logN);
}

__global__ void __launch_bounds__(KS_THREADS) k_mac(MacArgs a) {
    const u64 N = (u64)1 << a.logN;
    const int slots = a.L + a.K;
    const int tps = (int)(N / KS_THREADS);
    const TileMap m = map_tile(blockIdx.x, a.items, tps, slots);
    if (!m.valid) return;
    const int it = m.item, slot = m.slot;
    const long long i = (long long)m.tile * KS_THREADS + threadIdx.x;
    const int basis = slot < a.L ? slot : a.nq + (slot - a.L);
    const int key_limbs = a.nq + a.K;
    const BasisConst b = a.bc[basis];
    const long long src = a.galois ? (long long)galois_src((u32)i, a.galois, a.logN) : i;
    const u64* digits = a.digits + it * a.digits_stride;
    const u64* key = a.key[it];
    // Exact 128-bit sums of the digit products: at most 16 digits of
    // residues below q < 2^62 keep the sum below 2^128.  One reduction at the
    // end (REDC for Montgomery-form keys, Barrett otherwise) gives the
    // canonical residue of the sum.
    U128 sb{0, 0}, sa{0, 0};
    for (int d = 0; d < a.parts; ++d) {
        const u64 x = reduce_4q(digits[((long long)d * slots + slot) * N + src], b.q);
        const u64* kb = key + ((long long)d * 2 * key_limbs + basis) * N;
        const u64* ka = kb + (long long)key_limbs * N;
        const u64 kbv = reduce_4q(kb[src], b.q), kav = reduce_4q(ka[src], b.q);
        mac128(sb, x, kbv);
        mac128(sa, x, kav);
    }
    u64 acc_b, acc_a;
    if (a.montgomery) {
        if (a.parts <= 4) {
            // at most 4 products of residues below q: the sum is below 4 q^2 <
            // q 2^64 (q < 2^62), REDC's precondition, without reducing the high word
            acc_b = redc128(sb.lo, sb.hi, b.q, b.qninv);
            acc_a = redc128(sa.lo, sa.hi, b.q, b.qninv);
        } else {
            acc_b = redc128_any(sb.lo, sb.hi, b.q, b.qninv, b.mu_hi);
            acc_a = redc128_any(sa.lo, sa.hi, b.q, b.qninv, b.mu_hi);
        }
    } else {
        acc_b = reduce128(sb.lo, sb.hi, b.q, b.mu_lo, b.mu_hi);
        acc_a = reduce128(sa.lo, sa.hi, b.q, b.mu_lo, b.mu_hi);
    }
    const u64* ib = a.init_b[it];
    if (ib && slot < a.L) acc_b = add_mod(acc_b, ib[(long long)slot * N + src], b.q);
    if (a.qp_out) {
        const u64* ab = a.add_b[it];
        if (ab) acc_b = add_mod(acc_b, ab[(long long)slot * N + src], b.q);
        u64* pb = a.b_out[it] + (long long)slot * N + i;
        u64* pa = a.a_out[it] + (long long)slot * N + i;
        *pb = a.accumulate ? add_mod(*pb, acc_b, b.q) : acc_b;
        *pa = a.accumulate ? add_mod(*pa, acc_a, b.q) : acc_a;
    } else if (slot < a.L) {
        u64* pb = a.b_out[it] + (long long)slot * N + i;
        u64* pa = a.a_out[it] + (long long)slot * N + i;
        *pb = a.b_mode == 0 ? add_mod(*pb, acc_b, b.q) : acc_b;
        *pa = a.a_mode == 0 ? add_mod(*pa, acc_a, b.q) : acc_a;
    } else {
        const int pi = slot - a.L;
        u64* ap = a.accp + it * a.accp_stride;
        ap[(long long)pi * N + i] = acc_a;
        ap[((long long)a.K + pi) * N + i] = acc_b;
    }
}

void ks_mac(const Context& c, const MacArgs& args, int level, hipStream_t s) {
    if (args.items < 1 || args.items > KS_MAX_ITEMS) throw Error(-1, "key product batch size out of range");
    if (args.parts < 1 || args.parts > 16) throw Error(-1, "too many digits");  // 128-bit sums hold 16 products
    const int tps = (int)(c.N / KS_THREADS);
    const unsigned blocks = tile_grid(args.items, tps, level + 1 + c.np);
    hipLaunchKernelGGL(k_mac, dim3(blocks), dim3(KS_THREADS), 0, s, args);
    MI_HIP(hipGetLastError());
}

MacArgs mac_args(const Context& c, int level, const u64* digits, const u64* key, bool montgomery) {
    MacArgs a{};
    a.digits = digits;
    a.key[0] = key;
    a.items = 1;
    a.L = level + 1;
    a.K = c.np;
    a.nq = c.nq;
    a.parts = c.parts(level);
    a.logN = c.logN;
    a.montgomery = montgomery ? 1 : 0;
    a.bc = c.d_bc;
    return a;
}

// ---- folded ModDown ----
struct MdContribArgs {
    const u64* src[2 * KS_MAX_ITEMS];   // P limbs (coefficient domain) of each poly, K limbs apart by N
    u64* tmp;                           // [npoly][L][N]
    const u64* pinv;                    // [L][K]
    int L, K, nq, logN;
    const BasisConst* bc;
};

// One block: 256 coefficients of one poly, up to MODDOWN_TCH target limbs;
// the P-side constants and the targets' p_j^-1 rows are staged in LDS.
template <int K>
__global__ void __launch_bounds__(KS_THREADS) k_moddown_contrib(MdContribArgs a) {
    __shared__ u64 s_pinv[MODDOWN_TCH][K];
    __shared__ u64 s_tq[MODDOWN_TCH], s_ninv[MODDOWN_TCH], s_muhi[MODDOWN_TCH];
    __shared__ u64 s_pq[K];
    __shared__ double s_pinvd[K];
    const u64 N = (u64)1 << a.logN;
    const int poly = blockIdx.y;
    const int t0 = blockIdx.z * MODDOWN_TCH;
    if (t0 >= a.L) return;  // uniform over the block
    const int nt = t0 + MODDOWN_TCH < a.L ? MODDOWN_TCH : a.L - t0;
    for (int e = threadIdx.x; e < nt * K; e += KS_THREADS)
        s_pinv[e / K][e % K] = a.pinv[(long long)(t0 + e / K) * K + e % K];
    if (threadIdx.x < nt) {
        const BasisConst b = a.bc[t0 + threadIdx.x];
        s_tq[threadIdx.x] = b.q;
        s_ninv[threadIdx.x] = b.qninv;
        s_muhi[threadIdx.x] = b.mu_hi;
    }
    if (threadIdx.x >= 64 && threadIdx.x < 64 + K) {
        const BasisConst pb = a.bc[a.nq + threadIdx.x - 64];
        s_pq[threadIdx.x - 64] = pb.q;
        s_pinvd[threadIdx.x - 64] = pb.inv_d;
    }
    __syncthreads();
    // N is a multiple of KS_THREADS: every thread has a coefficient
    const long long i = (long long)blockIdx.x * KS_THREADS + threadIdx.x;
    const u64* src = a.src[poly];
    u64 x[K];
    double est = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        x[j] = reduce_4q(src[(long long)j * N + i], s_pq[j]);
        est = __dadd_rn(est, __dmul_rn((double)x[j], s_pinvd[j]));
    }
    u64 al = (u64)__dadd_rn(est, 0.5);
    if (al > (u64)K) al = (u64)K;
    for (int tt = 0; tt < nt; ++tt) {
        U128 acc{0, 0};
#pragma unroll
        for (int j = 0; j < K; ++j) mac128(acc, x[j], s_pinv[tt][j]);  // Montgomery-form p_j^-1
        const u64 q = s_tq[tt];
        const u64 v = redc128_any(acc.lo, acc.hi, q, s_ninv[tt], s_muhi[tt]);
        a.tmp[((long long)poly * a.L + t0 + tt) * N + i] = sub_mod(v, al, q);
    }
}

// forward NTT of the contributions, limb-major (all polys of one modulus together)
struct MdInJob {
    u64* tmp;
    int L, npoly;
    u64 N;
    FastDiv per_limb;  // npoly
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const {
        const int t = per_limb.div(k), poly = k - t * npoly;
        return View{t, tmp + ((long long)poly * L + t) * N};
    }
    __device__ bool skip(int) const { return false; }
};

struct MdOutJob {
    const u64* tmp;
    const u64* base[2 * KS_MAX_ITEMS];
    u64* out[2 * KS_MAX_ITEMS];
    int L, npoly;
    u64 N;
    const BasisConst* bc;
    FastDiv per_limb;  // npoly
    struct View {
        int basis;
        const u64* src;
        const u64* base;
        u64* out;
        u64 q;
        __device__ u64 load(int i) const { return src[i]; }
        __device__ u64 pre(int i) const { return base[i]; }
        __device__ void store_pre(int i, u64 v, u64 b) const { out[i] = sub_mod(b, v, q); }
    };
    __device__ View view(int k) const {
        const int t = per_limb.div(k), poly = k - t * npoly;
        return View{t, tmp + ((long long)poly * L + t) * N, base[poly] + (long long)t * N,
                    out[poly] + (long long)t * N, bc[t].q};
    }
    __device__ bool skip(int) const { return false; }
};

void ks_moddown_n(Context& c, int level, int npoly, u64* const* p, const u64* const* base, u64* const* out,
                  u64* T, hipStream_t s) {
    const int L = level + 1, K = c.np;
    const u64 N = c.N;
    if (npoly < 1 || npoly > 2 * KS_MAX_ITEMS) throw Error(-1, "ModDown batch size out of range");
    if (K > MAX_ALPHA) throw Error(-1, "more than 16 special moduli");
    {
        ListJob j{};
        for (int k = 0; k < npoly; ++k) j.p[k] = p[k];
        j.npoly = npoly;
        j.per_limb = fast_div(npoly);
        j.basis0 = c.nq;
        j.N = N;
        run_ntt_inv(c, j, j, K * npoly, s);
    }
    MdContribArgs a{};
    for (int k = 0; k < npoly; ++k) a.src[k] = p[k];
    a.tmp = T;
    a.pinv = c.moddown[level].pinv;
    a.L = L;
    a.K = K;
    a.nq = c.nq;
    a.logN = c.logN;
    a.bc = c.d_bc;
    dim3 grid((unsigned)(N / KS_THREADS), npoly, (L + MODDOWN_TCH - 1) / MODDOWN_TCH);
    switch (K) {
#define MI_MODDOWN_CASE(KK) \
        case KK: hipLaunchKernelGGL(k_moddown_contrib<KK>, grid, dim3(KS_THREADS), 0, s, a); break;
        MI_NS_CASES(MI_MODDOWN_CASE)
#undef MI_MODDOWN_CASE
        default: throw Error(-1, "more than 16 special moduli");
    }
    MI_HIP(hipGetLastError());
    MdInJob in{T, L, npoly, N, fast_div(npoly)};
    MdOutJob o{};
    o.tmp = T;
    for (int k = 0; k < npoly; ++k) {
        o.base[k] = base[k];
        o.out[k] = out[k];
    }
    o.L = L;
    o.npoly = npoly;
    o.per_limb = fast_div(npoly);
    o.N = N;
    o.bc = c.d_bc;
    run_ntt_fwd(c, in, o, npoly * L, s);
}

void ks_moddown(Context& c, int level, u64* p0, u64* p1, const u64* base0, const u64* base1, u64* out0,
                u64* out1, u64* T, hipStream_t s) {
    u64* p[2] = {p0, p1};
    const u64* b[2] = {base0, base1};
    u64* o[2] = {out0, out1};
    ks_moddown_n(c, level, p1 ? 2 : 1, p, b, o, T, s);
}

// ---- ModDown with the key inner product folded in ----
// The split form of k_mac (Q slots combined into the outputs, P slots kept
// for ModDown) computed where ModDown reads it: the P slots' products in the
// load of the inverse NTT's first pass, the Q slots' in the store of the
// forward NTT's last pass (with the combine mode, the b term and the
// subtraction of the NTT'd contribution).  The products are neither written
// nor read back.  Each sum holds PARTS <= 3 canonical products (below q
// 2^64: REDC's precondition), reduced once.
template <bool MONT, int PARTS>
__device__ __forceinline__ u64 key_product(const u64* dig, long long dstep, const u64* key, long long kstep, int i,
                                           const BasisConst& b) {
    U128 s{0, 0};
#pragma unroll
    for (int d = 0; d < PARTS; ++d) mac128(s, reduce_4q(dig[d * dstep + i], b.q), reduce_4q(key[d * kstep + i], b.q));
    return MONT ? redc128(s.lo, s.hi, b.q, b.qninv) : reduce128(s.lo, s.hi, b.q, b.mu_lo, b.mu_hi);
}

struct KeyProductLists {
    const u64* digits;  // item i's digits at digits + i * dstride: [parts][L+K][N]
    long long dstride;
    const u64* key[KS_MAX_ITEMS];
    int L, K, nq;
    u64 N;
    const BasisConst* bc;
    // slot `slot` of poly k (even: a, odd: b) of item k / 2
    __device__ void at(int k, int slot, int basis, const u64*& dig, const u64*& key_p) const {
        const int item = k >> 1;
        dig = digits + item * dstride + (long long)slot * N;
        key_p = key[item] + ((long long)((k & 1) ? 0 : nq + K) + basis) * N;
    }
};

// inverse NTT first (row) pass over the P slots: load = key product
template <bool MONT, int PARTS>
struct MacPJob {
    KeyProductLists kp;
    u64* p[2 * KS_MAX_ITEMS];  // poly k's P accumulator (K limbs)
    int npoly;
    FastDiv per_limb;  // npoly
    struct View {
        int basis;
        u64* p;
        const u64* dig;
        const u64* key;
        long long dstep, kstep;
        BasisConst b;
        __device__ u64 load(int i) const { return key_product<MONT, PARTS>(dig, dstep, key, kstep, i, b); }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const {
        const int pi = per_limb.div(k), poly = k - pi * npoly;
        const int basis = kp.nq + pi;
        View v;
        v.basis = basis;
        v.p = p[poly] + (long long)pi * kp.N;
        kp.at(poly, kp.L + pi, basis, v.dig, v.key);
        v.dstep = (long long)(kp.L + kp.K) * kp.N;
        v.kstep = 2LL * (kp.nq + kp.K) * kp.N;
        v.b = kp.bc[basis];
        return v;
    }
    __device__ bool skip(int) const { return false; }
};

// forward NTT last (row) pass over the Q slots: out = combine(old, key
// product + b term) - NTT(contribution); old and the b term read from a zero
// row where the mode assigns / there is no b term
template <bool MONT, int PARTS>
struct MdMacOutJob {
    KeyProductLists kp;
    const u64* tmp;                        // [npoly][L][N] NTT'd contributions
    u64* out[2 * KS_MAX_ITEMS];
    const u64* old[2 * KS_MAX_ITEMS];      // out (add mode) or the zero row
    const u64* add[2 * KS_MAX_ITEMS];      // the b term or the zero row
    long long old_step[2 * KS_MAX_ITEMS];  // limb step of old / add: N, or 0 for the zero row
    long long add_step[2 * KS_MAX_ITEMS];
    int npoly;
    FastDiv per_limb;  // npoly
    struct View {
        int basis;
        const u64* src;
        u64* out;
        const u64* old;
        const u64* add;
        const u64* dig;
        const u64* key;
        long long dstep, kstep;
        BasisConst b;
        __device__ u64 load(int i) const { return src[i]; }
        __device__ u64 pre(int i) const { return add_mod(old[i], add[i], b.q); }
        __device__ void store_pre(int i, u64 v, u64 base) const {
            const u64 acc = add_mod(key_product<MONT, PARTS>(dig, dstep, key, kstep, i, b), base, b.q);
            out[i] = sub_mod(acc, v, b.q);
        }
    };
    __device__ View view(int k) const {
        const int t = per_limb.div(k), poly = k - t * npoly;
        View v;
        v.basis = t;
        v.src = tmp + ((long long)poly * kp.L + t) * kp.N;
        v.out = out[poly] + (long long)t * kp.N;
        v.old = old[poly] + (long long)t * old_step[poly];
        v.add = add[poly] + (long long)t * add_step[poly];
        kp.at(poly, t, t, v.dig, v.key);
        v.dstep = (long long)(kp.L + kp.K) * kp.N;
        v.kstep = 2LL * (kp.nq + kp.K) * kp.N;
        v.b = kp.bc[t];
        return v;
    }
    __device__ bool skip(int) const { return false; }
};

template <bool MONT, int PARTS>
void moddown_mac_t(Context& c, int level, int n, const KsItem* items, const u64* D, long long dstride, u64* AP,
                   u64* T, hipStream_t s) {
    const int L = level + 1, K = c.np, npoly = 2 * n;
    const u64 N = c.N;
    KeyProductLists kp{};
    kp.digits = D;
    kp.dstride = dstride;
    for (int i = 0; i < n; ++i) kp.key[i] = items[i].key;
    kp.L = L;
    kp.K = K;
    kp.nq = c.nq;
    kp.N = N;
    kp.bc = c.d_bc;
    u64* p[2 * KS_MAX_ITEMS];
    for (int i = 0; i < n; ++i) {
        p[2 * i] = AP + (size_t)i * 2 * K * N;
        p[2 * i + 1] = p[2 * i] + (size_t)K * N;
    }
    {
        MacPJob<MONT, PARTS> in{};
        in.kp = kp;
        for (int k = 0; k < npoly; ++k) in.p[k] = p[k];
        in.npoly = npoly;
        in.per_limb = fast_div(npoly);
        ListJob o{};
        for (int k = 0; k < npoly; ++k) o.p[k] = p[k];
        o.npoly = npoly;
        o.per_limb = fast_div(npoly);
        o.basis0 = c.nq;
        o.N = N;
        const long long probe = c.probe_open(s, K * npoly);
        launch_rows_w4<false>(c, in, K * npoly, s);
        launch_pass<false, false>(c, o, K * npoly, s);
        c.probe_close(probe, s);
    }
    MdContribArgs a{};
    for (int k = 0; k < npoly; ++k) a.src[k] = p[k];
    a.tmp = T;
    a.pinv = c.moddown[level].pinv;
    a.L = L;
    a.K = K;
    a.nq = c.nq;
    a.logN = c.logN;
    a.bc = c.d_bc;
    dim3 grid((unsigned)(N / KS_THREADS), npoly, (L + MODDOWN_TCH - 1) / MODDOWN_TCH);
    switch (K) {
#define MI_MODDOWN_CASE(KK) \
        case KK: hipLaunchKernelGGL(k_moddown_contrib<KK>, grid, dim3(KS_THREADS), 0, s, a); break;
        MI_NS_CASES(MI_MODDOWN_CASE)
#undef MI_MODDOWN_CASE
        default: throw Error(-1, "more than 16 special moduli");
    }
    MI_HIP(hipGetLastError());
    MdInJob in{T, L, npoly, N, fast_div(npoly)};
    MdMacOutJob<MONT, PARTS> o{};
    o.kp = kp;
    o.tmp = T;
    for (int i = 0; i < n; ++i) {
        const KsItem& it = items[i];
        // poly 2i: a, poly 2i+1: b
        o.out[2 * i] = it.a_out;
        o.out[2 * i + 1] = it.b_out;
        o.old[2 * i] = it.a_mode == 0 ? it.a_out : c.d_zero;
        o.old_step[2 * i] = it.a_mode == 0 ? (long long)N : 0;
        o.old[2 * i + 1] = it.b_mode == 0 ? it.b_out : c.d_zero;
        o.old_step[2 * i + 1] = it.b_mode == 0 ? (long long)N : 0;
        o.add[2 * i] = c.d_zero;
        o.add_step[2 * i] = 0;
        o.add[2 * i + 1] = it.init_b ? it.init_b : c.d_zero;
        o.add_step[2 * i + 1] = it.init_b ? (long long)N : 0;
    }
    o.npoly = npoly;
    o.per_limb = fast_div(npoly);
    const long long probe = c.probe_open(s, npoly * L);
    launch_pass<true, false>(c, in, npoly * L, s);
    launch_rows_w4<true>(c, o, npoly * L, s);
    c.probe_close(probe, s);
}

void ks_moddown_mac(Context& c, int level, int n, const KsItem* items, const u64* D, long long dstride, u64* AP,
                    u64* T, bool montgomery, hipStream_t s) {
    if (n < 1 || n > KS_MAX_ITEMS) throw Error(-1, "key switch batch size out of range");
    if (c.np > MAX_ALPHA) throw Error(-1, "more than 16 special moduli");
    const int parts = c.parts(level);
#define MI_MDMAC(M, PP) moddown_mac_t<M, PP>(c, level, n, items, D, dstride, AP, T, s)
    if (montgomery) {
        if (parts == 1) MI_MDMAC(true, 1);
        else if (parts == 2) MI_MDMAC(true, 2);
        else if (parts == 3) MI_MDMAC(true, 3);
        else throw Error(-1, "folded key product takes at most 3 digits");
    } else {
        if (parts == 1) MI_MDMAC(false, 1);
        else if (parts == 2) MI_MDMAC(false, 2);
        else if (parts == 3) MI_MDMAC(false, 3);
        else throw Error(-1, "folded key product takes at most 3 digits");
    }
#undef MI_MDMAC
}

// ---------------------------------------------------------------- entry points
size_t ks_item_words(const Context& c, int level) {
    const size_t L = level + 1, K = c.np, parts = c.parts(level), N = c.N;
    return (L + parts * (L + K) + 2 * K + 2 * L) * N;
}

void keyswitch_items(Context& c, const KsItem* items, int n, bool montgomery, hipStream_t s) {
    if (n < 1 || n > KS_MAX_ITEMS) throw Error(-1, "key switch batch size out of range");
    u64* ws = (u64*)c.workspace(s, sizeof(u64) * ks_item_words(c, items[0].level) * (size_t)n);
    keyswitch_items_with(c, items, n, montgomery, ws, s);
}

void keyswitch_items_with(Context& c, const KsItem* items, int n, bool montgomery, u64* ws, hipStream_t s) {
    if (n < 1 || n > KS_MAX_ITEMS) throw Error(-1, "key switch batch size out of range");
    const int level = items[0].level;
    for (int i = 1; i < n; ++i)
        if (items[i].level != level || items[i].b_mode != items[0].b_mode || items[i].a_mode != items[0].a_mode)
            throw Error(-1, "key switch batch items must share level and combine modes");
    const int L = level + 1, K = c.np, parts = c.parts(level);
    const u64 N = c.N;
    const long long dw = (long long)parts * (L + K) * (long long)N;
    u64* C = ws;
    u64* D = C + (size_t)n * L * N;
    u64* AP = D + (size_t)n * dw;
    u64* T = AP + (size_t)n * 2 * K * N;
    const u64* kt[KS_MAX_ITEMS];
    for (int i = 0; i < n; ++i) kt[i] = items[i].key_term;
    ks_modup_n(c, level, n, kt, D, dw, C, s);
    if (parts <= 3 && c.fuse_mac) {
        ks_moddown_mac(c, level, n, items, D, dw, AP, T, montgomery, s);
        return;
    }
    MacArgs a = mac_args(c, level, D, items[0].key, montgomery);
    a.items = n;
    a.digits_stride = dw;
    a.accp = AP;
    a.accp_stride = 2LL * K * (long long)N;
    a.b_mode = items[0].b_mode;
    a.a_mode = items[0].a_mode;
    for (int i = 0; i < n; ++i) {
        a.key[i] = items[i].key;
        a.b_out[i] = items[i].b_out;
        a.a_out[i] = items[i].a_out;
        a.init_b[i] = items[i].init_b;
    }
    ks_mac(c, a, level, s);
    u64* p[2 * KS_MAX_ITEMS];
    const u64* base[2 * KS_MAX_ITEMS];
    u64* out[2 * KS_MAX_ITEMS];
    for (int i = 0; i < n; ++i) {
        p[2 * i] = AP + (size_t)i * 2 * K * N;
        p[2 * i + 1] = p[2 * i] + (size_t)K * N;
        base[2 * i] = out[2 * i] = items[i].a_out;
        base[2 * i + 1] = out[2 * i + 1] = items[i].b_out;
    }
    ks_moddown_n(c, level, 2 * n, p, base, out, T, s);
}

void keyswitch(Context& c, const KsItem& it, bool montgomery, hipStream_t s) {
    if (it.level < 0 || it.level >= c.nq) throw Error(-1, "invalid ciphertext level");
    keyswitch_items(c, &it, 1, montgomery, s);
}

namespace {
struct Span {
    const u64* lo;
    const u64* hi;
};
bool overlaps(const Span& x, const Span& y) { return x.lo < y.hi && y.lo < x.hi; }
}  // namespace

// Groups of one (level, modes) whose buffers are independent, up to the
// tuned group size (at most KS_MAX_ITEMS) each.  An item that touches a buffer an earlier item of the
// open group of its kind writes (or writes one it reads) closes that group
// and every other open group, so the order of dependent switches is kept.
void keyswitch_batch(Context& c, const KsItem* items, int n, bool montgomery, hipStream_t s) {
    struct Group {
        int level, b_mode, a_mode;
        std::vector<KsItem> v;
        std::vector<Span> writes, reads;
        bool full = false;
    };
    std::vector<Group> open;
    // Open groups (full ones included) are independent of each other: they
    // run spread over the caller's stream and the side streams, forked from
    // it and joined back, so groups overlap each other's fill and drain.
    auto flush_all = [&]() {
        std::vector<const Group*> gs;
        for (const auto& g : open)
            if (!g.v.empty()) gs.push_back(&g);
        if (gs.size() == 1 || (!gs.empty() && !c.multi_stream)) {
            for (const Group* g : gs) keyswitch_items(c, g->v.data(), (int)g->v.size(), montgomery, s);
        } else if (gs.size() > 1) {
            const auto& side = c.side_streams();
            const size_t ns = std::min(gs.size(), side.size() + 1);
            hipEvent_t fork;
            MI_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            MI_HIP(hipEventRecord(fork, s));
            for (size_t k = 1; k < ns; ++k) MI_HIP(hipStreamWaitEvent(side[k - 1], fork, 0));
            for (size_t g = 0; g < gs.size(); ++g) {
                hipStream_t st = g % ns == 0 ? s : side[g % ns - 1];
                keyswitch_items(c, gs[g]->v.data(), (int)gs[g]->v.size(), montgomery, st);
            }
            for (size_t k = 1; k < ns; ++k) {
                hipEvent_t j;
                MI_HIP(hipEventCreateWithFlags(&j, hipEventDisableTiming));
                MI_HIP(hipEventRecord(j, side[k - 1]));
                MI_HIP(hipStreamWaitEvent(s, j, 0));
                MI_HIP(hipEventDestroy(j));
            }
            MI_HIP(hipEventDestroy(fork));
        }
        open.clear();
    };
    for (int i = 0; i < n; ++i) {
        const KsItem& it = items[i];
        const size_t W = (size_t)(it.level + 1) * c.N;
        const Span rd{it.key_term, it.key_term + W};
        const Span ri{it.init_b, it.init_b ? it.init_b + W : it.init_b};
        const Span wb{it.b_out, it.b_out + W}, wa{it.a_out, it.a_out + W};
        bool conflict = false;
        for (const auto& g : open) {
            for (const auto& w : g.writes)
                if (overlaps(w, rd) || overlaps(w, ri) || overlaps(w, wb) || overlaps(w, wa)) conflict = true;
            for (const auto& r : g.reads)
                if (overlaps(r, wb) || overlaps(r, wa)) conflict = true;
        }
        if (conflict) flush_all();
        Group* g = nullptr;
        for (auto& x : open)
            if (!x.full && x.level == it.level && x.b_mode == it.b_mode && x.a_mode == it.a_mode) g = &x;
        if (!g) {
            open.push_back(Group{it.level, it.b_mode, it.a_mode, {}, {}, {}, false});
            g = &open.back();
        }
        g->v.push_back(it);
        g->writes.push_back(wb);
        g->writes.push_back(wa);
        g->reads.push_back(rd);
        if (it.init_b) g->reads.push_back(Span{it.init_b, it.init_b + W});
        if ((int)g->v.size() == c.tune.ks_group) g->full = true;  // runs at the next flush, with the others
    }
    flush_all();
}

void hoist(Context& c, int level, const u64* a_eval, u64* digits, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    u64* C = (u64*)c.workspace(s, sizeof(u64) * (size_t)(level + 1) * c.N);
    ks_modup(c, level, a_eval, digits, C, s);
}

void hoist_batch(Context& c, int level, int n, const u64* const* a_eval, u64* digits, long long dstride,
                 hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    if (n < 1) throw Error(-1, "need at least one ciphertext");
    for (int i0 = 0; i0 < n; i0 += KS_MAX_ITEMS) {
        const int m = std::min(KS_MAX_ITEMS, n - i0);
        u64* C = (u64*)c.workspace(s, sizeof(u64) * (size_t)m * (level + 1) * c.N);
        ks_modup_n(c, level, m, a_eval + i0, digits + (long long)i0 * dstride, dstride, C, s);
    }
}

void keyswitch_hoisted(Context& c, int level, const u64* digits, const u64* key, u64* b_out,
                       u64* a_out, int b_mode, int a_mode, u64 galois, bool montgomery, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    const int L = level + 1, K = c.np;
    const u64 N = c.N;
    const size_t words = (size_t)(2 * K + 2 * L + (galois ? 2 * L : 0)) * N;
    u64* AP = (u64*)c.workspace(s, sizeof(u64) * words);
    u64* T = AP + (size_t)2 * K * N;
    MacArgs a = mac_args(c, level, digits, key, montgomery);
    a.b_out[0] = b_out;
    a.a_out[0] = a_out;
    a.b_mode = b_mode;
    a.a_mode = a_mode;
    a.accp = AP;
    ks_mac(c, a, level, s);
    ks_moddown(c, level, AP, AP + (size_t)K * N, a_out, b_out, a_out, b_out, T, s);
    if (galois) {
        u64* tmp = T + (size_t)2 * L * N;
        const long long W = (long long)L * (long long)N;
        automorphism(c, L, galois, b_out, tmp, 1, W, W, s);
        automorphism(c, L, galois, a_out, tmp + W, 1, W, W, s);
        MI_HIP(hipMemcpyAsync(b_out, tmp, sizeof(u64) * W, hipMemcpyDeviceToDevice, s));
        MI_HIP(hipMemcpyAsync(a_out, tmp + W, sizeof(u64) * W, hipMemcpyDeviceToDevice, s));
    }
}

// dst_a[i], dst_b[i] = automorphism g_i of item i's (a, b) pair at src + i*2*L*N
struct AutoItemsArgs {
    const u64* src;
    u64 galois[KS_MAX_ITEMS];
    u64* dst_a[KS_MAX_ITEMS];
    u64* dst_b[KS_MAX_ITEMS];
    int L, logN;
};

__global__ void __launch_bounds__(KS_THREADS) k_automorphism_items(AutoItemsArgs a) {
    const u64 N = (u64)1 << a.logN;
    const int tps = (int)(N / KS_THREADS);
    const int tile = blockIdx.x % tps, rest = blockIdx.x / tps;
    const int l = rest % a.L, poly = (rest / a.L) & 1, it = rest / (2 * a.L);
    const u32 i = (u32)tile * KS_THREADS + threadIdx.x;
    const u32 sidx = galois_src(i, a.galois[it], a.logN);
    const u64* src = a.src + ((long long)it * 2 + poly) * a.L * N + (long long)l * N;
    u64* dst = (poly == 0 ? a.dst_a[it] : a.dst_b[it]) + (long long)l * N;
    dst[i] = src[sidx];
}

void rotate_hoisted_items(Context& c, int level, const u64* digits, long long dstride, int period,
                          const u64* const* b, int n, const u64* galois, const u64* const* keys, u64* const* b_out,
                          u64* const* a_out, bool montgomery, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    if (n < 1) throw Error(-1, "need at least one rotation");
    if (period < 1 || period > n) period = n;
    const int L = level + 1, K = c.np;
    const u64 N = c.N;
    const long long LN = (long long)L * (long long)N, KN = (long long)K * (long long)N;
    // item i reads the hoist at digits + (i mod period) * dstride: chunks of
    // up to the group size inside each run of `period` items
    const int chunk = std::min(period, c.tune.ks_group);
    const int per_block = (period + chunk - 1) / chunk;
    const int nblocks = (n + period - 1) / period;
    const bool fold = c.parts(level) <= 3 && c.fuse_mac;
    // per rotation: P accumulators (a, b) | switched pair (a, b) | ModDown temp
    const size_t ws_bytes = sizeof(u64) * (size_t)chunk * (2 * KN + 4 * LN);
    over_streams(c, s, nblocks * per_block, ws_bytes, [&](hipStream_t st, u64* ws, int ci) {
        const int blk = ci / per_block, cb = ci - blk * per_block;
        const int i0 = blk * period + cb * chunk;
        if (i0 >= n) return;
        const int m = std::min(std::min(chunk, period - cb * chunk), n - i0);
        u64* AP = ws;
        u64* SW = AP + chunk * 2 * KN;
        u64* Tm = SW + chunk * 2 * LN;
        const u64* dig = digits + (long long)(cb * chunk) * dstride;
        // key products of the hoisted digits, b added in the Q limbs
        // (ckks_hybrid_key_switch_hoisted_inplace with b Add, a Assign),
        // into the staging pairs
        if (fold) {
            KsItem items[KS_MAX_ITEMS];
            for (int it = 0; it < m; ++it)
                items[it] = KsItem{nullptr, keys[i0 + it], SW + it * 2 * LN + LN, SW + it * 2 * LN, level, 1, 1,
                                   b[i0 + it]};
            ks_moddown_mac(c, level, m, items, dig, dstride, AP, Tm, montgomery, st);
        } else {
            MacArgs ma = mac_args(c, level, dig, keys[i0], montgomery);
            ma.items = m;
            ma.digits_stride = dstride;
            ma.accp = AP;
            ma.accp_stride = 2 * KN;
            ma.b_mode = 1;
            ma.a_mode = 1;
            for (int it = 0; it < m; ++it) {
                ma.key[it] = keys[i0 + it];
                ma.a_out[it] = SW + it * 2 * LN;
                ma.b_out[it] = SW + it * 2 * LN + LN;
                ma.init_b[it] = b[i0 + it];
            }
            ks_mac(c, ma, level, st);
            u64* p[2 * KS_MAX_ITEMS];
            const u64* base[2 * KS_MAX_ITEMS];
            u64* out[2 * KS_MAX_ITEMS];
            for (int it = 0; it < m; ++it) {
                p[2 * it] = AP + it * 2 * KN;
                p[2 * it + 1] = AP + it * 2 * KN + KN;
                base[2 * it] = out[2 * it] = SW + it * 2 * LN;
                base[2 * it + 1] = out[2 * it + 1] = SW + it * 2 * LN + LN;
            }
            ks_moddown_n(c, level, 2 * m, p, base, out, Tm, st);
        }
        // the automorphisms (the rotation's NTT-domain map)
        AutoItemsArgs aa{};
        aa.src = SW;
        for (int it = 0; it < m; ++it) {
            aa.galois[it] = galois[i0 + it] % (2 * N);
            aa.dst_a[it] = a_out[i0 + it];
            aa.dst_b[it] = b_out[i0 + it];
        }
        aa.L = L;
        aa.logN = c.logN;
        hipLaunchKernelGGL(k_automorphism_items, dim3((unsigned)(m * 2 * L * (N / KS_THREADS))), dim3(KS_THREADS),
                           0, st, aa);
        MI_HIP(hipGetLastError());
    });
}

void rotate_hoisted_batch(Context& c, int level, const u64* digits, const u64* b, int n, const u64* galois,
                          const u64* const* keys, u64* const* b_out, u64* const* a_out, bool montgomery,
                          hipStream_t s) {
    std::vector<const u64*> bs((size_t)std::max(n, 0), b);
    rotate_hoisted_items(c, level, digits, 0, n, bs.data(), n, galois, keys, b_out, a_out, montgomery, s);
}

// In-place rotations of independent ciphertexts at one level
// (ckks_rotate_inplace, ckks_rotation.cpp:41-96, for each): per chunk of up to
// KS_MAX_ITEMS, the key switches of all a run as one batch into a staging
// area (b enters as the switch's init_b: staged b = b + switched b, a
// assigned), and one launch applies each rotation's automorphism from the
// staging area back into the ciphertexts.
void rotate_batch(Context& c, int level, int n, u64* const* b, u64* const* a, const u64* galois,
                  const u64* const* keys, bool montgomery, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    if (n < 1) throw Error(-1, "need at least one rotation");
    const int L = level + 1;
    const u64 N = c.N;
    const long long LN = (long long)L * (long long)N;
    const int chunk = std::min(n, c.tune.ks_group);
    const size_t kw = ks_item_words(c, level);
    over_streams(c, s, (n + chunk - 1) / chunk, sizeof(u64) * (size_t)chunk * (kw + 2 * LN),
                 [&](hipStream_t st, u64* ws, int ci) {
        const int i0 = ci * chunk;
        const int m = std::min(chunk, n - i0);
        u64* SW = ws + (size_t)chunk * kw;  // per item: a' then b'
        KsItem items[KS_MAX_ITEMS];
        for (int it = 0; it < m; ++it) {
            u64* a_sw = SW + it * 2 * LN;
            u64* b_sw = a_sw + LN;
            items[it] = KsItem{a[i0 + it], keys[i0 + it], b_sw, a_sw, level, 1, 1, b[i0 + it]};
        }
        keyswitch_items_with(c, items, m, montgomery, ws, st);
        AutoItemsArgs aa{};
        aa.src = SW;
        for (int it = 0; it < m; ++it) {
            aa.galois[it] = galois[i0 + it] % (2 * N);
            aa.dst_a[it] = a[i0 + it];
            aa.dst_b[it] = b[i0 + it];
        }
        aa.L = L;
        aa.logN = c.logN;
        hipLaunchKernelGGL(k_automorphism_items, dim3((unsigned)(m * 2 * L * (N / KS_THREADS))), dim3(KS_THREADS),
                           0, st, aa);
        MI_HIP(hipGetLastError());
    });
}

// dst[k] = automorphism g of src[k] (L limbs each), for up to 2 KS_MAX_ITEMS polys
struct AutoListArgs {
    const u64* src[2 * KS_MAX_ITEMS];
    u64* dst[2 * KS_MAX_ITEMS];
    u64 galois;
    int L, logN;
};

__global__ void __launch_bounds__(KS_THREADS) k_automorphism_list(AutoListArgs a) {
    const u64 N = (u64)1 << a.logN;
    const int tps = (int)(N / KS_THREADS);
    const int tile = blockIdx.x % tps, rest = blockIdx.x / tps;
    const int l = rest % a.L, k = rest / a.L;
    const u32 i = (u32)tile * KS_THREADS + threadIdx.x;
    const long long off = (long long)l * N;
    a.dst[k][off + i] = a.src[k][off + galois_src(i, a.galois, a.logN)];
}

void conjugate_batch(Context& c, int level, int n, u64* const* b, u64* const* a, const u64* key, bool montgomery,
                     hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    if (n < 1) return;
    const int L = level + 1;
    const u64 N = c.N;
    const long long LN = (long long)L * (long long)N;
    const u64 g = 2 * N - 1;
    // conjugated (a, b) of every ciphertext, live across the key-switch batch
    // (which uses the stream workspace itself)
    u64* tmp = (u64*)c.pool_alloc(sizeof(u64) * (size_t)n * 2 * LN);
    try {
        for (int i0 = 0; i0 < n; i0 += KS_MAX_ITEMS) {
            const int m = std::min(KS_MAX_ITEMS, n - i0);
            AutoListArgs al{};
            for (int it = 0; it < m; ++it) {
                al.src[2 * it] = a[i0 + it];
                al.src[2 * it + 1] = b[i0 + it];
                al.dst[2 * it] = tmp + (long long)(i0 + it) * 2 * LN;
                al.dst[2 * it + 1] = tmp + (long long)(i0 + it) * 2 * LN + LN;
            }
            al.galois = g;
            al.L = L;
            al.logN = c.logN;
            hipLaunchKernelGGL(k_automorphism_list, dim3((unsigned)(2 * m * L * (N / KS_THREADS))), dim3(KS_THREADS),
                               0, s, al);
            MI_HIP(hipGetLastError());
        }
        // (b, a) <- (sigma(b) + KS_b(sigma(a)), KS_a(sigma(a)))
        std::vector<KsItem> items;
        items.reserve(n);
        for (int k = 0; k < n; ++k) {
            const u64* ca = tmp + (long long)k * 2 * LN;
            items.push_back(KsItem{ca, key, b[k], a[k], level, 1, 1, ca + LN});
        }
        keyswitch_batch(c, items.data(), n, montgomery, s);
    } catch (...) {
        c.pool_release(tmp);
        throw;
    }
    c.pool_release(tmp);
}

void keyswitch_hoisted_qp(Context& c, int level, const u64* digits, const u64* key, const u64* b_q,
                          u64* b_qp, u64* a_qp, u64 galois, bool montgomery, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    MacArgs a = mac_args(c, level, digits, key, montgomery);
    a.qp_out = 1;
    a.b_out[0] = b_qp;
    a.a_out[0] = a_qp;
    a.galois = galois % (2 * c.N);
    a.init_b[0] = b_q;
    ks_mac(c, a, level, s);
}

void keyswitch_qp(Context& c, int level, const u64* key_term, const u64* key, u64* b_qp, u64* a_qp,
                  bool montgomery, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    const int L = level + 1, K = c.np, parts = c.parts(level);
    const u64 N = c.N;
    u64* C = (u64*)c.workspace(s, sizeof(u64) * (size_t)(L + parts * (L + K)) * N);
    u64* D = C + (size_t)L * N;
    ks_modup(c, level, key_term, D, C, s);
    MacArgs a = mac_args(c, level, D, key, montgomery);
    a.qp_out = 1;
    a.b_out[0] = b_qp;
    a.a_out[0] = a_qp;
    ks_mac(c, a, level, s);
}

void moddown_qp(Context& c, int level, u64* qp, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    const int L = level + 1;
    u64* T = (u64*)c.workspace(s, sizeof(u64) * (size_t)L * c.N);
    ks_moddown(c, level, qp + (size_t)L * c.N, nullptr, qp, nullptr, qp, nullptr, T, s);
}

}  // namespace mi
