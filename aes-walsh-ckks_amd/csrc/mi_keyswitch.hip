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

namespace mi {

constexpr int KS_THREADS = 256;
constexpr int MAX_ALPHA = 16;   // digit width held in registers
constexpr int MODUP_DCH = 8;    // destination limbs per ModUp thread
constexpr int MODDOWN_TCH = 4;  // target limbs per ModDown-contribution thread

// ---- out-of-place inverse NTT (key term -> C) ----
struct CopyJob {
    const u64* src;
    u64* dst;
    u64 N;
    struct View {
        int basis;
        const u64* s;
        u64* d;
        __device__ u64 load(int i) const { return s[i]; }
        __device__ void store(int i, u64 v) const { d[i] = v; }
    };
    __device__ View view(int k) const {
        return View{k, src + (long long)k * N, dst + (long long)k * N};
    }
    __device__ bool skip(int) const { return false; }
};

// ---- ModUp of all digits ----
struct ModUpArgs {
    const u64* coeff;     // C [L][N]
    const u64* eval;      // key term [L][N]
    u64* digits;          // D [parts][L+K][N]
    int L, K, parts;
    int logN;
    ModUpTables t[16];
    const BasisConst* bc;
};

__global__ void __launch_bounds__(KS_THREADS) k_modup(ModUpArgs a) {
    const int c = blockIdx.x * KS_THREADS + threadIdx.x;
    const int d = blockIdx.y;
    const int chunk = blockIdx.z;
    const u64 N = (u64)1 << a.logN;
    if (c >= (int)N) return;
    const ModUpTables& t = a.t[d];
    const int j0 = chunk * MODUP_DCH;
    if (j0 >= t.nd) return;
    const int j1 = j0 + MODUP_DCH < t.nd ? j0 + MODUP_DCH : t.nd;
    const int slots = a.L + a.K;
    u64* D = a.digits + (long long)d * slots * N;
    u64 y[MAX_ALPHA];
    double est = 0.0;
#pragma unroll
    for (int i = 0; i < MAX_ALPHA; ++i) {
        y[i] = 0;
        if (i < t.ns) {
            const int limb = t.s0 + i;
            const u64 q = a.bc[limb].q;
            const u64 x = reduce_4q(a.coeff[(long long)limb * N + c], q);
            const u64 s = shoup(x, t.qhat_inv[i], t.qhat_inv_shoup[i], q);
            y[i] = s;
            est = __dadd_rn(est, __dmul_rn((double)s, t.src_inv[i]));
            if (chunk == 0) D[(long long)limb * N + c] = a.eval[(long long)limb * N + c];
        }
    }
    u64 al = (u64)__dadd_rn(est, 0.5);
    if (al > (u64)t.ns) al = (u64)t.ns;
    for (int j = j0; j < j1; ++j) {
        const BasisConst b = a.bc[t.dst_basis[j]];
        const u64* row = t.qhat_mod_dst + (long long)j * t.ns;
        U128 acc{0, 0};
#pragma unroll
        for (int i = 0; i < MAX_ALPHA; ++i)
            if (i < t.ns) mac128(acc, y[i], row[i]);
        u64 v = reduce128(acc.lo, acc.hi, b.q, b.mu_lo, b.mu_hi);
        v = sub_mod(v, t.alpha_corr[(long long)j * (t.ns + 1) + al], b.q);
        const int slot =
            j < t.s0 ? j : (j < t.s0 + (a.L - t.s1) ? t.s1 + (j - t.s0) : a.L + (j - (a.L - t.ns)));
        D[(long long)slot * N + c] = v;
    }
}

// ---- forward NTT of the ModUp'd slots ----
struct DigitJob {
    u64* digits;
    int L, K, nq, alpha, parts;
    u64 N;
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    // slot-major: every digit's slot 0, then slot 1, ... (shared twiddles)
    __device__ View view(int k) const {
        const int l = k / parts, d = k - l * parts;
        return View{l < L ? l : nq + (l - L), digits + ((long long)d * (L + K) + l) * N};
    }
    __device__ bool skip(int k) const {
        const int l = k / parts, d = k - l * parts;
        const int s0 = d * alpha;
        return l >= s0 && l < s0 + alpha && l < L;
    }
};

void ks_modup(Context& c, int level, const u64* key_term, u64* D, u64* C, hipStream_t s) {
    const int L = level + 1, K = c.np, parts = c.parts(level);
    const u64 N = c.N;
    if (parts > 16) throw Error(-1, "too many digits");
    for (int d = 0; d < parts; ++d)
        if (c.modup[level][d].ns > MAX_ALPHA) throw Error(-1, "digit wider than 16 limbs");
    {
        CopyJob in{key_term, C, N};
        PlainJob out{LimbBatch{C, (long long)L * (long long)N, L, 0, L}, N};
        run_ntt_inv(c, in, out, L, s);
    }
    {
        ModUpArgs a{};
        a.coeff = C;
        a.eval = key_term;
        a.digits = D;
        a.L = L;
        a.K = K;
        a.parts = parts;
        a.logN = c.logN;
        a.bc = c.d_bc;
        int maxnd = 0;
        for (int d = 0; d < parts; ++d) {
            a.t[d] = c.modup[level][d];
            maxnd = a.t[d].nd > maxnd ? a.t[d].nd : maxnd;
        }
        dim3 grid((unsigned)((N + KS_THREADS - 1) / KS_THREADS), parts,
                  (maxnd + MODUP_DCH - 1) / MODUP_DCH);
        hipLaunchKernelGGL(k_modup, grid, dim3(KS_THREADS), 0, s, a);
        MI_HIP(hipGetLastError());
    }
    DigitJob j{D, L, K, c.nq, c.alpha, parts, N};
    run_ntt_fwd(c, j, j, parts * (L + K), s);
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
    const long long i = (long long)blockIdx.x * KS_THREADS + threadIdx.x;
    const int slot = blockIdx.y;
    if (i >= (long long)N) return;
    const int slots = a.L + a.K;
    const int basis = slot < a.L ? slot : a.nq + (slot - a.L);
    const int key_limbs = a.nq + a.K;
    const BasisConst b = a.bc[basis];
    const long long src = a.galois ? (long long)galois_src((u32)i, a.galois, a.logN) : i;
    // Lazy 128-bit sums of the digit products, brought below q every three
    // products (3 q^2 + q < q 2^64 for q < 2^62) and reduced once at the end:
    // REDC for Montgomery-form keys, Barrett otherwise.  The result is the
    // canonical residue of the per-product sum.
    U128 sb{0, 0}, sa{0, 0};
    for (int d = 0; d < a.parts; ++d) {
        u64 x = a.digits[((long long)d * slots + slot) * N + src];
        const u64* kb = a.key + ((long long)d * 2 * key_limbs + basis) * N;
        const u64* ka = kb + (long long)key_limbs * N;
        u64 kbv = kb[src], kav = ka[src];
        if (!a.montgomery) {
            x = reduce_4q(x, b.q);
            kbv = reduce_4q(kbv, b.q);
            kav = reduce_4q(kav, b.q);
        }
        mac128(sb, x, kbv);
        mac128(sa, x, kav);
        if (d % 3 == 2 && d + 1 < a.parts) {
            sb = U128{reduce128(sb.lo, sb.hi, b.q, b.mu_lo, b.mu_hi), 0};
            sa = U128{reduce128(sa.lo, sa.hi, b.q, b.mu_lo, b.mu_hi), 0};
        }
    }
    u64 acc_b, acc_a;
    if (a.montgomery) {
        acc_b = redc128(sb.lo, sb.hi, b.q, b.qninv);
        acc_a = redc128(sa.lo, sa.hi, b.q, b.qninv);
    } else {
        acc_b = reduce128(sb.lo, sb.hi, b.q, b.mu_lo, b.mu_hi);
        acc_a = reduce128(sa.lo, sa.hi, b.q, b.mu_lo, b.mu_hi);
    }
    if (a.qp_out) {
        if (a.init_b && slot < a.L) acc_b = add_mod(acc_b, a.init_b[(long long)slot * N + src], b.q);
        if (a.add_b) acc_b = add_mod(acc_b, a.add_b[(long long)slot * N + src], b.q);
        u64* pb = a.b_out + (long long)slot * N + i;
        u64* pa = a.a_out + (long long)slot * N + i;
        *pb = a.accumulate ? add_mod(*pb, acc_b, b.q) : acc_b;
        *pa = a.accumulate ? add_mod(*pa, acc_a, b.q) : acc_a;
    } else if (slot < a.L) {
        u64* pb = a.b_out + (long long)slot * N + i;
        u64* pa = a.a_out + (long long)slot * N + i;
        *pb = a.b_mode == 0 ? add_mod(*pb, acc_b, b.q) : acc_b;
        *pa = a.a_mode == 0 ? add_mod(*pa, acc_a, b.q) : acc_a;
    } else {
        const int pi = slot - a.L;
        a.accp_a[(long long)pi * N + i] = acc_a;
        a.accp_b[(long long)pi * N + i] = acc_b;
    }
}

void ks_mac(const Context& c, const MacArgs& args, int level, hipStream_t s) {
    const u64 N = c.N;
    dim3 grid((unsigned)((N + KS_THREADS - 1) / KS_THREADS), level + 1 + c.np);
    hipLaunchKernelGGL(k_mac, grid, dim3(KS_THREADS), 0, s, args);
    MI_HIP(hipGetLastError());
}

MacArgs mac_args(const Context& c, int level, const u64* digits, const u64* key, bool montgomery) {
    MacArgs a{};
    a.digits = digits;
    a.key = key;
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
    const u64* src[2];    // P limbs (coefficient domain) of each poly, K limbs apart by N
    u64* tmp;             // [npoly][L][N]
    const u64* pinv;      // [L][K]
    int L, K, nq, logN;
    const BasisConst* bc;
};

__global__ void __launch_bounds__(KS_THREADS) k_moddown_contrib(MdContribArgs a) {
    const u64 N = (u64)1 << a.logN;
    const long long i = (long long)blockIdx.x * KS_THREADS + threadIdx.x;
    const int poly = blockIdx.y;
    const int t0 = blockIdx.z * MODDOWN_TCH;
    if (i >= (long long)N || t0 >= a.L) return;
    const int t1 = t0 + MODDOWN_TCH < a.L ? t0 + MODDOWN_TCH : a.L;
    const u64* src = a.src[poly];
    u64 x[MAX_ALPHA];
    double est = 0.0;
#pragma unroll
    for (int j = 0; j < MAX_ALPHA; ++j) {
        x[j] = 0;
        if (j < a.K) {
            const BasisConst pb = a.bc[a.nq + j];
            x[j] = reduce_4q(src[(long long)j * N + i], pb.q);
            est = __dadd_rn(est, __dmul_rn((double)x[j], pb.inv_d));
        }
    }
    u64 al = (u64)__dadd_rn(est, 0.5);
    if (al > (u64)a.K) al = (u64)a.K;
    for (int t = t0; t < t1; ++t) {
        const BasisConst b = a.bc[t];
        const u64* pinv = a.pinv + (long long)t * a.K;
        U128 acc{0, 0};
#pragma unroll
        for (int j = 0; j < MAX_ALPHA; ++j)
            if (j < a.K) mac128(acc, x[j], pinv[j]);
        const u64 v = reduce128(acc.lo, acc.hi, b.q, b.mu_lo, b.mu_hi);
        a.tmp[((long long)poly * a.L + t) * N + i] = sub_mod(v, al, b.q);
    }
}

// forward NTT of the contributions, limb-major (a and b of one modulus together)
struct MdInJob {
    u64* tmp;
    int L, npoly;
    u64 N;
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const {
        const int t = k / npoly, poly = k - t * npoly;
        return View{t, tmp + ((long long)poly * L + t) * N};
    }
    __device__ bool skip(int) const { return false; }
};

struct MdOutJob {
    const u64* tmp;
    const u64* base[2];
    u64* out[2];
    int L, npoly;
    u64 N;
    const BasisConst* bc;
    struct View {
        int basis;
        const u64* src;
        const u64* base;
        u64* out;
        u64 q;
        __device__ u64 load(int i) const { return src[i]; }
        __device__ void store(int i, u64 v) const { out[i] = sub_mod(base[i], v, q); }
    };
    __device__ View view(int k) const {
        const int t = k / npoly, poly = k - t * npoly;
        return View{t, tmp + ((long long)poly * L + t) * N, base[poly] + (long long)t * N,
                    out[poly] + (long long)t * N, bc[t].q};
    }
    __device__ bool skip(int) const { return false; }
};

void ks_moddown(Context& c, int level, u64* p0, u64* p1, const u64* base0, const u64* base1, u64* out0,
                u64* out1, u64* T, hipStream_t s) {
    const int L = level + 1, K = c.np, npoly = p1 ? 2 : 1;
    const u64 N = c.N;
    // both polys' P limbs in one batch (the callers' P blocks never overlap)
    if (p1)
        ntt_inverse(c, LimbBatch{p0, (long long)(p1 - p0), K, c.nq, 2 * K}, s);
    else
        ntt_inverse(c, LimbBatch{p0, (long long)K * (long long)N, K, c.nq, K}, s);
    MdContribArgs a{};
    a.src[0] = p0;
    a.src[1] = p1 ? p1 : p0;
    a.tmp = T;
    a.pinv = c.moddown[level].pinv;
    a.L = L;
    a.K = K;
    a.nq = c.nq;
    a.logN = c.logN;
    a.bc = c.d_bc;
    if (K > MAX_ALPHA) throw Error(-1, "more than 16 special moduli");
    dim3 grid((unsigned)((N + KS_THREADS - 1) / KS_THREADS), npoly, (L + MODDOWN_TCH - 1) / MODDOWN_TCH);
    hipLaunchKernelGGL(k_moddown_contrib, grid, dim3(KS_THREADS), 0, s, a);
    MI_HIP(hipGetLastError());
    MdInJob in{T, L, npoly, N};
    MdOutJob out{T, {base0, base1 ? base1 : base0}, {out0, out1 ? out1 : out0}, L, npoly, N, c.d_bc};
    run_ntt_fwd(c, in, out, npoly * L, s);
}

// ---------------------------------------------------------------- entry points
static size_t ks_scratch_words(const Context& c, int level) {
    const size_t L = level + 1, K = c.np, parts = c.parts(level), N = c.N;
    return (L + parts * (L + K) + 2 * K + 2 * L) * N;
}

void keyswitch_with(Context& c, const KsItem& it, bool montgomery, u64* ws, hipStream_t s) {
    const int level = it.level;
    const int L = level + 1, K = c.np, parts = c.parts(level);
    const u64 N = c.N;
    u64* C = ws;
    u64* D = C + (size_t)L * N;
    u64* AP = D + (size_t)parts * (L + K) * N;
    u64* T = AP + (size_t)2 * K * N;
    ks_modup(c, level, it.key_term, D, C, s);
    MacArgs a = mac_args(c, level, D, it.key, montgomery);
    a.b_out = it.b_out;
    a.a_out = it.a_out;
    a.b_mode = it.b_mode;
    a.a_mode = it.a_mode;
    a.accp_a = AP;
    a.accp_b = AP + (size_t)K * N;
    ks_mac(c, a, level, s);
    ks_moddown(c, level, AP, AP + (size_t)K * N, it.a_out, it.b_out, it.a_out, it.b_out, T, s);
}

void keyswitch(Context& c, const KsItem& it, bool montgomery, hipStream_t s) {
    if (it.level < 0 || it.level >= c.nq) throw Error(-1, "invalid ciphertext level");
    u64* ws = (u64*)c.workspace(s, sizeof(u64) * ks_scratch_words(c, it.level));
    keyswitch_with(c, it, montgomery, ws, s);
}

void hoist(Context& c, int level, const u64* a_eval, u64* digits, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    u64* C = (u64*)c.workspace(s, sizeof(u64) * (size_t)(level + 1) * c.N);
    ks_modup(c, level, a_eval, digits, C, s);
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
    a.b_out = b_out;
    a.a_out = a_out;
    a.b_mode = b_mode;
    a.a_mode = a_mode;
    a.accp_a = AP;
    a.accp_b = AP + (size_t)K * N;
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

void keyswitch_hoisted_qp(Context& c, int level, const u64* digits, const u64* key, const u64* b_q,
                          u64* b_qp, u64* a_qp, u64 galois, bool montgomery, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    MacArgs a = mac_args(c, level, digits, key, montgomery);
    a.qp_out = 1;
    a.b_out = b_qp;
    a.a_out = a_qp;
    a.galois = galois % (2 * c.N);
    a.init_b = b_q;
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
    a.b_out = b_qp;
    a.a_out = a_qp;
    ks_mac(c, a, level, s);
}

void moddown_qp(Context& c, int level, u64* qp, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    const int L = level + 1;
    u64* T = (u64*)c.workspace(s, sizeof(u64) * (size_t)L * c.N);
    ks_moddown(c, level, qp + (size_t)L * c.N, nullptr, qp, nullptr, qp, nullptr, T, s);
}

}  // namespace mi
