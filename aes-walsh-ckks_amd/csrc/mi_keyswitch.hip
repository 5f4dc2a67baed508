// Hybrid key switching (reference ckks_hybrid_key_switch_inplace,
// ckks_keyswitch.cpp:1330-1641, fold_hybrid_keyswitch = true).
//
// Per item at level l (L = l+1 Q limbs, K P limbs, beta digits of alpha limbs):
//   1. C  = INTT(key_term)                                  (copy-NTT job)
//   2. D[d] = HPS ModUp of digit d's limbs of C to the other Q limbs and P,
//      with the digit's own limbs copied from the eval-domain key term
//      (ckks_keyswitch.cpp:278-332, 1452)
//   3. NTT of every ModUp'd limb of D (source limbs skipped)
//   4. per slot: acc = sum_d D[d] * key[d] (Montgomery REDC or Barrett);
//      Q slots are combined into b_out / a_out (Add or Assign), P slots kept
//   5. INTT of the 2K P slots
//   6. folded ModDown (ckks_keyswitch.cpp:521-607): contribution
//      sum_j x_j p_j^-1 - alpha per Q limb, NTT, subtracted from the outputs.
// Reordering the final "+ acc" before the "- NTT(contribution)" leaves every
// canonical residue unchanged.
#include "mi_ntt_kernels.hpp"

namespace mi {

constexpr int KS_THREADS = 256;

// ---- step 1: out-of-place inverse NTT ----
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

// ---- step 2: ModUp of all digits ----
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
    const u64 N = (u64)1 << a.logN;
    if (c >= (int)N) return;
    const ModUpTables& t = a.t[d];
    const int slots = a.L + a.K;
    u64* D = a.digits + (long long)d * slots * N;
    u64 y[32];
    double est = 0.0;
    for (int i = 0; i < t.ns; ++i) {
        const int limb = t.s0 + i;
        const BasisConst b = a.bc[limb];
        const u64 x = reduce_4q(a.coeff[(long long)limb * N + c], b.q);
        const u64 s = shoup(x, t.qhat_inv[i], t.qhat_inv_shoup[i], b.q);
        y[i] = s;
        est = __dadd_rn(est, __dmul_rn((double)s, t.src_inv[i]));
        D[(long long)limb * N + c] = a.eval[(long long)limb * N + c];
    }
    u64 al = (u64)__dadd_rn(est, 0.5);
    if (al > (u64)t.ns) al = (u64)t.ns;
    for (int j = 0; j < t.nd; ++j) {
        const int basis = t.dst_basis[j];
        const BasisConst b = a.bc[basis];
        const u64* row = t.qhat_mod_dst + (long long)j * t.ns;
        U128 acc{0, 0};
        for (int i = 0; i < t.ns; ++i) mac128(acc, y[i], row[i]);
        u64 v = reduce128(acc.lo, acc.hi, b.q, b.mu_lo, b.mu_hi);
        v = sub_mod(v, t.alpha_corr[(long long)j * (t.ns + 1) + al], b.q);
        const int slot = j < t.s0 ? j : (j < t.s0 + (a.L - t.s1) ? t.s1 + (j - t.s0) : a.L + (j - (a.L - t.ns)));
        D[(long long)slot * N + c] = v;
    }
}

// ---- step 3: forward NTT of the ModUp'd slots ----
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
    // k walks slot-major (every digit's slot 0, then slot 1, ...) so that
    // co-resident workgroups share one modulus's twiddles.
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

// ---- step 4: key inner product ----
struct MacArgs {
    const u64* digits;    // [parts][L+K][N]
    const u64* key;       // [parts][2][nq+np][N]
    u64* b_out;
    u64* a_out;
    u64* accp;            // [2][K][N]: a then b
    int L, K, nq, parts, logN;
    int b_mode, a_mode, montgomery;
    const BasisConst* bc;
};

__global__ void __launch_bounds__(KS_THREADS) k_mac(MacArgs a) {
    const u64 N = (u64)1 << a.logN;
    const long long i = (long long)blockIdx.x * KS_THREADS + threadIdx.x;
    const int slot = blockIdx.y;
    if (i >= (long long)N) return;
    const int slots = a.L + a.K;
    const int basis = slot < a.L ? slot : a.nq + (slot - a.L);
    const int kl = basis;  // key limb index: Q limbs then P limbs at nq + pi
    const int key_limbs = a.nq + a.K;
    const BasisConst b = a.bc[basis];
    u64 acc_a = 0, acc_b = 0;
    for (int d = 0; d < a.parts; ++d) {
        const u64 x = a.digits[((long long)d * slots + slot) * N + i];
        const u64* kb = a.key + ((long long)d * 2 * key_limbs + kl) * N;
        const u64* ka = kb + (long long)key_limbs * N;
        u64 tb, ta;
        if (a.montgomery) {
            tb = mont_mul(x, kb[i], b.q, b.qninv);
            ta = mont_mul(x, ka[i], b.q, b.qninv);
        } else {
            const u64 xr = reduce_4q(x, b.q);
            tb = mul_mod(xr, reduce_4q(kb[i], b.q), b.q, b.mu_lo, b.mu_hi);
            ta = mul_mod(xr, reduce_4q(ka[i], b.q), b.q, b.mu_lo, b.mu_hi);
        }
        acc_b = add_mod(acc_b, tb, b.q);
        acc_a = add_mod(acc_a, ta, b.q);
    }
    if (slot < a.L) {
        u64* pb = a.b_out + (long long)slot * N + i;
        u64* pa = a.a_out + (long long)slot * N + i;
        *pb = a.b_mode == 0 ? add_mod(*pb, acc_b, b.q) : acc_b;
        *pa = a.a_mode == 0 ? add_mod(*pa, acc_a, b.q) : acc_a;
    } else {
        const int pi = slot - a.L;
        a.accp[((long long)0 * a.K + pi) * N + i] = acc_a;
        a.accp[((long long)1 * a.K + pi) * N + i] = acc_b;
    }
}

// ---- step 6: folded ModDown fused with the forward NTT ----
struct MdInJob {
    const u64* accp;      // [2][K][N] coefficient domain
    u64* tmp;             // [2][L][N]
    const u64* pinv;      // [L][K]
    int L, K, nq;
    u64 N;
    const BasisConst* bc;
    struct View {
        int basis;
        const u64* src;   // K limbs
        u64* out;
        const u64* pinv;  // K entries for this q limb
        const BasisConst* bc;
        int K, nq;
        u64 N, q, mu_lo, mu_hi;
        __device__ u64 load(int i) const {
            U128 acc{0, 0};
            double est = 0.0;
            for (int j = 0; j < K; ++j) {
                const BasisConst pb = bc[nq + j];
                const u64 x = reduce_4q(src[(long long)j * N + i], pb.q);
                est = __dadd_rn(est, __dmul_rn((double)x, pb.inv_d));
                mac128(acc, x, pinv[j]);
            }
            u64 al = (u64)__dadd_rn(est, 0.5);
            if (al > (u64)K) al = (u64)K;
            const u64 v = reduce128(acc.lo, acc.hi, q, mu_lo, mu_hi);
            return sub_mod(v, al, q);
        }
        __device__ void store(int i, u64 v) const { out[i] = v; }
    };
    // k = 2*i + poly: limb-major, a and b of one modulus side by side
    __device__ View view(int k) const {
        const int i = k >> 1, poly = k & 1;
        const BasisConst b = bc[i];
        return View{i, accp + (long long)poly * K * N, tmp + ((long long)poly * L + i) * N,
                    pinv + (long long)i * K, bc, K, nq, N, b.q, b.mu_lo, b.mu_hi};
    }
    __device__ bool skip(int) const { return false; }
};

struct MdOutJob {
    const u64* tmp;
    u64* a_out;
    u64* b_out;
    int L;
    u64 N;
    const BasisConst* bc;
    struct View {
        int basis;
        const u64* src;
        u64* out;
        u64 q;
        __device__ u64 load(int i) const { return src[i]; }
        __device__ void store(int i, u64 v) const { out[i] = sub_mod(out[i], v, q); }
    };
    __device__ View view(int k) const {
        const int i = k >> 1, poly = k & 1;
        u64* o = (poly == 0 ? a_out : b_out) + (long long)i * N;
        return View{i, tmp + ((long long)poly * L + i) * N, o, bc[i].q};
    }
    __device__ bool skip(int) const { return false; }
};

void keyswitch(Context& c, const KsItem& it, bool montgomery, hipStream_t s) {
    const int level = it.level;
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    const int L = level + 1, K = c.np, parts = c.parts(level);
    if (parts > 16) throw Error(-1, "too many digits");
    const u64 N = c.N;
    const size_t words_C = (size_t)L * N, words_D = (size_t)parts * (L + K) * N,
                 words_P = (size_t)2 * K * N, words_T = (size_t)2 * L * N;
    u64* ws = (u64*)c.workspace(s, sizeof(u64) * (words_C + words_D + words_P + words_T));
    u64* C = ws;
    u64* D = C + words_C;
    u64* AP = D + words_D;
    u64* T = AP + words_P;

    // 1. C = INTT(key_term)
    {
        CopyJob in{it.key_term, C, N};
        PlainJob out{LimbBatch{C, (long long)L * (long long)N, L, 0, L}, N};
        run_ntt_inv(c, in, out, L, s);
    }
    // 2. ModUp
    {
        ModUpArgs a{};
        a.coeff = C;
        a.eval = it.key_term;
        a.digits = D;
        a.L = L;
        a.K = K;
        a.parts = parts;
        a.logN = c.logN;
        a.bc = c.d_bc;
        for (int d = 0; d < parts; ++d) a.t[d] = c.modup[level][d];
        dim3 grid((unsigned)((N + KS_THREADS - 1) / KS_THREADS), parts);
        hipLaunchKernelGGL(k_modup, grid, dim3(KS_THREADS), 0, s, a);
        MI_HIP(hipGetLastError());
    }
    // 3. NTT of the ModUp'd slots
    {
        DigitJob j{D, L, K, c.nq, c.alpha, parts, N};
        run_ntt_fwd(c, j, j, parts * (L + K), s);
    }
    // 4. inner product with the key, combine Q slots
    {
        MacArgs a{};
        a.digits = D;
        a.key = it.key;
        a.b_out = it.b_out;
        a.a_out = it.a_out;
        a.accp = AP;
        a.L = L;
        a.K = K;
        a.nq = c.nq;
        a.parts = parts;
        a.logN = c.logN;
        a.b_mode = it.b_mode;
        a.a_mode = it.a_mode;
        a.montgomery = montgomery ? 1 : 0;
        a.bc = c.d_bc;
        dim3 grid((unsigned)((N + KS_THREADS - 1) / KS_THREADS), L + K);
        hipLaunchKernelGGL(k_mac, grid, dim3(KS_THREADS), 0, s, a);
        MI_HIP(hipGetLastError());
    }
    // 5. INTT of the P accumulators (a then b)
    ntt_inverse(c, LimbBatch{AP, (long long)K * (long long)N, K, c.nq, 2 * K}, s);
    // 6. ModDown contribution -> NTT -> subtract
    {
        MdInJob in{AP, T, c.moddown[level].pinv, L, K, c.nq, N, c.d_bc};
        MdOutJob out{T, it.a_out, it.b_out, L, N, c.d_bc};
        run_ntt_fwd(c, in, out, 2 * L, s);
    }
}

}  // namespace mi
