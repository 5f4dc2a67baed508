// Bootstrapping pieces on the device: ModRaise (modraise_eval_poly_from_q0,
// ckks_modraise.cpp:31-75) and the level-0 dense-to-sparse switch over
// {q_0, p_0} (ckks_q0p0_key_switch_level0_inplace, ckks_keyswitch.cpp:1661-1816).
#include "mi_ntt_kernels.hpp"

namespace mi {

constexpr int BT_THREADS = 256;

// one limb, out of place: used as the input job of an inverse NTT
struct OneLimbCopy {
    const u64* src;
    u64* dst;
    int basis;
    struct View {
        int basis;
        const u64* s;
        u64* d;
        __device__ u64 load(int i) const { return s[i]; }
        __device__ void store(int i, u64 v) const { d[i] = v; }
    };
    __device__ View view(int) const { return View{basis, src, dst}; }
    __device__ bool skip(int) const { return false; }
};

// ---------------------------------------------------------------- ModRaise
// Forward-NTT column-pass input: the centred q_0 coefficient reduced to q_k.
struct LiftJob {
    const u64* coeff0;   // q_0 limb, coefficient domain, canonical
    u64* out;            // [target+1][N]
    u64 N, q0;
    const BasisConst* bc;
    struct View {
        int basis;
        const u64* c0;
        u64* o;
        u64 q0, q, mu_hi;
        __device__ u64 load(int i) const {
            const u64 v = c0[i];
            if (v > (q0 >> 1)) {
                const u64 m = reduce64(q0 - v, q, mu_hi);
                return m == 0 ? 0 : q - m;
            }
            return reduce64(v, q, mu_hi);
        }
        __device__ void store(int i, u64 v) const { o[i] = v; }
    };
    __device__ View view(int k) const {
        const BasisConst b = bc[k];
        return View{k, coeff0, out + (long long)k * N, q0, b.q, b.mu_hi};
    }
    __device__ bool skip(int) const { return false; }
};

void modraise(Context& c, int target_level, const u64* poly_q0, u64* out, hipStream_t s) {
    if (target_level < 1 || target_level >= c.nq) throw Error(-1, "invalid target level");
    const u64 N = c.N;
    const int L = target_level + 1;
    u64* S = (u64*)c.workspace(s, sizeof(u64) * N);
    {
        OneLimbCopy in{poly_q0, S, 0};
        const PlainJob o = plain_job(LimbBatch{S, (long long)N, 1, 0, 1}, N);
        run_ntt_inv(c, in, o, 1, s);
    }
    LiftJob lift{S, out, N, c.q[0], c.d_bc};
    const PlainJob rows = plain_job(LimbBatch{out, (long long)L * (long long)N, L, 0, L}, N);
    run_ntt_fwd(c, lift, rows, L, s);
}

// ModRaise of up to KS_MAX_ITEMS polys as one launch sequence: their q_0
// limbs copied to the workspace by the first inverse pass, then every item's
// target limbs lifted and transformed together (limb-major, item-minor).
struct Q0ListCopy {
    const u64* src[KS_MAX_ITEMS];
    u64* dst;  // [n][N]
    u64 N;
    struct View {
        int basis;
        const u64* s;
        u64* d;
        __device__ u64 load(int i) const { return s[i]; }
        __device__ void store(int i, u64 v) const { d[i] = v; }
    };
    __device__ View view(int k) const { return View{0, src[k], dst + (long long)k * N}; }
    __device__ bool skip(int) const { return false; }
};
struct LiftListJob {
    const u64* coeff0;  // [n][N], coefficient domain
    u64* out[KS_MAX_ITEMS];
    u64 N, q0;
    int n;
    FastDiv per_limb;  // n
    const BasisConst* bc;
    __device__ LiftJob::View view(int k) const {
        const int l = per_limb.div(k), it = k - l * n;
        const BasisConst b = bc[l];
        return LiftJob::View{l, coeff0 + (long long)it * N, out[it] + (long long)l * N, q0, b.q, b.mu_hi};
    }
    __device__ bool skip(int) const { return false; }
};
struct LimbListJob {
    u64* p[KS_MAX_ITEMS];
    u64 N;
    int n;
    FastDiv per_limb;  // n
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const {
        const int l = per_limb.div(k), it = k - l * n;
        return View{l, p[it] + (long long)l * N};
    }
    __device__ bool skip(int) const { return false; }
};

void modraise_batch(Context& c, int target_level, int n, const u64* const* poly_q0, u64* const* out, hipStream_t s) {
    if (target_level < 1 || target_level >= c.nq) throw Error(-1, "invalid target level");
    if (n < 1) throw Error(-1, "need at least one poly");
    const u64 N = c.N;
    const int L = target_level + 1;
    for (int i0 = 0; i0 < n; i0 += KS_MAX_ITEMS) {
        const int m = std::min(KS_MAX_ITEMS, n - i0);
        u64* S = (u64*)c.workspace(s, sizeof(u64) * (size_t)m * N);
        Q0ListCopy in{};
        for (int k = 0; k < m; ++k) in.src[k] = poly_q0[i0 + k];
        in.dst = S;
        in.N = N;
        const PlainJob o = plain_job(LimbBatch{S, (long long)N, 1, 0, m}, N);
        run_ntt_inv(c, in, o, m, s);
        LiftListJob lift{};
        LimbListJob rows{};
        for (int k = 0; k < m; ++k) lift.out[k] = rows.p[k] = out[i0 + k];
        lift.coeff0 = S;
        lift.N = rows.N = N;
        lift.q0 = c.q[0];
        lift.n = rows.n = m;
        lift.per_limb = rows.per_limb = fast_div(m);
        lift.bc = c.d_bc;
        run_ntt_fwd(c, lift, rows, L * m, s);
    }
}

// ---------------------------------------------------------------- q0p0 switch
// hps_base_change_single_limb (ckks_keyswitch.cpp:1647-1659): centred change
// of a residue mod src to mod dst.
__device__ __forceinline__ u64 limb_change(u64 v, u64 src, u64 dst, u64 dst_mu_hi) {
    const u64 r = reduce_4q(v, src);
    const u64 r_dst = reduce64(r, dst, dst_mu_hi);
    const u64 src_dst = reduce64(src, dst, dst_mu_hi);
    return r >= (src >> 1) + 1 ? sub_mod(r_dst, src_dst, dst) : r_dst;
}

// Forward-NTT (over p_0) column-pass input: the q_0 coefficients moved to p_0.
struct ChangeJob {
    const u64* coeff;
    u64* out;
    int basis_p0;
    u64 q0, p0, p0_mu_hi;
    struct View {
        int basis;
        const u64* c;
        u64* o;
        u64 q0, p0, mu;
        __device__ u64 load(int i) const { return limb_change(c[i], q0, p0, mu); }
        __device__ void store(int i, u64 v) const { o[i] = v; }
    };
    __device__ View view(int) const { return View{basis_p0, coeff, out, q0, p0, p0_mu_hi}; }
    __device__ bool skip(int) const { return false; }
};

// Per item (ciphertext) k of a batch: a[k], b[k] its q_0 limbs (eval), and
// workspace rows moved/accq/accp/ct at k (N words per row, two rows where
// noted), so every phase covers the whole batch in one launch.
struct Q0P0Args {
    const u64* a[KS_MAX_ITEMS];   // q_0 limb of a (eval)
    u64* b[KS_MAX_ITEMS];
    u64* a_out[KS_MAX_ITEMS];
    const u64* ap;   // [n][N] a moved to p_0 (eval)
    const u64* key;  // [b(q0), b(p0), a(q0), a(p0)]
    u64* accq;       // [n][2][N] (a, b) products over q_0
    u64* accp;       // [n][2][N] (a, b) products over p_0
    const u64* ct;   // [n][2][N] NTT'd ModDown contributions (a, b)
    u64 N;
    int montgomery;
    BasisConst q0, p0;
    u64 inv, inv_shoup;   // p_0^-1 mod q_0
};

__device__ __forceinline__ u64 mac1(u64 x, u64 k, const BasisConst& b, int mont) {
    if (mont) return mont_mul(x, k, b.q, b.qninv);
    return mul_mod(reduce_4q(x, b.q), reduce_4q(k, b.q), b.q, b.mu_lo, b.mu_hi);
}

__global__ void __launch_bounds__(BT_THREADS) k_q0p0_mac(Q0P0Args g) {
    const long long i = (long long)blockIdx.x * BT_THREADS + threadIdx.x;
    if (i >= (long long)g.N) return;
    const int k = blockIdx.y;
    const u64 N = g.N;
    const u64 x = g.a[k][i], y = g.ap[k * N + i];
    u64* aq = g.accq + 2 * k * N;
    u64* ap = g.accp + 2 * k * N;
    aq[i] = mac1(x, g.key[2 * N + i], g.q0, g.montgomery);
    aq[N + i] = mac1(x, g.key[i], g.q0, g.montgomery);
    ap[i] = mac1(y, g.key[3 * N + i], g.p0, g.montgomery);
    ap[N + i] = mac1(y, g.key[N + i], g.p0, g.montgomery);
}

// folded ModDown over one P limb: (sum x p^-1 - alpha) mod q_0, alpha <= 1
__global__ void __launch_bounds__(BT_THREADS) k_q0p0_contrib(Q0P0Args g, u64* ct) {
    const long long i = (long long)blockIdx.x * BT_THREADS + threadIdx.x;
    if (i >= (long long)g.N) return;
    const long long row = blockIdx.y;  // item * 2 + poly
    const u64 v = reduce_4q(g.accp[row * g.N + i], g.p0.q);
    u64 al = (u64)__dadd_rn(__dmul_rn((double)v, g.p0.inv_d), 0.5);
    if (al > 1) al = 1;
    const u64 sv = shoup(reduce64(v, g.q0.q, g.q0.mu_hi), g.inv, g.inv_shoup, g.q0.q);
    ct[row * g.N + i] = sub_mod(sv, al, g.q0.q);
}

__global__ void __launch_bounds__(BT_THREADS) k_q0p0_finish(Q0P0Args g) {
    const long long i = (long long)blockIdx.x * BT_THREADS + threadIdx.x;
    if (i >= (long long)g.N) return;
    const int k = blockIdx.y;
    const u64 q = g.q0.q, N = g.N;
    const u64 na = sub_mod(g.accq[2 * k * N + i], g.ct[2 * k * N + i], q);
    const u64 nb = sub_mod(g.accq[(2 * k + 1) * N + i], g.ct[(2 * k + 1) * N + i], q);
    g.b[k][i] = add_mod(g.b[k][i], nb, q);
    g.a_out[k][i] = na;
}

// Forward-NTT (over p_0) column-pass input of a batch: item k's q_0
// coefficients moved to p_0.
struct ChangeListJob {
    const u64* coeff;  // [n][N]
    u64* out;          // [n][N]
    int basis_p0;
    u64 N, q0, p0, p0_mu_hi;
    __device__ ChangeJob::View view(int k) const {
        return ChangeJob::View{basis_p0, coeff + (long long)k * N, out + (long long)k * N, q0, p0, p0_mu_hi};
    }
    __device__ bool skip(int) const { return false; }
};

void q0p0_keyswitch_batch(Context& c, int n, u64* const* b, u64* const* a, const u64* key, bool montgomery,
                          hipStream_t s) {
    if (n < 1) throw Error(-1, "need at least one ciphertext");
    const u64 N = c.N;
    for (int i0 = 0; i0 < n; i0 += KS_MAX_ITEMS) {
        const int m = std::min(KS_MAX_ITEMS, n - i0);
        const long long mN = (long long)m * (long long)N;
        // scratch: coeff [m] | moved [m] | accq [m][2] | accp [m][2] | contributions [m][2]
        u64* S = (u64*)c.workspace(s, sizeof(u64) * 8 * (size_t)mN);
        u64* coeff = S;
        u64* moved = S + mN;
        u64* accq = S + 2 * mN;
        u64* accp = S + 4 * mN;
        u64* ct = S + 6 * mN;
        {
            Q0ListCopy in{};
            for (int k = 0; k < m; ++k) in.src[k] = a[i0 + k];
            in.dst = coeff;
            in.N = N;
            const PlainJob o = plain_job(LimbBatch{coeff, (long long)N, 1, 0, m}, N);
            run_ntt_inv(c, in, o, m, s);
        }
        {
            ChangeListJob in{coeff, moved, c.nq, N, c.q[0], c.p[0], c.h_bc[c.nq].mu_hi};
            const PlainJob o = plain_job(LimbBatch{moved, (long long)N, 1, c.nq, m}, N);
            run_ntt_fwd(c, in, o, m, s);
        }
        Q0P0Args g{};
        for (int k = 0; k < m; ++k) {
            g.a[k] = a[i0 + k];
            g.b[k] = b[i0 + k];
            g.a_out[k] = a[i0 + k];
        }
        g.ap = moved;
        g.key = key;
        g.accq = accq;
        g.accp = accp;
        g.ct = ct;
        g.N = N;
        g.montgomery = montgomery ? 1 : 0;
        g.q0 = c.h_bc[0];
        g.p0 = c.h_bc[c.nq];
        g.inv = c.q0p0_inv;
        g.inv_shoup = c.q0p0_inv_shoup;
        const unsigned blocks = (unsigned)((N + BT_THREADS - 1) / BT_THREADS);
        hipLaunchKernelGGL(k_q0p0_mac, dim3(blocks, m), dim3(BT_THREADS), 0, s, g);
        MI_HIP(hipGetLastError());
        ntt_inverse(c, LimbBatch{accp, (long long)N, 1, c.nq, 2 * m}, s);
        hipLaunchKernelGGL(k_q0p0_contrib, dim3(blocks, 2 * m), dim3(BT_THREADS), 0, s, g, ct);
        MI_HIP(hipGetLastError());
        ntt_forward(c, LimbBatch{ct, (long long)N, 1, 0, 2 * m}, s);
        hipLaunchKernelGGL(k_q0p0_finish, dim3(blocks, m), dim3(BT_THREADS), 0, s, g);
        MI_HIP(hipGetLastError());
    }
}

void q0p0_keyswitch(Context& c, u64* b, u64* a, const u64* key, bool montgomery, hipStream_t s) {
    q0p0_keyswitch_batch(c, 1, &b, &a, key, montgomery, s);
}

}  // namespace mi
