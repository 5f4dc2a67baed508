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

struct Q0P0Args {
    const u64* a;    // q_0 limb of a (eval)
    const u64* ap;   // a moved to p_0 (eval)
    const u64* key;  // [b(q0), b(p0), a(q0), a(p0)]
    u64* acc;        // [aq_a, aq_b, ap_a, ap_b][N]
    u64* b;
    u64* a_out;
    const u64* ct;   // [2][N] NTT'd ModDown contributions (a, b)
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
    const u64 N = g.N;
    const u64 x = g.a[i], y = g.ap[i];
    g.acc[i] = mac1(x, g.key[2 * N + i], g.q0, g.montgomery);
    g.acc[N + i] = mac1(x, g.key[i], g.q0, g.montgomery);
    g.acc[2 * N + i] = mac1(y, g.key[3 * N + i], g.p0, g.montgomery);
    g.acc[3 * N + i] = mac1(y, g.key[N + i], g.p0, g.montgomery);
}

// folded ModDown over one P limb: (sum x p^-1 - alpha) mod q_0, alpha <= 1
__global__ void __launch_bounds__(BT_THREADS) k_q0p0_contrib(Q0P0Args g, u64* ct) {
    const long long i = (long long)blockIdx.x * BT_THREADS + threadIdx.x;
    if (i >= (long long)g.N) return;
    const int poly = blockIdx.y;
    const u64 v = reduce_4q(g.acc[(2 + poly) * g.N + i], g.p0.q);
    u64 al = (u64)__dadd_rn(__dmul_rn((double)v, g.p0.inv_d), 0.5);
    if (al > 1) al = 1;
    const u64 sv = shoup(reduce64(v, g.q0.q, g.q0.mu_hi), g.inv, g.inv_shoup, g.q0.q);
    ct[poly * g.N + i] = sub_mod(sv, al, g.q0.q);
}

__global__ void __launch_bounds__(BT_THREADS) k_q0p0_finish(Q0P0Args g) {
    const long long i = (long long)blockIdx.x * BT_THREADS + threadIdx.x;
    if (i >= (long long)g.N) return;
    const u64 q = g.q0.q;
    const u64 na = sub_mod(g.acc[i], g.ct[i], q);
    const u64 nb = sub_mod(g.acc[g.N + i], g.ct[g.N + i], q);
    g.b[i] = add_mod(g.b[i], nb, q);
    g.a_out[i] = na;
}

void q0p0_keyswitch(Context& c, u64* b, u64* a, const u64* key, bool montgomery, hipStream_t s) {
    const u64 N = c.N;
    // scratch: coeff | moved | acc[4] | contributions[2]
    u64* S = (u64*)c.workspace(s, sizeof(u64) * 8 * N);
    u64* coeff = S;
    u64* moved = S + N;
    u64* acc = S + 2 * N;
    u64* ct = S + 6 * N;
    {
        OneLimbCopy in{a, coeff, 0};
        const PlainJob o = plain_job(LimbBatch{coeff, (long long)N, 1, 0, 1}, N);
        run_ntt_inv(c, in, o, 1, s);
    }
    {
        ChangeJob in{coeff, moved, c.nq, c.q[0], c.p[0], c.h_bc[c.nq].mu_hi};
        const PlainJob o = plain_job(LimbBatch{moved, (long long)N, 1, c.nq, 1}, N);
        run_ntt_fwd(c, in, o, 1, s);
    }
    Q0P0Args g{};
    g.a = a;
    g.ap = moved;
    g.key = key;
    g.acc = acc;
    g.b = b;
    g.a_out = a;
    g.ct = ct;
    g.N = N;
    g.montgomery = montgomery ? 1 : 0;
    g.q0 = c.h_bc[0];
    g.p0 = c.h_bc[c.nq];
    g.inv = c.q0p0_inv;
    g.inv_shoup = c.q0p0_inv_shoup;
    const unsigned blocks = (unsigned)((N + BT_THREADS - 1) / BT_THREADS);
    hipLaunchKernelGGL(k_q0p0_mac, dim3(blocks), dim3(BT_THREADS), 0, s, g);
    MI_HIP(hipGetLastError());
    ntt_inverse(c, LimbBatch{acc + 2 * N, (long long)N, 1, c.nq, 2}, s);
    hipLaunchKernelGGL(k_q0p0_contrib, dim3(blocks, 2), dim3(BT_THREADS), 0, s, g, ct);
    MI_HIP(hipGetLastError());
    ntt_forward(c, LimbBatch{ct, (long long)N, 1, 0, 2}, s);
    hipLaunchKernelGGL(k_q0p0_finish, dim3(blocks), dim3(BT_THREADS), 0, s, g);
    MI_HIP(hipGetLastError());
}

}  // namespace mi
