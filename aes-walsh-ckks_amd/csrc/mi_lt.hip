// BSGS linear-transform stage on the QP evaluator
// (ckks_apply_linear_transform_owned_stage_qp, ckks_linear_transform.cpp:1174-1407).
//
// With baby_steps = bs and the terms' diagonals d = g*bs + r:
//   baby[0]  = ([a;0], [b;0])                         (make_direct_input_qp, :487-532)
//   baby[r]  = hoisted QP rotation of (b, a) by r      (ckks_keyswitch.cpp:675-852)
//   inner_g  = sum_{terms of g} baby[r] * diag_t       (QP basis, per limb)
//   g = 0:     out += inner_0
//   g > 0:     KS = QP key switch of ModDown(inner_g.a) with the key of g*bs,
//              out.a += rot_{g*bs}(KS.a), out.b += rot_{g*bs}(KS.b + inner_g.b)
//              (automorphism_qp_pair_accumulate, :373-416)
//   result   = folded ModDown of out                  (moddown_qp_pair_to_q_eval, :641-718)
// The automorphism is folded into the inner-product kernel's loads, so the
// rotated baby steps and giant-step accumulations never exist unpermuted.
#include "mi_ntt_kernels.hpp"

#include <vector>

namespace mi {

constexpr int LT_THREADS = 256;
constexpr int LT_MAX_TERMS = 32;  // terms per inner-product launch

struct LtTerm {
    const u64* a;   // baby step a (QP, or Q only when q_only)
    const u64* b;
    const u64* pt;  // diagonal, L+K limbs
    int q_only;     // baby step 0: P limbs are zero
    int pad;
};

struct LtInnerArgs {
    LtTerm t[LT_MAX_TERMS];
    int nterms;
    u64* out_a;
    u64* out_b;
    int accumulate;
    int L, K, nq, logN, pt_mont;
    const BasisConst* bc;
};

__global__ void __launch_bounds__(LT_THREADS) k_lt_inner(LtInnerArgs a) {
    const u64 N = (u64)1 << a.logN;
    const long long i = (long long)blockIdx.x * LT_THREADS + threadIdx.x;
    const int slot = blockIdx.y;
    if (i >= (long long)N) return;
    const int basis = slot < a.L ? slot : a.nq + (slot - a.L);
    const BasisConst b = a.bc[basis];
    const long long o = (long long)slot * N + i;
    // lazy 128-bit sums, brought below q every three products and reduced
    // once (REDC for Montgomery-form plaintexts, Barrett otherwise)
    U128 sa{0, 0}, sb{0, 0};
    int pending = 0;
    for (int t = 0; t < a.nterms; ++t) {
        const LtTerm& tm = a.t[t];
        if (tm.q_only && slot >= a.L) continue;
        const u64 p = tm.pt[o];
        mac128(sa, tm.a[o], p);
        mac128(sb, tm.b[o], p);
        if (++pending == 3) {
            sa = U128{reduce128(sa.lo, sa.hi, b.q, b.mu_lo, b.mu_hi), 0};
            sb = U128{reduce128(sb.lo, sb.hi, b.q, b.mu_lo, b.mu_hi), 0};
            pending = 0;
        }
    }
    u64 ra, rb;
    if (a.pt_mont) {
        ra = redc128(sa.lo, sa.hi, b.q, b.qninv);
        rb = redc128(sb.lo, sb.hi, b.q, b.qninv);
    } else {
        ra = reduce128(sa.lo, sa.hi, b.q, b.mu_lo, b.mu_hi);
        rb = reduce128(sb.lo, sb.hi, b.q, b.mu_lo, b.mu_hi);
    }
    a.out_a[o] = a.accumulate ? add_mod(a.out_a[o], ra, b.q) : ra;
    a.out_b[o] = a.accumulate ? add_mod(a.out_b[o], rb, b.q) : rb;
}

static u64 powmod_u64(u64 a, u64 e, u64 m) {
    unsigned __int128 r = 1, x = a % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (u64)r;
}

// Galois element of a slot rotation (rotation_key_generation.cpp:73-78):
// 5^(r mod N/2) mod 2N.
u64 rotation_galois(u64 N, long long rotation) {
    const long long slots = (long long)(N >> 1);
    long long r = rotation % slots;
    if (r < 0) r += slots;
    return powmod_u64(5, (u64)r, N << 1);
}

void linear_transform_qp(Context& c, const LtStage& st, const u64* a, const u64* b, u64* a_out,
                         u64* b_out, hipStream_t s) {
    const int level = st.level;
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    if (st.nterms < 1 || st.baby_steps < 1 || st.slots < 1)
        throw Error(-1, "need nterms >= 1, baby_steps >= 1, slots >= 1");
    if (!a || !b || !a_out || !b_out || !st.diag_idx || !st.diags)
        throw Error(-1, "null pointer argument");
    if (st.nkeys > 0 && (!st.key_rot || !st.keys)) throw Error(-1, "null key table");
    const int L = level + 1, K = c.np, T = L + K, parts = c.parts(level), bs = st.baby_steps;
    const u64 N = c.N;
    const size_t W = (size_t)T * N;
    const int giants = (st.slots + bs - 1) / bs;

    // compile_linear_transform_stage_metadata (:818-859): needed baby steps
    std::vector<int> baby_slot(bs, -1);
    int nbaby = 0;
    for (int t = 0; t < st.nterms; ++t) {
        const long long d = st.diag_idx[t];
        if (d < 0) throw Error(-1, "diagonal indices must be non-negative");
        if (!st.diags[t]) throw Error(-1, "null diagonal");
        if (d / bs >= giants) continue;
        const int r = (int)(d % bs);
        if (r > 0 && baby_slot[r] < 0) baby_slot[r] = nbaby++;
    }
    auto key_for = [&](long long r) -> const u64* {
        for (int i = 0; i < st.nkeys; ++i)
            if (st.key_rot[i] == r) {
                if (!st.keys[i]) throw Error(-1, "null rotation key");
                return st.keys[i];
            }
        throw Error(-1, "missing rotation key for step " + std::to_string(r));
    };
    // check every key up front so that no work is queued on failure
    for (int r = 1; r < bs; ++r)
        if (baby_slot[r] >= 0) key_for(r);
    for (int g = 1; g < giants; ++g)
        for (int t = 0; t < st.nterms; ++t)
            if (st.diag_idx[t] / bs == g) {
                key_for((long long)g * bs);
                break;
            }

    // scratch: digits | C | babies (a, b) | out (a, b) | inner (a, b) | T
    const size_t words = (size_t)parts * W + (size_t)L * N + 2 * (size_t)nbaby * W + 4 * W +
                         (size_t)2 * L * N;
    u64* ws = (u64*)c.workspace(s, sizeof(u64) * words);
    u64* D = ws;
    u64* C = D + (size_t)parts * W;
    u64* BA = C + (size_t)L * N;
    u64* BB = BA + (size_t)nbaby * W;
    u64* OA = BB + (size_t)nbaby * W;
    u64* OB = OA + W;
    u64* IA = OB + W;
    u64* IB = IA + W;
    u64* Tm = IB + W;

    if (nbaby > 0) {
        ks_modup(c, level, a, D, C, s);
        for (int r = 1; r < bs; ++r) {
            if (baby_slot[r] < 0) continue;
            MacArgs m = mac_args(c, level, D, key_for(r), st.montgomery != 0);
            m.qp_out = 1;
            m.b_out = BB + (size_t)baby_slot[r] * W;
            m.a_out = BA + (size_t)baby_slot[r] * W;
            m.galois = rotation_galois(N, r);
            m.init_b = b;
            ks_mac(c, m, level, s);
        }
    }

    bool have_out = false;
    const dim3 grid((unsigned)((N + LT_THREADS - 1) / LT_THREADS), T);
    std::vector<int> terms;
    for (int g = 0; g < giants; ++g) {
        terms.clear();
        for (int t = 0; t < st.nterms; ++t)
            if (st.diag_idx[t] / bs == g) terms.push_back(t);
        if (terms.empty()) continue;
        u64* dst_a = g == 0 ? OA : IA;
        u64* dst_b = g == 0 ? OB : IB;
        bool acc = g == 0 ? have_out : false;
        for (size_t t0 = 0; t0 < terms.size(); t0 += LT_MAX_TERMS) {
            LtInnerArgs ia{};
            ia.nterms = (int)std::min<size_t>(LT_MAX_TERMS, terms.size() - t0);
            for (int j = 0; j < ia.nterms; ++j) {
                const int t = terms[t0 + j];
                const int r = (int)(st.diag_idx[t] - (long long)g * bs);
                LtTerm& tm = ia.t[j];
                tm.pt = st.diags[t];
                if (r == 0) {
                    tm.a = a;
                    tm.b = b;
                    tm.q_only = 1;
                } else {
                    tm.a = BA + (size_t)baby_slot[r] * W;
                    tm.b = BB + (size_t)baby_slot[r] * W;
                    tm.q_only = 0;
                }
            }
            ia.out_a = dst_a;
            ia.out_b = dst_b;
            ia.accumulate = acc ? 1 : 0;
            ia.L = L;
            ia.K = K;
            ia.nq = c.nq;
            ia.logN = c.logN;
            ia.pt_mont = st.pt_montgomery ? 1 : 0;
            ia.bc = c.d_bc;
            hipLaunchKernelGGL(k_lt_inner, grid, dim3(LT_THREADS), 0, s, ia);
            MI_HIP(hipGetLastError());
            acc = true;
        }
        if (g == 0) {
            have_out = true;
            continue;
        }
        const long long gr = (long long)g * bs;
        // key term = ModDown(inner.a); QP switch; rotate-accumulate with inner.b
        ks_moddown(c, level, IA + (size_t)L * N, nullptr, IA, nullptr, IA, nullptr, Tm, s);
        ks_modup(c, level, IA, D, C, s);
        MacArgs m = mac_args(c, level, D, key_for(gr), st.montgomery != 0);
        m.qp_out = 1;
        m.b_out = OB;
        m.a_out = OA;
        m.galois = rotation_galois(N, gr);
        m.add_b = IB;
        m.accumulate = have_out ? 1 : 0;
        ks_mac(c, m, level, s);
        have_out = true;
    }
    if (!have_out) throw Error(-1, "no diagonal falls inside the slot range");
    ks_moddown(c, level, OA + (size_t)L * N, OB + (size_t)L * N, OA, OB, a_out, b_out, Tm, s);
}

}  // namespace mi
