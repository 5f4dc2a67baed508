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

#include <algorithm>
#include <vector>

namespace mi {

constexpr int LT_THREADS = 256;
constexpr int LT_NB = 16;  // baby-step slots held in registers per launch (slot 0 = the input)
constexpr int LT_NG = 16;  // giant steps per launch

// Inner products of every giant step.  One block computes one giant step's
// sum for 256 coefficients of one limb of one ciphertext, streaming its (up
// to 16) baby/diagonal pairs: few registers, and a block touches only the
// memory of its own terms.  Block order (map_tile_g) puts the blocks of every
// giant step and ciphertext for one coefficient tile consecutively on one
// XCD, so the tile's baby steps (read once per giant step) and diagonals (read
// once per ciphertext) are served from that XCD's L2: HBM sees each baby step
// once per ciphertext and each diagonal once per batch.
struct LtAllArgs {
    const u64* diag[LT_NG * LT_NB];     // [giant][baby slot], null: no term
    const u64* in_a[KS_MAX_ITEMS];      // baby slot 0 (Q limbs; its P limbs are zero)
    const u64* in_b[KS_MAX_ITEMS];
    const u64* babies;                  // item i's baby s (s >= 1): a at babies +
    long long baby_item_stride;         //   i*item_stride + (s-1)*slot_stride, b T*N later
    long long baby_slot_stride;
    u64* out[LT_NG];                    // giant g of item i: a at out[g] + i*out_stride[g], b T*N later
    long long out_stride[LT_NG];
    int slot0;                          // baby slot of this launch's first term (multiple of 16)
    int nb, ng, accumulate;
    int items, L, K, nq, logN, pt_mont;
    const BasisConst* bc;
};

__global__ void __launch_bounds__(LT_THREADS) k_lt_inner_all(LtAllArgs a) {
    const u64 N = (u64)1 << a.logN;
    const int T = a.L + a.K;
    const int tps = (int)(N / LT_THREADS);
    // block -> (giant, item, slot, tile): giant fastest, then item, per XCD
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int g = local % a.ng, rest = local / a.ng;
    const int it = rest % a.items, tl = rest / a.items;
    const int tile = tl * 8 + xcd;
    const int slot = tile / tps;
    if (slot >= T) return;
    const long long i = (long long)(tile - slot * tps) * LT_THREADS + threadIdx.x;
    const int basis = slot < a.L ? slot : a.nq + (slot - a.L);
    const BasisConst b = a.bc[basis];
    const long long o = (long long)slot * N + i;
    const long long TN = (long long)T * N;
    // exact 128-bit sums: at most LT_NB = 16 products of residues below
    // 2^62 (plaintexts brought below q first) stay below 2^128; one reduction
    // at the end (REDC for Montgomery-form plaintexts, Barrett otherwise)
    // gives the canonical residue of the sum
    U128 sa{0, 0}, sb{0, 0};
    for (int r = 0; r < a.nb; ++r) {
        const u64* dp = a.diag[g * LT_NB + r];
        if (!dp) continue;
        const int sl = a.slot0 + r;
        u64 xa, xb;
        if (sl == 0) {
            if (slot >= a.L) continue;  // the input's P limbs are zero
            xa = a.in_a[it][o];
            xb = a.in_b[it][o];
        } else {
            const u64* p = a.babies + it * a.baby_item_stride + (long long)(sl - 1) * a.baby_slot_stride + o;
            xa = p[0];
            xb = p[TN];
        }
        const u64 d = reduce_4q(dp[o], b.q);
        mac128(sa, xa, d);
        mac128(sb, xb, d);
    }
    u64 ra, rb;
    if (a.pt_mont) {
        ra = redc128_any(sa.lo, sa.hi, b.q, b.qninv, b.mu_hi);
        rb = redc128_any(sb.lo, sb.hi, b.q, b.qninv, b.mu_hi);
    } else {
        ra = reduce128(sa.lo, sa.hi, b.q, b.mu_lo, b.mu_hi);
        rb = reduce128(sb.lo, sb.hi, b.q, b.mu_lo, b.mu_hi);
    }
    u64* pa = a.out[g] + it * a.out_stride[g] + o;
    u64* pb = pa + TN;
    *pa = a.accumulate ? add_mod(*pa, ra, b.q) : ra;
    *pb = a.accumulate ? add_mod(*pb, rb, b.q) : rb;
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

namespace {

// compile_linear_transform_stage_metadata (ckks_linear_transform.cpp:720-756)
// plus the key lookups: baby steps in use get slots 1.. in order of first use
// (slot 0 is the input itself); giant steps in use are listed in order.
struct Compiled {
    int bs = 0, nslots = 1, ngp = 0;
    bool g0 = false;
    std::vector<int> baby_slot;            // rotation r < bs -> slot, -1 unused
    std::vector<int> glist;                // giant indices in use, ascending
    std::vector<const u64*> table;         // [giant index][slot] diagonal
    std::vector<const u64*> baby_key;      // [slot]
    std::vector<u64> baby_gal;             // [slot]
    std::vector<const u64*> giant_key;     // [giant index]
};

Compiled compile_stage(const Context& c, const LtStage& st, bool q_eval) {
    const int level = st.level;
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid ciphertext level");
    if (st.nterms < 1 || st.baby_steps < 1 || st.slots < 1)
        throw Error(-1, "need nterms >= 1, baby_steps >= 1, slots >= 1");
    if (!st.diag_idx || !st.diags) throw Error(-1, "null pointer argument");
    if (st.nkeys > 0 && (!st.key_rot || !st.keys)) throw Error(-1, "null key table");
    Compiled k;
    const int bs = k.bs = st.baby_steps;
    const int giants = (st.slots + bs - 1) / bs;
    k.baby_slot.assign(bs, -1);
    k.baby_slot[0] = 0;
    std::vector<int> gidx(giants, -1);
    for (int t = 0; t < st.nterms; ++t) {
        const long long d = st.diag_idx[t];
        if (d < 0) throw Error(-1, "diagonal indices must be non-negative");
        if (!st.diags[t]) throw Error(-1, "null diagonal");
        if (d / bs >= giants) {
            // the Q evaluator's validation (:758-814) rejects such a term;
            // the QP evaluator leaves it out
            if (q_eval) throw Error(-1, "diagonal index exceeds slot count");
            continue;
        }
        const int r = (int)(d % bs), g = (int)(d / bs);
        if (k.baby_slot[r] < 0) k.baby_slot[r] = k.nslots++;
        if (gidx[g] < 0) {
            gidx[g] = 0;
            k.glist.push_back(g);
        }
    }
    if (k.glist.empty()) throw Error(-1, "no diagonal falls inside the slot range");
    std::sort(k.glist.begin(), k.glist.end());
    for (size_t j = 0; j < k.glist.size(); ++j) gidx[k.glist[j]] = (int)j;
    k.g0 = k.glist[0] == 0;
    k.ngp = (int)k.glist.size() - (k.g0 ? 1 : 0);
    k.table.assign(k.glist.size() * (size_t)k.nslots, nullptr);
    for (int t = 0; t < st.nterms; ++t) {
        const long long d = st.diag_idx[t];
        if (d / bs >= giants) continue;
        const u64*& e = k.table[(size_t)gidx[d / bs] * k.nslots + k.baby_slot[d % bs]];
        if (e) throw Error(-1, "duplicate diagonal index " + std::to_string(d));
        e = st.diags[t];
    }
    auto key_for = [&](long long r) -> const u64* {
        for (int i = 0; i < st.nkeys; ++i)
            if (st.key_rot[i] == r) {
                if (!st.keys[i]) throw Error(-1, "null rotation key");
                return st.keys[i];
            }
        throw Error(-1, "missing rotation key for step " + std::to_string(r));
    };
    // every key is looked up before any work is queued
    k.baby_key.assign(k.nslots, nullptr);
    k.baby_gal.assign(k.nslots, 0);
    for (int r = 1; r < bs; ++r)
        if (k.baby_slot[r] > 0) {
            k.baby_key[k.baby_slot[r]] = key_for(r);
            k.baby_gal[k.baby_slot[r]] = rotation_galois(c.N, r);
        }
    k.giant_key.assign(k.glist.size(), nullptr);
    for (size_t j = 0; j < k.glist.size(); ++j)
        if (k.glist[j] > 0) k.giant_key[j] = key_for((long long)k.glist[j] * bs);
    return k;
}

// Inner products of every giant step for m items.  Giant 0 goes to out0 +
// i*out0_stride, giant j > 0 (in use order) to outg + (j-1)*outg_giant +
// i*outg_item; K = 0 for the Q evaluator.
void launch_inner(Context& c, const Compiled& k, const LtStage& st, int m, int K, const u64* const* a,
                  const u64* const* b, const u64* babies, long long baby_item, long long baby_slot, u64* out0,
                  long long out0_stride, u64* outg, long long outg_giant, long long outg_item, hipStream_t s) {
    const int L = st.level + 1, T = L + K;
    const int tps = (int)(c.N / LT_THREADS);
    for (int s0 = 0; s0 < k.nslots; s0 += LT_NB) {
        for (size_t gk0 = 0; gk0 < k.glist.size(); gk0 += LT_NG) {
            LtAllArgs ia{};
            ia.nb = std::min(LT_NB, k.nslots - s0);
            ia.ng = (int)std::min<size_t>(LT_NG, k.glist.size() - gk0);
            ia.slot0 = s0;
            for (int g = 0; g < ia.ng; ++g) {
                const size_t gk = gk0 + g;
                for (int r = 0; r < ia.nb; ++r) ia.diag[g * LT_NB + r] = k.table[gk * k.nslots + s0 + r];
                if (k.glist[gk] == 0) {
                    ia.out[g] = out0;
                    ia.out_stride[g] = out0_stride;
                } else {
                    ia.out[g] = outg + (long long)(gk - (k.g0 ? 1 : 0)) * outg_giant;
                    ia.out_stride[g] = outg_item;
                }
            }
            for (int it = 0; it < m; ++it) {
                ia.in_a[it] = a[it];
                ia.in_b[it] = b[it];
            }
            ia.babies = babies;
            ia.baby_item_stride = baby_item;
            ia.baby_slot_stride = baby_slot;
            ia.accumulate = s0 > 0 ? 1 : 0;
            ia.items = m;
            ia.L = L;
            ia.K = K;
            ia.nq = c.nq;
            ia.logN = c.logN;
            ia.pt_mont = st.pt_montgomery ? 1 : 0;
            ia.bc = c.d_bc;
            hipLaunchKernelGGL(k_lt_inner_all, dim3(tile_grid(m * ia.ng, tps, T)), dim3(LT_THREADS), 0, s, ia);
            MI_HIP(hipGetLastError());
        }
    }
}

void check_batch(int n, const u64* const* a, const u64* const* b, u64* const* a_out, u64* const* b_out) {
    if (n < 1 || !a || !b || !a_out || !b_out) throw Error(-1, "null pointer argument");
    for (int i = 0; i < n; ++i)
        if (!a[i] || !b[i] || !a_out[i] || !b_out[i]) throw Error(-1, "null pointer argument");
}

}  // namespace

void linear_transform_qp_batch(Context& c, const LtStage& st, int n, const u64* const* a,
                               const u64* const* b, u64* const* a_out, u64* const* b_out, hipStream_t s) {
    const Compiled k = compile_stage(c, st, false);
    check_batch(n, a, b, a_out, b_out);
    const int level = st.level;
    const int L = level + 1, K = c.np, T = L + K, parts = c.parts(level);
    const u64 N = c.N;
    const long long W = (long long)T * (long long)N, LN = (long long)L * (long long)N;
    const int nbaby = k.nslots - 1, ngp = k.ngp;

    // per-item scratch: digits | C | babies (a, b) | out (a, b) | giant inners (a, b) | ModDown temp
    const long long dw = (long long)parts * W;
    const long long item_words = dw + LN + 2 * (long long)nbaby * W + 2 * W + 2 * (long long)ngp * W + 2 * LN;
    const long long budget_words = (32LL << 30) / 8;
    int chunk = (int)std::min<long long>(KS_MAX_ITEMS, std::max<long long>(1, budget_words / item_words));
    chunk = std::min(chunk, n);
    u64* ws = (u64*)c.workspace(s, sizeof(u64) * (size_t)(item_words * chunk));

    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = std::min(chunk, n - i0);
        u64* D = ws;
        u64* C = D + m * dw;
        u64* BAB = C + m * LN;
        u64* OAB = BAB + m * 2 * (long long)nbaby * W;
        u64* IAB = OAB + m * 2 * W;
        u64* Tm = IAB + m * 2 * (long long)ngp * W;
        auto baby = [&](int it, int sl) { return BAB + it * 2LL * nbaby * W + (long long)(sl - 1) * 2 * W; };
        auto inner = [&](int it, int gp) { return IAB + it * 2LL * ngp * W + (long long)gp * 2 * W; };
        auto outp = [&](int it) { return OAB + it * 2 * W; };

        // baby steps: one hoisted ModUp, then one key product per baby step
        // (hoisted QP rotations, :675-852), b added in the Q limbs
        if (nbaby > 0) {
            ks_modup_n(c, level, m, a + i0, D, dw, C, s);
            // the baby steps' key products are independent (each writes its
            // own baby slots from the shared digits): spread over the streams
            over_streams(c, s, k.nslots - 1, 0, [&](hipStream_t bst, u64*, int i) {
                const int sl = i + 1;
                MacArgs ma = mac_args(c, level, D, k.baby_key[sl], st.montgomery != 0);
                ma.items = m;
                ma.digits_stride = dw;
                ma.qp_out = 1;
                ma.galois = k.baby_gal[sl];
                for (int it = 0; it < m; ++it) {
                    ma.key[it] = k.baby_key[sl];
                    ma.a_out[it] = baby(it, sl);
                    ma.b_out[it] = baby(it, sl) + W;
                    ma.init_b[it] = b[i0 + it];
                }
                ks_mac(c, ma, level, bst);
            });
        }

        launch_inner(c, k, st, m, K, a + i0, b + i0, BAB, 2LL * nbaby * W, 2 * W, OAB, 2 * W, IAB, 2 * W,
                     2LL * ngp * W, s);

        // giant steps > 0: key term = ModDown(inner.a); QP switch; rotate-accumulate
        // with inner.b (automorphism_qp_pair_accumulate, :373-416)
        bool have_out = k.g0;
        u64* p[2 * KS_MAX_ITEMS];
        const u64* base[2 * KS_MAX_ITEMS];
        u64* outs[2 * KS_MAX_ITEMS];
        const u64* kt[KS_MAX_ITEMS];
        for (int gp = 0; gp < ngp; ++gp) {
            const size_t gk = (size_t)gp + (k.g0 ? 1 : 0);
            for (int it = 0; it < m; ++it) {
                p[it] = inner(it, gp) + LN;
                base[it] = outs[it] = inner(it, gp);
                kt[it] = inner(it, gp);
            }
            ks_moddown_n(c, level, m, p, base, outs, Tm, s);
            ks_modup_n(c, level, m, kt, D, dw, C, s);
            MacArgs ma = mac_args(c, level, D, k.giant_key[gk], st.montgomery != 0);
            ma.items = m;
            ma.digits_stride = dw;
            ma.qp_out = 1;
            ma.galois = rotation_galois(N, (long long)k.glist[gk] * k.bs);
            ma.accumulate = have_out ? 1 : 0;
            for (int it = 0; it < m; ++it) {
                ma.key[it] = k.giant_key[gk];
                ma.a_out[it] = outp(it);
                ma.b_out[it] = outp(it) + W;
                ma.add_b[it] = inner(it, gp) + W;
            }
            ks_mac(c, ma, level, s);
            have_out = true;
        }

        // result = folded ModDown of out (moddown_qp_pair_to_q_eval, :641-718)
        for (int it = 0; it < m; ++it) {
            p[2 * it] = outp(it) + LN;
            p[2 * it + 1] = outp(it) + W + LN;
            base[2 * it] = outp(it);
            base[2 * it + 1] = outp(it) + W;
            outs[2 * it] = a_out[i0 + it];
            outs[2 * it + 1] = b_out[i0 + it];
        }
        ks_moddown_n(c, level, 2 * m, p, base, outs, Tm, s);
    }
}

void linear_transform_q_batch(Context& c, const LtStage& st, int n, const u64* const* a,
                              const u64* const* b, u64* const* a_out, u64* const* b_out, hipStream_t s) {
    const Compiled k = compile_stage(c, st, true);
    check_batch(n, a, b, a_out, b_out);
    const int level = st.level;
    const int L = level + 1, K = c.np, parts = c.parts(level);
    const u64 N = c.N;
    const long long LN = (long long)L * (long long)N, KN = (long long)K * (long long)N;
    const int nbaby = k.nslots - 1, ngp = k.ngp;
    const long long dw = (long long)parts * (L + K) * (long long)N;

    // per-item scratch: digits | C | P accumulators | ModDown temp | switched
    // pair | babies [slot][item] | out | giant inners [giant][item]
    const long long item_words = dw + LN + 2 * KN + 2 * LN + 2 * LN + 2 * (long long)nbaby * LN + 2 * LN +
                                 2 * (long long)ngp * LN;
    const long long budget_words = (32LL << 30) / 8;
    int chunk = (int)std::min<long long>(KS_MAX_ITEMS, std::max<long long>(1, budget_words / item_words));
    chunk = std::min(chunk, n);
    u64* ws = (u64*)c.workspace(s, sizeof(u64) * (size_t)(item_words * chunk));

    for (int i0 = 0; i0 < n; i0 += chunk) {
        const int m = std::min(chunk, n - i0);
        u64* D = ws;
        u64* C = D + m * dw;
        u64* AP = C + m * LN;
        u64* Tm = AP + m * 2 * KN;
        u64* SW = Tm + m * 2 * LN;
        u64* BQ = SW + m * 2 * LN;
        u64* OQ = BQ + m * 2 * (long long)nbaby * LN;
        u64* IQ = OQ + m * 2 * LN;
        u64* p[2 * KS_MAX_ITEMS];
        const u64* base[2 * KS_MAX_ITEMS];
        u64* outs[2 * KS_MAX_ITEMS];

        // key switch of m key terms into SW (a at +0, b at +LN per item, b
        // starting from b_init), modes Assign: the hoisted or full switch
        // of ckks_hybrid_key_switch_(hoisted_)inplace
        auto mac_moddown = [&](const u64* key, const u64* const* b_init) {
            if (parts <= 3 && c.fuse_mac) {
                // the key product folded into ModDown's NTT passes
                KsItem items[KS_MAX_ITEMS];
                for (int it = 0; it < m; ++it)
                    items[it] = KsItem{nullptr, key, SW + it * 2 * LN + LN, SW + it * 2 * LN, level, 1, 1, b_init[it]};
                ks_moddown_mac(c, level, m, items, D, dw, AP, Tm, st.montgomery != 0, s);
                return;
            }
            MacArgs ma = mac_args(c, level, D, key, st.montgomery != 0);
            ma.items = m;
            ma.digits_stride = dw;
            ma.accp = AP;
            ma.accp_stride = 2 * KN;
            ma.b_mode = 1;
            ma.a_mode = 1;
            for (int it = 0; it < m; ++it) {
                ma.key[it] = key;
                ma.a_out[it] = SW + it * 2 * LN;
                ma.b_out[it] = SW + it * 2 * LN + LN;
                ma.init_b[it] = b_init[it];
            }
            ks_mac(c, ma, level, s);
            for (int it = 0; it < m; ++it) {
                p[2 * it] = AP + it * 2 * KN;
                p[2 * it + 1] = AP + it * 2 * KN + KN;
                base[2 * it] = outs[2 * it] = SW + it * 2 * LN;
                base[2 * it + 1] = outs[2 * it + 1] = SW + it * 2 * LN + LN;
            }
            ks_moddown_n(c, level, 2 * m, p, base, outs, Tm, s);
        };

        // baby steps: hoisted rotations in Q (ckks_hybrid_key_switch_hoisted_inplace
        // with b Add and a Assign, then the automorphism), [slot][item] layout
        if (nbaby > 0) {
            ks_modup_n(c, level, m, a + i0, D, dw, C, s);
            for (int sl = 1; sl < k.nslots; ++sl) {
                mac_moddown(k.baby_key[sl], b + i0);
                automorphism(c, L, k.baby_gal[sl], SW, BQ + (long long)(sl - 1) * m * 2 * LN, 2 * m, LN, LN, s);
            }
        }

        launch_inner(c, k, st, m, 0, a + i0, b + i0, BQ, 2 * LN, (long long)m * 2 * LN, OQ, 2 * LN, IQ,
                     (long long)m * 2 * LN, 2 * LN, s);

        // giant steps > 0: rotate each inner sum by a full key switch
        // (ckks_rotate_inplace: switch a, then the automorphism) and add
        bool have_out = k.g0;
        for (int gp = 0; gp < ngp; ++gp) {
            const size_t gk = (size_t)gp + (k.g0 ? 1 : 0);
            u64* in = IQ + (long long)gp * m * 2 * LN;
            const u64* kt[KS_MAX_ITEMS];
            const u64* bi[KS_MAX_ITEMS];
            for (int it = 0; it < m; ++it) {
                kt[it] = in + it * 2 * LN;
                bi[it] = in + it * 2 * LN + LN;
            }
            ks_modup_n(c, level, m, kt, D, dw, C, s);
            mac_moddown(k.giant_key[gk], bi);
            automorphism_acc(c, L, rotation_galois(N, (long long)k.glist[gk] * k.bs), SW, OQ, 2 * m, LN, LN,
                             have_out, s);
            have_out = true;
        }
        for (int it = 0; it < m; ++it) {
            MI_HIP(hipMemcpyAsync(a_out[i0 + it], OQ + it * 2 * LN, sizeof(u64) * LN, hipMemcpyDeviceToDevice, s));
            MI_HIP(hipMemcpyAsync(b_out[i0 + it], OQ + it * 2 * LN + LN, sizeof(u64) * LN,
                                  hipMemcpyDeviceToDevice, s));
        }
    }
}

void linear_transform_qp(Context& c, const LtStage& st, const u64* a, const u64* b, u64* a_out,
                         u64* b_out, hipStream_t s) {
    linear_transform_qp_batch(c, st, 1, &a, &b, &a_out, &b_out, s);
}

}  // namespace mi
