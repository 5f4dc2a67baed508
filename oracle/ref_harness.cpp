// Golden-vector harness over the reference's own C++ sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) from the
// reference sources where they lie under /root/reference, with output only in
// oracle/_ref/.  It drives the reference's public API (CKKSParams,
// CKKSContext, ckks_* evaluator functions) and writes raw little-endian
// arrays that tests/golden/make_golden.py packs into committed fixtures.
// Inputs the harness chooses are drawn from splitmix64 so the tests can
// regenerate them; the reference's own randomness (keys, encryption masks) is
// captured as data.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <optional>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <sys/stat.h>

#include "ckks_addition.hpp"
#include "ckks_conjugate.hpp"
#include "ckks_context.hpp"
#include "ckks_encoding.hpp"
#include "ckks_keyswitch.hpp"
#include "ckks_mul.hpp"
#include "ckks_params.hpp"
#include "ckks_aes_params.hpp"
#include "ckks_relin.hpp"
#include "ckks_rescale.hpp"
#include "ckks_rotation.hpp"
#include "ckks_linear_transform.hpp"
#include "ckks_modraise.hpp"
#include "ckks_evalpoly.hpp"
#include "ckks_s2c_generalized.hpp"
#include "ckks_c2s_generalized.hpp"
#include "ckks_evalmod_scale_plan.hpp"
#include "ntt.hpp"
#include "random.hpp"

static std::string g_out;

static uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void dump(const std::string& name, const void* data, size_t bytes) {
    std::string path = g_out + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    std::fwrite(data, 1, bytes, f);
    std::fclose(f);
}
template <class V> static void dumpv(const std::string& name, const V& v) {
    dump(name, v.data(), v.size() * sizeof(v[0]));
}
static void dump_ct(const std::string& name, const CKKSCiphertext& ct) {
    std::vector<uint64_t> flat;
    for (size_t k = 0; k < ct.getNumPolys(); ++k) {
        const auto& p = ct.getPoly(k);
        flat.insert(flat.end(), p.begin(), p.end());
    }
    dumpv(name, flat);
    uint64_t meta[3] = {ct.getNumPolys(), ct.getLevel(), 0};
    long double sc = ct.getScale();
    double scd = (double)sc;
    std::memcpy(&meta[2], &scd, 8);
    dump(name + "_meta", meta, sizeof(meta));
}
static void dump_key(const std::string& name, std::span<const CKKSCiphertext> key) {
    std::vector<uint64_t> flat;
    for (const auto& d : key) {
        flat.insert(flat.end(), d.getB().begin(), d.getB().end());
        flat.insert(flat.end(), d.getA().begin(), d.getA().end());
    }
    dumpv(name, flat);
}

// Prime lists of the parameter presets and of the benchmark configurations.
static void run_primes() {
    auto dump_params = [](const std::string& name, const CKKSParams& p) {
        dumpv(name + "_q", p.getModuli());
        dumpv(name + "_p", p.getPModuli());
    };
    dump_params("walsh15", CKKSParams::AESWalsh15());
    dump_params("xboot13", CKKSParams::XBootParamAes13Reference());
    dump_params("xboot14", CKKSParams::AESXBoot14Secure());
    dump_params("xboot15", CKKSParams::AESXBoot15());
    {
        CKKSParamsConfig c;
        c.N = 1 << 14; c.mult_depth = 7; c.q0_bit = 50; c.q_bit = 50;
        c.p_bits = {50, 50}; c.dnum = 4; c.secret_key_hamming_weight = 192;
        c.sigma = 3.2; c.secure = false; c.rotate_values = {1};
        dump_params("chain14", CKKSParams(c));
    }
    {
        std::vector<uint64_t> qb(40, 60);
        auto [p, q] = find_p_and_q_primes({}, qb, 1 << 16);
        dumpv("ntt16_q", q);
        std::vector<uint64_t> qb3(30, 50), pb3(10, 50);
        auto [p3, q3] = find_p_and_q_primes(pb3, qb3, 1 << 16);
        dumpv("ks16_q", q3);
        dumpv("ks16_p", p3);
    }
}

// Twiddle tables and NTT outputs for splitmix64 inputs.
static void run_ntt(size_t N, const std::vector<uint64_t>& bits, const std::string& tag, bool dump_tables) {
    auto [p, q] = find_p_and_q_primes({}, bits, N);
    dumpv(tag + "_q", q);
    for (size_t i = 0; i < q.size(); ++i) {
        auto tw = generate_negacyclic_ntt_twiddles(N, q[i]);
        if (dump_tables) {
            dumpv(tag + "_fwd" + std::to_string(i), tw.forward_ntt);
            dumpv(tag + "_fwds" + std::to_string(i), tw.forward_shoup);
            dumpv(tag + "_inv" + std::to_string(i), tw.inverse_ntt);
            dumpv(tag + "_invs" + std::to_string(i), tw.inverse_shoup);
        }
        uint64_t seed = 1000 + i;
        std::vector<uint64_t> x(N);
        for (auto& v : x) v = splitmix64(seed) % q[i];
        std::vector<uint64_t> y = x;
        ntt_forward_dit2_dispatch(y.data(), N, q[i], tw, false);
        dumpv(tag + "_fwdout" + std::to_string(i), y);
        std::vector<uint64_t> yl = x;
        ntt_forward_dit2_dispatch(yl.data(), N, q[i], tw, true);
        dumpv(tag + "_fwdlazy" + std::to_string(i), yl);
        std::vector<uint64_t> z = x;
        ntt_inverse_dif2_dispatch(z.data(), N, q[i], tw, false);
        dumpv(tag + "_invout" + std::to_string(i), z);
    }
}

// Evaluator operations on a small context with captured keys.
static void run_eval(size_t logN, size_t depth, size_t dnum, std::vector<uint64_t> p_bits,
                     const std::string& tag) {
    CKKSParamsConfig c;
    c.N = size_t{1} << logN;
    c.mult_depth = depth;
    c.q0_bit = 50;
    c.q_bit = 40;
    c.p_bits = p_bits;
    c.dnum = dnum;
    c.secret_key_hamming_weight = 192;
    c.sigma = 3.2;
    c.secure = false;
    c.rotate_values = {1, 5};
    CKKSContext ctx{CKKSParams(c)};
    const auto& P = ctx.getParams();
    const size_t N = P.getN(), L = P.getMaxLevel();
    dumpv(tag + "_q", P.getModuli());
    dumpv(tag + "_p", P.getPModuli());
    std::vector<uint64_t> sk;
    for (const auto& limb : ctx.getSecretKey()) sk.insert(sk.end(), limb.begin(), limb.end());
    dumpv(tag + "_sk", sk);
    dump_key(tag + "_relin", ctx.getRelinHybrid(L));
    dump_key(tag + "_rot1", ctx.getRotationKey(1));
    dump_key(tag + "_rot5", ctx.getRotationKey(5));
    dump_key(tag + "_conj", ctx.getConjugationKey());

    uint64_t seed = 77;
    auto rnd_slots = [&](double amp) {
        MyVector<Complex> v(N / 2);
        for (auto& z : v) {
            double re = (double)(splitmix64(seed) >> 11) * 0x1.0p-53 * 2 - 1;
            double im = (double)(splitmix64(seed) >> 11) * 0x1.0p-53 * 2 - 1;
            z = Complex(re * amp, im * amp);
        }
        return v;
    };
    auto m1 = rnd_slots(1.0), m2 = rnd_slots(1.0);
    dumpv(tag + "_m1", m1);
    dumpv(tag + "_m2", m2);
    auto ct1 = ckks_encrypt(ctx, m1, L);
    auto ct2 = ckks_encrypt(ctx, m2, L);
    dump_ct(tag + "_ct1", ct1);
    dump_ct(tag + "_ct2", ct2);

    {   // encode (eval domain), no randomness
        auto pt = ckks_encode(ctx, m1, L, true);
        dumpv(tag + "_encode_m1", pt.getPlaintext());
    }
    {   // decrypt of ct1 (FP output)
        auto dec = ckks_decrypt(ctx, ct1);
        dumpv(tag + "_dec_ct1", dec);
    }
    {   // tensor
        auto x = ct1.clone();
        ckks_mul_inplace(ctx, x, ct2);
        dump_ct(tag + "_mul", x);
        ckks_relin_hybrid_inplace(ctx, x);
        dump_ct(tag + "_mul_relin", x);
        ckks_rescale(ctx, x);
        dump_ct(tag + "_mul_relin_rescale", x);
        ckks_rotate_inplace(ctx, x, 1);
        dump_ct(tag + "_chain_rot1", x);
    }
    {
        auto x = ct1.clone();
        ckks_square_inplace(ctx, x);
        dump_ct(tag + "_square", x);
    }
    {
        auto x = ct1.clone();
        ckks_rotate_inplace(ctx, x, 1);
        dump_ct(tag + "_rot1_out", x);
        auto y = ct1.clone();
        ckks_rotate_inplace(ctx, y, 5);
        dump_ct(tag + "_rot5_out", y);
        auto z = ct1.clone();
        ckks_conjugate_inplace(ctx, z);
        dump_ct(tag + "_conj_out", z);
    }
    {
        auto x = ct1.clone();
        ckks_rescale(ctx, x);
        dump_ct(tag + "_rescale_ct1", x);
        ckks_rescale(ctx, x);
        dump_ct(tag + "_rescale2_ct1", x);
    }
    {
        auto x = ct1.clone();
        ckks_add_inplace(ctx, x, ct2);
        dump_ct(tag + "_add", x);
        auto y = ct1.clone();
        ckks_sub_inplace(ctx, y, ct2);
        dump_ct(tag + "_sub", y);
        auto z = ct1.clone();
        ckks_mul_inplace(ctx, z, int64_t{-7});
        dump_ct(tag + "_mulint", z);
        auto w = ct1.clone();
        ckks_add_integer_constant_inplace(ctx, w, 3);
        dump_ct(tag + "_addconst", w);
        auto pt = ckks_encode(ctx, m2, L, true);
        dumpv(tag + "_pt_m2", pt.getPlaintext());
        auto v = ct1.clone();
        ckks_mul_inplace(ctx, v, pt);
        dump_ct(tag + "_mulplain", v);
    }
    {   // key switch of a lower-level ciphertext (fewer digits)
        auto x = ct1.clone();
        ckks_drop_to_level_inplace(x, 1);
        ckks_rotate_inplace(ctx, x, 1);
        dump_ct(tag + "_rot1_level1", x);
    }
}

// splitmix64 residues of one QP plaintext of a linear-transform stage:
// limb l of term t is drawn from seed 5000 + 64 t + l.
static MyVector<uint64_t> lt_diag_qp(size_t N, const MyVector<uint64_t>& qp, size_t term) {
    MyVector<uint64_t> out(N * qp.size());
    for (size_t l = 0; l < qp.size(); ++l) {
        uint64_t seed = 5000 + 64 * term + l;
        for (size_t i = 0; i < N; ++i) out[l * N + i] = splitmix64(seed) % qp[l];
    }
    return out;
}

// Hoisted rotations and a QP-evaluator BSGS stage on a small context.
static void run_hoist_lt() {
    CKKSParamsConfig c;
    c.N = size_t{1} << 10;
    c.mult_depth = 3;
    c.q0_bit = 50;
    c.q_bit = 40;
    c.p_bits = {60, 60};
    c.dnum = 2;
    c.secret_key_hamming_weight = 192;
    c.sigma = 3.2;
    c.secure = false;
    c.rotate_values = {1, 2, 4};
    CKKSContext ctx{CKKSParams(c)};
    const auto& P = ctx.getParams();
    const size_t N = P.getN(), L = P.getMaxLevel();
    const std::string tag = "lt10";
    dumpv(tag + "_q", P.getModuli());
    dumpv(tag + "_p", P.getPModuli());
    dump_key(tag + "_rot1", ctx.getRotationKey(1));
    dump_key(tag + "_rot2", ctx.getRotationKey(2));
    dump_key(tag + "_rot4", ctx.getRotationKey(4));
    uint64_t seed = 91;
    MyVector<Complex> m(N / 2);
    for (auto& z : m) {
        double re = (double)(splitmix64(seed) >> 11) * 0x1.0p-53 * 2 - 1;
        z = Complex(re, -re / 2);
    }
    auto ct = ckks_encrypt(ctx, m, L);
    dump_ct(tag + "_ct", ct);

    // hoisted key switches of ct's a (rotation by 1, Q result and QP result)
    CKKSHoistWorkspace ws;
    CKKSEvalHoistedCiphertextQP h;
    ckks_hoist_ciphertext_a_qp_eval(ctx, ct, ws, h);
    {
        std::vector<uint64_t> flat;
        for (const auto& d : h.a_partition_qp_eval) flat.insert(flat.end(), d.begin(), d.end());
        dumpv(tag + "_hoist_digits_lazy", flat);
        MyVector<uint64_t> b = ct.getB();
        MyVector<uint64_t> a(N * (L + 1), 0);
        ckks_hybrid_key_switch_hoisted_inplace(ctx, h, ctx.getRotationKey(1), b, a,
                                               CKKSHybridKeySwitchCombine::Add,
                                               CKKSHybridKeySwitchCombine::Assign, P.montgomery,
                                               ctx.getRotationNttMap(1));
        std::vector<uint64_t> o(b.begin(), b.end());
        o.insert(o.end(), a.begin(), a.end());
        dumpv(tag + "_hoisted_rot1", o);
        MyVector<uint64_t> bqp, aqp;
        ckks_hybrid_key_switch_hoisted_qp_rotated(ctx, h, ctx.getRotationKey(2), ct.getB(), bqp, aqp,
                                                  P.montgomery, ctx.getRotationNttMap(2));
        std::vector<uint64_t> o2(bqp.begin(), bqp.end());
        o2.insert(o2.end(), aqp.begin(), aqp.end());
        dumpv(tag + "_hoisted_qp_rot2", o2);
        MyVector<uint64_t> sb, sa;
        ckks_hybrid_key_switch_qp_from_key_term(ctx, L, ct.getA(), ctx.getRotationKey(4), sb, sa,
                                                P.montgomery);
        std::vector<uint64_t> o3(sb.begin(), sb.end());
        o3.insert(o3.end(), sa.begin(), sa.end());
        dumpv(tag + "_qp_from_key_term_rot4", o3);
    }

    // BSGS stage: diagonals 0..5, baby steps 2 (rotations 1; giants 2, 4)
    MyVector<uint64_t> qp(P.getModuli().begin(), P.getModuli().end());
    qp.insert(qp.end(), P.getPModuli().begin(), P.getPModuli().end());
    for (int rescale = 0; rescale < 2; ++rescale) {
        CKKSLinearTransformPlan plan;
        plan.use_qp_evaluator = true;
        plan.stages.emplace_back();
        auto& st = plan.stages.back();
        for (size_t t = 0; t < 6; ++t) {
            MyVector<Complex> d(N / 2, Complex(0.5, 0.25));
            st.diagonals.emplace_back(ckks_encode(ctx, d, L, true));
            st.diagonal_qp_plaintexts.emplace_back(lt_diag_qp(N, qp, t));
            st.diagonal_indices.emplace_back(t);
        }
        st.baby_step_count = 2;
        st.qp_limb_count = qp.size();
        st.qp_p_limb_count = P.getPModuli().size();
        st.diagonal_qp_plaintexts_montgomery = true;
        st.rescale_after = rescale != 0;
        auto out = ckks_apply_linear_transform_plan(ctx, ct.clone(), plan);
        dump_ct(tag + (rescale ? "_lt_rescaled" : "_lt"), out);
    }
}

// Seeded residues of one Q-basis diagonal plaintext (level+1 limbs).
static MyVector<uint64_t> lt_diag_q(size_t N, const MyVector<uint64_t>& q, size_t limbs, size_t term) {
    MyVector<uint64_t> out(N * limbs);
    for (size_t l = 0; l < limbs; ++l) {
        uint64_t seed = 7000 + 64 * term + l;
        for (size_t i = 0; i < N; ++i) out[l * N + i] = splitmix64(seed) % q[l];
    }
    return out;
}

// Q-evaluator BSGS stages: an owned stage through a plan with the Q
// evaluator (Montgomery-form diagonals, no rescale) and
// ckks_apply_linear_transform_bsgs (normal-form diagonals, rescale after).
static void run_lt_q() {
    CKKSParamsConfig c;
    c.N = size_t{1} << 10;
    c.mult_depth = 3;
    c.q0_bit = 50;
    c.q_bit = 40;
    c.p_bits = {60, 60};
    c.dnum = 2;
    c.secret_key_hamming_weight = 192;
    c.sigma = 3.2;
    c.secure = false;
    c.rotate_values = {1, 2, 3, 6};
    CKKSContext ctx{CKKSParams(c)};
    const auto& P = ctx.getParams();
    const size_t N = P.getN(), L = P.getMaxLevel();
    const std::string tag = "lq10";
    dumpv(tag + "_q", P.getModuli());
    dumpv(tag + "_p", P.getPModuli());
    for (int r : {1, 2, 3, 6}) dump_key(tag + "_rot" + std::to_string(r), ctx.getRotationKey(r));
    uint64_t seed = 93;
    MyVector<Complex> m(N / 2);
    for (auto& z : m) {
        double re = (double)(splitmix64(seed) >> 11) * 0x1.0p-53 * 2 - 1;
        z = Complex(re, re / 3);
    }
    auto ct = ckks_encrypt(ctx, m, L);
    dump_ct(tag + "_ct", ct);
    const long double pt_scale = 1099511627776.0L;  // 2^40
    {
        // diagonals {0, 1, 2, 4, 5, 7}, baby steps 3: giants 0, 3, 6
        CKKSLinearTransformPlan plan;
        plan.use_qp_evaluator = false;
        plan.stages.emplace_back();
        auto& st = plan.stages.back();
        const size_t idx[6] = {0, 1, 2, 4, 5, 7};
        for (size_t t = 0; t < 6; ++t) {
            st.diagonals.emplace_back(N, lt_diag_q(N, P.getModuli(), L + 1, t), pt_scale, L, true);
            st.diagonal_indices.emplace_back(idx[t]);
        }
        st.baby_step_count = 3;
        st.rescale_after = false;
        auto out = ckks_apply_linear_transform_plan(ctx, ct.clone(), plan);
        dump_ct(tag + "_ltq", out);
    }
    {
        // diagonals {1, 3, 5, 6} in normal form through the stage view API
        MyVector<CKKSEncoding> diags;
        const size_t idx[4] = {1, 3, 5, 6};
        for (size_t t = 0; t < 4; ++t)
            diags.emplace_back(N, lt_diag_q(N, P.getModuli(), L + 1, 10 + t), pt_scale, L, false);
        MyVector<size_t> indices(idx, idx + 4);
        CKKSLinearTransformStage st;
        st.diagonals = std::span<const CKKSEncoding>(diags.data(), diags.size());
        st.diagonal_indices = std::span<const size_t>(indices.data(), indices.size());
        st.baby_step_count = 3;
        st.rescale_after = true;
        auto out = ckks_apply_linear_transform_bsgs(ctx, ct, st);
        dump_ct(tag + "_bsgs", out);
    }
}

// Pieces of bootstrapping on a context with bootstrap keys and EvalMod depth
// 0 (no EvalMod polynomial, so nothing reaches the MPFR fit).
static void run_bootstrap_pieces() {
    CKKSParamsConfig c;
    c.N = size_t{1} << 10;
    c.mult_depth = 1;
    c.q0_bit = 40;
    c.q_bit = 35;
    c.p_bits = {60, 60};
    c.dnum = 2;
    c.secret_key_hamming_weight = 192;
    c.sigma = 3.2;
    c.secure = false;
    c.bts_keys = true;
    auto bp = ckks_params_detail::make_evalmod_bootstrap_params(
        16, {CKKSBootstrapDftScaleSchedule{{25}, {25}, {25}}, CKKSBootstrapDftScaleSchedule{{25}, {25}, {25}}},
        0, 0, {0, 0});
    bp.linear_transform_use_qp = {true, true};
    c.bootstrap_params = bp;
    CKKSContext ctx{CKKSParams(c)};
    const auto& P = ctx.getParams();
    const size_t N = P.getN();
    const std::string tag = "bt10";
    dumpv(tag + "_q", P.getModuli());
    dumpv(tag + "_p", P.getPModuli());
    dump_key(tag + "_d2s", ctx.getDenseToSparseSwitchKey());
    uint64_t seed = 123;
    MyVector<Complex> m(N / 2);
    for (auto& z : m) {
        double re = (double)(splitmix64(seed) >> 11) * 0x1.0p-53 * 2 - 1;
        z = Complex(re * 0.01, 0.0);
    }
    auto ct = ckks_encrypt(ctx, m, 0);
    dump_ct(tag + "_ct0", ct);
    ckks_key_switch_dense_to_sparse_inplace(ctx, ct);
    dump_ct(tag + "_q0p0", ct);
    ckks_modraise_inplace(ctx, ct, P.getMaxLevel());
    dump_ct(tag + "_modraise", ct);
    // the S2C DFT plan's QP diagonals (digests pin host plan encoding later)
    const auto& plan = ctx.getEvalModS2CPlan();
    std::vector<uint64_t> meta;
    for (const auto& st : plan.stages) {
        meta.push_back(st.diagonal_indices.size());
        meta.push_back(st.baby_step_count);
        meta.push_back(st.qp_limb_count);
        meta.insert(meta.end(), st.diagonal_indices.begin(), st.diagonal_indices.end());
        std::vector<uint64_t> flat;
        for (const auto& d : st.diagonal_qp_plaintexts) flat.insert(flat.end(), d.begin(), d.end());
        dumpv(tag + "_s2c_stage" + std::to_string(&st - plan.stages.data()), flat);
    }
    dumpv(tag + "_s2c_meta", meta);
}

// The scaled C2S plan of a context's stages: metadata and QP diagonals.
static void dump_plan(const std::string& tag, const CKKSLinearTransformPlan& plan) {
    std::vector<uint64_t> meta;
    for (const auto& st : plan.stages) {
        meta.push_back(st.diagonal_indices.size());
        meta.push_back(st.baby_step_count);
        meta.push_back(st.qp_limb_count);
        meta.insert(meta.end(), st.diagonal_indices.begin(), st.diagonal_indices.end());
        std::vector<uint64_t> flat;
        for (const auto& d : st.diagonal_qp_plaintexts) flat.insert(flat.end(), d.begin(), d.end());
        dumpv(tag + "_stage" + std::to_string(&st - plan.stages.data()), flat);
    }
    dumpv(tag + "_meta", meta);
}

// The bootstrap driver ckks_batch_bootstrap_evalmod_from_complex_packed
// (ckks_bootstrap_evalmod_inplace.cpp:354-434) as its sequence of public
// functions, on a bootstrapping context whose EvalMod polynomial is a fixed
// harness-chosen one: the context is built without EvalMod limbs (the
// polynomial fit needs MPFR), the mult limbs carry the Chebyshev evaluation,
// and evalmod_binary_inplace (:306-326) is restated with those coefficients.
// Dumps the keys the sequence uses, its input, each step's output and the
// S2C / scaled C2S plans.
static void run_bootstrap_sequence() {
    CKKSParamsConfig c;
    c.N = size_t{1} << 10;
    c.mult_depth = 3;
    c.q0_bit = 40;
    c.q_bit = 35;
    c.p_bits = {60, 60, 60};
    c.dnum = 2;
    c.secret_key_hamming_weight = 192;
    c.sigma = 3.2;
    c.secure = false;
    c.bts_keys = true;
    auto bp = ckks_params_detail::make_evalmod_bootstrap_params(
        16, {CKKSBootstrapDftScaleSchedule{{25}, {25}, {25}}, CKKSBootstrapDftScaleSchedule{{25}, {25}, {25}}},
        0, 0, {0, 0});
    bp.linear_transform_use_qp = {true, true};
    c.bootstrap_params = bp;
    CKKSContext ctx{CKKSParams(c)};
    const auto& P = ctx.getParams();
    const size_t N = P.getN(), L = P.getMaxLevel();
    const std::string tag = "bs10";
    dumpv(tag + "_q", P.getModuli());
    dumpv(tag + "_p", P.getPModuli());
    dump_key(tag + "_relin", ctx.getRelinHybrid(L));
    dump_key(tag + "_conj", ctx.getConjugationKey());
    dump_key(tag + "_d2s", ctx.getDenseToSparseSwitchKey());
    dump_key(tag + "_s2d", ctx.getSparseToDenseSwitchKey());
    // the rotation keys of both plans' baby and giant steps
    std::vector<uint64_t> rots;
    for (const CKKSLinearTransformPlan* plan : {&ctx.getEvalModS2CPlan(), &ctx.getEvalModScaledC2SPlan()})
        for (const auto& st : plan->stages)
            for (size_t d : st.diagonal_indices)
                for (size_t r : {d % st.baby_step_count, (d / st.baby_step_count) * st.baby_step_count})
                    if (r != 0 && std::find(rots.begin(), rots.end(), (uint64_t)r) == rots.end()) rots.push_back(r);
    std::sort(rots.begin(), rots.end());
    dumpv(tag + "_rots", rots);
    for (uint64_t r : rots) dump_key(tag + "_rot" + std::to_string(r), ctx.getRotationKey((int64_t)r));
    dump_plan(tag + "_s2c", ctx.getEvalModS2CPlan());
    dump_plan(tag + "_c2s", ctx.getEvalModScaledC2SPlan());
    double scales[2] = {ctx.getEvalModScaledC2STransformScale(), std::exp2((double)ckks_evalmod_log_scale_bits(P))};
    dump(tag + "_scales", scales, sizeof(scales));
    // complex-packed bits at the S2C level and the default scale (input scale ratio 1)
    uint64_t seed = 777;
    MyVector<Complex> m(N / 2);
    for (auto& z : m)
        z = Complex((double)(splitmix64(seed) & 1), (double)(splitmix64(seed) & 1));
    dumpv(tag + "_m", m);
    auto packed = ckks_encrypt(ctx, m, P.getBootstrapS2CDepth());
    dump_ct(tag + "_in", packed);
    // evalmod_s2c_for_input_scale (:156-178) with scale ratio 1
    auto s2c = ckks_s2c_generalized(ctx, packed, P.getBootstrapS2CDepth());
    s2c.setScale(P.getScale());
    dump_ct(tag + "_s2c_out", s2c);
    // finish_batch_evalmod_from_s2c_coeff_packed (:354-389)
    ckks_key_switch_dense_to_sparse_inplace(ctx, s2c);
    dump_ct(tag + "_d2s_out", s2c);
    ckks_modraise_inplace(ctx, s2c, P.getDepth());
    dump_ct(tag + "_modraise_out", s2c);
    ckks_key_switch_sparse_to_dense_inplace(ctx, s2c);
    dump_ct(tag + "_s2d_out", s2c);
    auto c2s = ckks_c2s_generalized(ctx, s2c, ctx.getEvalModScaledC2STransformScale());
    dump_ct(tag + "_c2s_out", c2s);
    auto conjugated = ckks_conjugate(ctx, c2s);
    auto out0 = c2s.clone();
    ckks_add_inplace(ctx, out0, conjugated);
    out0.setScale(c2s.getScale());
    auto out1 = c2s.clone();
    ckks_sub_inplace(ctx, out1, conjugated);
    out1 = ckks_mul(ctx, out1, Complex(0.0, -1.0), 1.0);
    out1.setScale(c2s.getScale());
    dump_ct(tag + "_split0", out0);
    dump_ct(tag + "_split1", out1);
    // an even degree-6 polynomial in place of the fitted one (three levels)
    std::vector<double> co(7, 0.0);
    for (size_t i = 0; i <= 6; i += 2) co[i] = ((double)(splitmix64(seed) >> 11) * 0x1.0p-53 - 0.5) / (double)(i + 1);
    dumpv(tag + "_coeffs", co);
    int k = 0;
    for (CKKSCiphertext* ct : {&out0, &out1}) {
        CKKSEvalPolyOptions opt;
        opt.target_scale = P.getScale();
        ct->setScale(std::exp2((double)ckks_evalmod_log_scale_bits(P)));
        *ct = ckks_evalchebyshev_preconditioned(ctx, *ct, co, opt);
        ct->setScale(P.getScale());
        dump_ct(tag + "_out" + std::to_string(k++), *ct);
    }
}

// Chebyshev evaluation (Paterson-Stockmeyer with exact-scale tracking) of
// harness-chosen coefficients on an encrypted input in [-1, 1].
static void run_evalpoly() {
    CKKSParamsConfig c;
    c.N = size_t{1} << 10;
    c.mult_depth = 8;
    c.q0_bit = 50;
    c.q_bit = 40;
    c.p_bits = {60, 60, 60};
    c.dnum = 3;
    c.secret_key_hamming_weight = 192;
    c.sigma = 3.2;
    c.secure = false;
    CKKSContext ctx{CKKSParams(c)};
    const auto& P = ctx.getParams();
    const size_t N = P.getN(), L = P.getMaxLevel();
    const std::string tag = "ep10";
    dumpv(tag + "_q", P.getModuli());
    dumpv(tag + "_p", P.getPModuli());
    std::vector<uint64_t> sk;
    for (const auto& limb : ctx.getSecretKey()) sk.insert(sk.end(), limb.begin(), limb.end());
    dumpv(tag + "_sk", sk);
    dump_key(tag + "_relin", ctx.getRelinHybrid(L));
    uint64_t seed = 4242;
    MyVector<Complex> m(N / 2);
    for (auto& z : m) z = Complex((double)(splitmix64(seed) >> 11) * 0x1.0p-53 * 2 - 1, 0.0);
    dumpv(tag + "_m", m);
    auto ct = ckks_encrypt(ctx, m, L);
    dump_ct(tag + "_ct", ct);
    // coefficient sets: generic degree 31, even degree 30, odd degree 15,
    // generic degree 63, degree 2, degree 1
    const size_t degrees[6] = {31, 30, 15, 63, 2, 1};
    const int parity[6] = {0, 2, 1, 0, 0, 0};  // 0 any, 1 odd, 2 even
    for (int k = 0; k < 6; ++k) {
        std::vector<double> co(degrees[k] + 1, 0.0);
        for (size_t i = 0; i <= degrees[k]; ++i) {
            const double u = (double)(splitmix64(seed) >> 11) * 0x1.0p-53 - 0.5;
            if ((parity[k] == 1 && i % 2 == 0) || (parity[k] == 2 && i % 2 == 1)) continue;
            co[i] = u / (double)(i + 1);
        }
        dumpv(tag + "_coeffs" + std::to_string(k), co);
        auto out = ckks_evalchebyshev_preconditioned(ctx, ct, co);
        dump_ct(tag + "_out" + std::to_string(k), out);
    }
}

// ---------------------------------------------------------------------------
// CPU timing of one walsh15 S-box (ckks_walsh_sbox_complex_blocks_mask_packed,
// ckks_walsh_sbox.cpp) through the reference's own evaluator on the host's
// OpenMP threads: the CPU baseline of the AES evaluation (160 S-boxes and 4
// output refreshes per 1024 blocks).  The reference fits its EvalMod
// polynomial with MPFR (poly_approx.cpp), which this image lacks, so the
// context is AESWalsh15's with the seven 38-bit EvalMod limbs declared as
// mult limbs: the modulus chain, keys, plans and every operation's shape are
// those of AESWalsh15 (checked below), and the degree-100 polynomial (its
// coefficients read from a file) is evaluated on those limbs with
// ckks_evalchebyshev_preconditioned as evalmod_binary_inplace does.  The
// S-box's private steps (parity packing, hoisted unpacking, the Walsh product
// sums) are composed from the same public calls they make.
#include <chrono>
#include <fstream>
#include <omp.h>

static double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Eight encrypted bit ciphertexts at the S2C entry level (512 blocks per lane).
static MyVector<CKKSCiphertext> sbox_input(const CKKSContext& ctx, uint64_t seed) {
    const auto& P = ctx.getParams();
    MyVector<CKKSCiphertext> bits;
    for (int k = 0; k < 8; ++k) {
        MyVector<Complex> m(P.getSlots(), Complex(0.0, 0.0));
        for (size_t j = 0; j < 512; ++j)
            m[j] = Complex((double)(splitmix64(seed) & 1), (double)(splitmix64(seed) & 1));
        bits.push_back(ckks_encrypt(ctx, m, P.getBootstrapS2CDepth()));
    }
    return bits;
}

// One S-box; t[0..5] = the seconds of its steps.
static void time_one_sbox(const CKKSContext& ctx, const std::vector<double>& coeffs,
                          const MyVector<CKKSCiphertext>& bits, double* t) {
    const auto& P = ctx.getParams();
    const size_t slots = P.getSlots(), bc = 512, level = P.getBootstrapS2CDepth();
    auto t0 = std::chrono::steady_clock::now();
    // parity packing: 30 parity sums, each rotated to its segment and added
    std::optional<CKKSCiphertext> packed;
    for (size_t half = 0; half < 2; ++half)
        for (size_t mask = 1; mask < 16; ++mask) {
            std::optional<CKKSCiphertext> sum;
            for (size_t b = 0; b < 4; ++b)
                if ((mask >> b) & 1) {
                    if (!sum) sum.emplace(bits[4 * half + b].clone());
                    else ckks_add_inplace(ctx, *sum, bits[4 * half + b]);
                }
            const size_t segment = half * 15 + mask - 1;
            if (segment) ckks_rotate_inplace(ctx, *sum, (int64_t)(segment * bc));
            if (!packed) packed.emplace(std::move(*sum));
            else ckks_add_inplace(ctx, *packed, *sum);
        }
    t[0] = seconds_since(t0);
    // the complex-packed bootstrap
    t0 = std::chrono::steady_clock::now();
    auto s2c = ckks_s2c_generalized(ctx, *packed, level);
    s2c.setScale(P.getScale());
    ckks_key_switch_dense_to_sparse_inplace(ctx, s2c);
    ckks_modraise_inplace(ctx, s2c, P.getDepth());
    ckks_key_switch_sparse_to_dense_inplace(ctx, s2c);
    auto c2s = ckks_c2s_generalized(ctx, s2c, ctx.getEvalModScaledC2STransformScale());
    auto conjugated = ckks_conjugate(ctx, c2s);
    CKKSCiphertext lanes[2] = {c2s.clone(), c2s.clone()};
    ckks_add_inplace(ctx, lanes[0], conjugated);
    ckks_sub_inplace(ctx, lanes[1], conjugated);
    lanes[1] = ckks_mul(ctx, lanes[1], Complex(0.0, -1.0), 1.0);
    t[1] = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    for (auto& ct : lanes) {
        CKKSEvalPolyOptions opt;
        opt.target_scale = P.getScale();
        ct.setScale(std::exp2((double)ckks_evalmod_log_scale_bits(P)));
        ct = ckks_evalchebyshev_preconditioned(ctx, ct, coeffs, opt);
        ct.setScale(P.getScale());
        ckks_mul_inplace(ctx, ct, int64_t{-2});  // bit -> sign
        ckks_add_integer_constant_inplace(ctx, ct, int64_t{1});
    }
    t[2] = seconds_since(t0);
    std::vector<CKKSCiphertext> outputs;
    for (auto& lane : lanes) {
        // hoisted unpacking of the 30 segments
        t0 = std::chrono::steady_clock::now();
        CKKSHoistWorkspace ws;
        CKKSEvalHoistedCiphertextQP hoisted;
        ckks_hoist_ciphertext_a_qp_eval(ctx, lane, ws, hoisted);
        std::vector<CKKSCiphertext> signs;  // [0..14] low masks 1..15, [15..29] high
        const size_t N = P.getN(), L = lane.getLevel() + 1;
        for (size_t segment = 0; segment < 30; ++segment) {
            if (segment == 0) { signs.push_back(lane.clone()); continue; }
            const int64_t r = ((-(int64_t)(segment * bc)) % (int64_t)slots + (int64_t)slots) % (int64_t)slots;
            MyVector<uint64_t> b = lane.getB();
            MyVector<uint64_t> a(N * L, 0);
            ckks_hybrid_key_switch_hoisted_inplace(ctx, hoisted, ctx.getRotationKey(r), b, a,
                                                   CKKSHybridKeySwitchCombine::Add,
                                                   CKKSHybridKeySwitchCombine::Assign, P.montgomery,
                                                   ctx.getRotationNttMap(r));
            signs.emplace_back(N, std::move(a), std::move(b), lane.getScale(), lane.getLevel(),
                               lane.getSecretOwner(), lane.getMessageEncodingState());
        }
        t[3] += seconds_since(t0);
        // four output pairs: per high mask, the low signs' integer
        // combination times that high sign, summed raw, one relin, rescale
        t0 = std::chrono::steady_clock::now();
        for (int pair = 0; pair < 4; ++pair) {
            std::optional<CKKSCiphertext> raw;
            for (size_t h = 0; h < 15; ++h) {
                std::optional<CKKSCiphertext> comb;
                for (size_t m = 0; m < 15; ++m) {
                    auto term = signs[m].clone();
                    ckks_mul_inplace(ctx, term, (int64_t)(((h * 7 + m * 3 + pair) % 5) * 2 - 4));
                    if (!comb) comb.emplace(std::move(term));
                    else ckks_add_inplace(ctx, *comb, term);
                }
                ckks_mul_inplace(ctx, *comb, signs[15 + h]);
                if (!raw) raw.emplace(std::move(*comb));
                else ckks_add_inplace(ctx, *raw, *comb);
            }
            ckks_relin_hybrid_inplace(ctx, *raw);
            ckks_rescale(ctx, *raw);
            outputs.push_back(std::move(*raw));
        }
        t[4] += seconds_since(t0);
    }
    // real/imaginary split of the eight outputs (one conjugation each) and recombination
    t0 = std::chrono::steady_clock::now();
    for (auto& o : outputs) {
        auto conj = ckks_conjugate(ctx, o);
        auto re = o.clone();
        ckks_add_inplace(ctx, re, conj);
        auto im = o.clone();
        ckks_sub_inplace(ctx, im, conj);
        im = ckks_mul(ctx, im, Complex(0.0, -1.0), 1.0);
        im.setScale(re.getScale());
        ckks_add_inplace(ctx, re, im);
    }
    t[5] = seconds_since(t0);
}

static void run_sbox15_timing(const char* coeff_path, int concurrent) {
    std::vector<double> coeffs;
    {
        std::ifstream f(coeff_path, std::ios::binary);
        double v;
        while (f.read(reinterpret_cast<char*>(&v), sizeof v)) coeffs.push_back(v);
    }
    if (coeffs.size() < 2) { std::fprintf(stderr, "sbox15time: no coefficients in %s\n", coeff_path); std::exit(2); }
    const auto t_setup = std::chrono::steady_clock::now();
    CKKSParamsConfig c = ckks_aes_params_detail::AESWalsh15Config();
    const CKKSParams original(c);
    auto bp = *c.bootstrap_params;
    const size_t evalmod_limbs = bp.evalmod_depth;
    std::vector<uint64_t> mult_bits = c.q_mult_bits;
    for (size_t i = 0; i < evalmod_limbs; ++i) mult_bits.push_back(bp.evalmod_bit_size);
    bp.evalmod_depth = 0;
    c.bootstrap_params = bp;
    c.q_mult_bits = mult_bits;
    c.mult_depth = mult_bits.size();
    CKKSContext ctx{CKKSParams(c)};
    const auto& P = ctx.getParams();
    if (P.getModuli() != original.getModuli() || P.getPModuli() != original.getPModuli()) {
        std::fprintf(stderr, "sbox15time: the modulus chain differs from AESWalsh15's\n");
        std::exit(3);
    }
    const double setup_s = seconds_since(t_setup);
    // as sub_bytes_walsh_complex_blocks_mask_packed runs them (:659-721): the
    // S-boxes of a round concurrently, one per OpenMP thread (nested regions
    // inside each run on that thread alone)
    omp_set_max_active_levels(1);
    std::vector<std::array<double, 6>> steps(concurrent);
    std::vector<MyVector<CKKSCiphertext>> inputs;
    for (int i = 0; i < concurrent; ++i) inputs.push_back(sbox_input(ctx, 99 + (uint64_t)i));
    const auto t_all = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < concurrent; ++i) {
        double t[6] = {0, 0, 0, 0, 0, 0};
        time_one_sbox(ctx, coeffs, inputs[i], t);
        for (int k = 0; k < 6; ++k) steps[i][k] = t[k];
    }
    const double wall = seconds_since(t_all);
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (const auto& st : steps)
        for (int k = 0; k < 6; ++k) t[k] += st[k] / concurrent;
    std::printf("{\"concurrent_sboxes\": %d, \"wall_s\": %.6f, \"sbox_mean_s\": %.6f, \"parity_pack_s\": %.6f, "
                "\"bootstrap_linear_s\": %.6f, \"evalmod_s\": %.6f, \"unpack_s\": %.6f, "
                "\"walsh_products_s\": %.6f, \"split_s\": %.6f, \"setup_s\": %.3f, \"omp_threads\": %d, "
                "\"limbs\": %zu}\n",
                concurrent, wall, t[0] + t[1] + t[2] + t[3] + t[4] + t[5], t[0], t[1], t[2], t[3], t[4], t[5],
                setup_s, omp_get_max_threads(), P.getModuli().size());
    std::fflush(stdout);
}

int main(int argc, char** argv) {
    if (argc == 4 && std::strcmp(argv[1], "sbox15time") == 0) {
        run_sbox15_timing(argv[2], std::max(1, std::atoi(argv[3])));
        return 0;
    }
    if (argc < 2) { std::fprintf(stderr, "usage: ref_harness OUTDIR [case ...]\n"); return 2; }
    g_out = argv[1];
    mkdir(g_out.c_str(), 0755);
    // cases: primes ntt eval lt ltq boot evalpoly (all when none is named)
    auto want = [&](const char* name) {
        if (argc < 3) return true;
        for (int i = 2; i < argc; ++i)
            if (std::strcmp(argv[i], name) == 0) return true;
        return false;
    };
    if (want("primes")) run_primes();
    if (want("ntt")) {
        run_ntt(1 << 10, {60, 59, 50, 36}, "ntt10", true);
        run_ntt(1 << 12, {61, 55, 45, 30}, "ntt12", false);
        run_ntt(1 << 16, {60}, "ntt16one", false);
    }
    if (want("eval")) {
        run_eval(10, 3, 2, {60, 60}, "ev10");
        run_eval(10, 4, 3, {61, 61}, "ev10b");
    }
    if (want("lt")) run_hoist_lt();
    if (want("ltq")) run_lt_q();
    if (want("boot")) run_bootstrap_pieces();
    if (want("evalpoly")) run_evalpoly();
    if (want("bootseq")) run_bootstrap_sequence();
    std::printf("ref_harness: wrote %s\n", g_out.c_str());
    return 0;
}
