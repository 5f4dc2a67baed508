// Bootstrapping driver on the host API (ckks_bootstrap_evalmod_inplace.cpp:
// 156-434, ckks_s2c_generalized.cpp, ckks_c2s_generalized.cpp,
// ckks_modraise.cpp:76-110, ckks_keyswitch.cpp:1844-1870).
//
//   S2C (slots -> coefficients, scaled for the input scale)
//   dense -> sparse switch at level 0 over {q0, p0}
//   ModRaise to the top of the chain
//   sparse -> dense switch
//   trace over the inactive slots (sparse packing only)
//   C2S (coefficients -> slots)
//   split into the real and imaginary halves (conjugate, add / subtract, * -i)
//   EvalMod on each half (Chebyshev polynomial, then the double-angle steps)
// Every ring step is a C-ABI kernel; this file only carries the schedule,
// scales and metadata.
#include <cmath>
#include <limits>

#include "ckks.hpp"

namespace {

uint64_t scaled_real_mod_q(double value, long double scale, uint64_t q) {
    const long double scaled = (long double)value * scale;
    if (!std::isfinite(scaled)) throw std::invalid_argument("evalmod_binary_inplace: non-finite double-angle scalar");
    long double r = std::fmod(std::round(scaled), (long double)q);
    if (r < 0.0L) r += (long double)q;
    uint64_t o = (uint64_t)r;
    if (o >= q) o %= q;
    return o;
}

// ct <- ct * scalar with the result's scale set to output_scale
// (multiply_by_real_scalar_to_scale_inplace, ckks_bootstrap_evalmod_inplace.cpp:228-270)
void mul_real_to_scale(const CKKSContext& c, CKKSCiphertext& ct, double scalar, long double output_scale) {
    if (!std::isfinite(scalar) || !std::isfinite(output_scale) || output_scale <= 0.0L || ct.getScale() <= 0.0)
        throw std::invalid_argument("evalmod_binary_inplace: invalid double-angle scalar scaling");
    const auto& q = c.getParams().getModuli();
    const size_t lv = ct.getLevel();
    const long double ss = output_scale / ct.getScale();
    std::vector<uint64_t> r(lv + 1);
    for (size_t l = 0; l <= lv; ++l) r[l] = scaled_real_mod_q(scalar, ss, q[l]);
    ckks_detail::poly_list_op(c, MI_LIST_MUL_SCALAR, ct, nullptr, r.data(), "evalmod_binary_inplace");
    ct.setScale((double)output_scale);
}

void double_angle(const CKKSContext& c, CKKSCiphertext& ct, size_t steps) {
    const long double target = ct.getScale();
    for (size_t i = 0; i < steps; ++i) {
        if (ct.getLevel() == 0)
            throw std::invalid_argument("evalmod_binary_inplace: insufficient level for BinBoot double-angle");
        // the square of ct written into a new ciphertext (the same product as
        // squaring a copy of ct in place)
        CKKSCiphertext sq = ckks_detail::mul_into_at_level(c, ct, nullptr, ct.getLevel());
        ckks_relin_hybrid_inplace(c, sq);
        ckks_rescale(c, sq);
        mul_real_to_scale(c, sq, -4.0, target);
        ckks_drop_to_level_inplace(ct, sq.getLevel());
        mul_real_to_scale(c, ct, 4.0, target);
        ckks_add_inplace(c, ct, sq);
        ct.setScale(target);
    }
}

void validate_packed_input(const CKKSContext& c, const CKKSCiphertext& ct, const char* label) {
    const auto& P = c.getParams();
    const std::string l(label);
    if (P.getBootstrapS2CDepth() == 0) throw std::invalid_argument(l + ": S2C depth must be positive");
    if (ct.getN() != P.getN()) throw std::invalid_argument(l + ": ciphertext/context size mismatch");
    if (ct.getNumPolys() != 2) throw std::invalid_argument(l + ": expected a 2-polynomial ciphertext");
    if (ct.getLevel() != P.getBootstrapS2CDepth())
        throw std::invalid_argument(l + ": expected ciphertext level to equal S2C depth");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense) throw std::invalid_argument(l + ": expected a dense-secret ciphertext");
    if (ct.getMessageEncodingState() != CKKSMessageEncodingState::Slots)
        throw std::invalid_argument(l + ": expected slot-encoded input");
}

uint64_t input_scale_ratio(const CKKSContext& c, const CKKSCiphertext& ct, const char* label) {
    const long double target = c.getParams().getScale();
    const long double in = ct.getScale();
    const std::string l(label);
    if (!std::isfinite(in) || in <= 0.0L) throw std::invalid_argument(l + ": invalid ciphertext scale");
    if (in < target) throw std::invalid_argument(l + ": EvalMod input scale is below the S2C target scale");
    const long double ratio = in / target;
    const long double r = std::round(ratio);
    if (r < 1.0L || r > (long double)std::numeric_limits<uint64_t>::max() ||
        std::abs(ratio - r) > std::max<long double>(1.0L, std::abs(ratio)) * 1e-3L)
        throw std::invalid_argument(l + ": EvalMod entry scale ratio is not an exact integer");
    return (uint64_t)r;
}

void sparse_trace(const CKKSContext& c, CKKSCiphertext& ct) {
    const auto& P = c.getParams();
    if (!P.usesSparseBootstrapSlots()) return;
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks EvalMod sparse trace: expected a dense-secret ciphertext");
    for (size_t r = P.getBootstrapActiveSlotCount(); r < P.getSlots(); r <<= 1) {
        CKKSCiphertext rot = ckks_rotate(c, ct, (int64_t)r);
        ckks_add_inplace(c, ct, rot);
    }
}

}  // namespace

void ckks_key_switch_dense_to_sparse_inplace(const CKKSContext& c, CKKSCiphertext& ct) {
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_key_switch: dense-to-sparse expects a dense-secret ciphertext");
    if (ct.getNumPolys() != 2 || ct.getLevel() != 0)
        throw std::invalid_argument("ckks_key_switch: dense-to-sparse expects a level-0 2-polynomial ciphertext");
    std::vector<uint64_t*> b, a;
    for (size_t m = 0; m < ct.getBatchSize(); ++m) {
        b.push_back(ct.polyData(0, m));
        a.push_back(ct.polyData(1, m));
    }
    mi_check(mi_q0p0_keyswitch_batch(c.handle(), (int)b.size(), b.data(), a.data(), c.getDenseToSparseSwitchKeyDevice(),
                                     c.getParams().montgomery ? 1 : 0, c.stream()),
             "ckks_key_switch");
    ct.setSecretOwner(CKKSSecretOwner::BootstrapSparse);
}

void ckks_key_switch_sparse_to_dense_inplace(const CKKSContext& c, CKKSCiphertext& ct) {
    if (ct.getSecretOwner() != CKKSSecretOwner::BootstrapSparse)
        throw std::invalid_argument("ckks_key_switch: sparse-to-dense expects a sparse-secret ciphertext");
    if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_key_switch: expected a 2-polynomial ciphertext");
    if (ct.getBatchSize() == 1) {
        mi_check(mi_keyswitch(c.handle(), (int)ct.getLevel(), ct.polyData(1), c.getSparseToDenseSwitchKeyDevice(),
                              ct.polyData(0), ct.polyData(1), 0, 1, c.getParams().montgomery ? 1 : 0, c.stream()),
                 "ckks_key_switch");
    } else {
        // b += KS_b(a), a = KS_a(a) of every member as one key-switch batch
        std::vector<mi_keyswitch_item> items(ct.getBatchSize());
        for (size_t m = 0; m < items.size(); ++m)
            items[m] = mi_keyswitch_item{ct.polyData(1, m), c.getSparseToDenseSwitchKeyDevice(), ct.polyData(0, m),
                                         ct.polyData(1, m),  (int32_t)ct.getLevel(), 0, 1, 0};
        mi_check(mi_keyswitch_batch(c.handle(), items.data(), (int)items.size(), c.getParams().montgomery ? 1 : 0,
                                    c.stream()),
                 "ckks_key_switch");
    }
    ct.setSecretOwner(CKKSSecretOwner::Dense);
}

void ckks_modraise_inplace(const CKKSContext& c, CKKSCiphertext& ct, size_t target_level) {
    const auto& P = c.getParams();
    if (ct.getN() != P.getN()) throw std::invalid_argument("ckks_modraise: ciphertext/context size mismatch");
    if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_modraise: expected a 2-polynomial ciphertext");
    if (ct.getLevel() != 0) throw std::invalid_argument("ckks_modraise: expected a level-0 ciphertext");
    if (ct.getSecretOwner() != CKKSSecretOwner::BootstrapSparse)
        throw std::invalid_argument("ckks_modraise: expected a sparse-secret ciphertext");
    if (target_level == 0 || target_level > P.getMaxLevel())
        throw std::invalid_argument("ckks_modraise: target level must be in [1, max_level]");
    CKKSCiphertext out(c, 2, ct.getScale(), target_level, ct.getSecretOwner(), ct.getMessageEncodingState(),
                       ct.getBatchSize());
    // every poly of every member as one batch
    std::vector<const uint64_t*> src;
    std::vector<uint64_t*> dst;
    for (size_t m = 0; m < ct.getBatchSize(); ++m)
        for (size_t i = 0; i < 2; ++i) {
            src.push_back(ct.polyData(i, m));
            dst.push_back(out.polyData(i, m));
        }
    mi_check(mi_modraise_batch(c.handle(), (int)target_level, (int)src.size(), src.data(), dst.data(), c.stream()),
             "ckks_modraise");
    ct = std::move(out);
}

CKKSCiphertext ckks_s2c_generalized(const CKKSContext& c, const CKKSCiphertext& ct, size_t stage_count) {
    const auto& P = c.getParams();
    const size_t lv = ct.getLevel(), depth = P.getBootstrapS2CDepth();
    if (ct.getN() != P.getN()) throw std::invalid_argument("ckks_s2c_generalized: ciphertext/context size mismatch");
    if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_s2c_generalized: expected a 2-polynomial ciphertext");
    if (lv == 0 || lv > P.getMaxLevel())
        throw std::invalid_argument(
            "ckks_s2c_generalized: S2C requires a ciphertext with at least one available level");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_s2c_generalized: expected a dense-secret input ciphertext");
    if (ct.getMessageEncodingState() != CKKSMessageEncodingState::Slots)
        throw std::invalid_argument("ckks_s2c_generalized: expected slot-encoded input ciphertext");
    if (stage_count == 0 || stage_count != depth || depth > lv)
        throw std::invalid_argument("ckks_s2c_generalized: invalid S2C stage count for input level");
    if (lv != depth)
        throw std::runtime_error("ckks_s2c_generalized: cached EvalMod S2C DFT plan requires input level " +
                                 std::to_string(depth) + ", got level " + std::to_string(lv));
    return ckks_detail::apply_linear_transform_plan_from(c, ct, c.getEvalModS2CPlan());
}

CKKSCiphertext ckks_c2s_generalized(const CKKSContext& c, const CKKSCiphertext& ct, double transform_scale) {
    const auto& P = c.getParams();
    const size_t lv = ct.getLevel(), depth = P.getBootstrapC2SDepth();
    if (ct.getN() != P.getN()) throw std::invalid_argument("ckks_c2s_generalized: ciphertext/context size mismatch");
    if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_c2s_generalized: expected a 2-polynomial ciphertext");
    if (depth == 0) throw std::invalid_argument("ckks_c2s_generalized: C2S depth must be positive");
    if (lv == 0 || lv > P.getMaxLevel())
        throw std::invalid_argument(
            "ckks_c2s_generalized: C2S requires a ciphertext with at least one available level");
    if (depth > lv) throw std::invalid_argument("ckks_c2s_generalized: C2S depth exceeds input ciphertext level");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_c2s_generalized: expected a dense-secret input ciphertext");
    if (ct.getMessageEncodingState() != CKKSMessageEncodingState::Coefficients)
        throw std::invalid_argument("ckks_c2s_generalized: expected coefficient-encoded input ciphertext");
    if (!(transform_scale > 0.0)) throw std::invalid_argument("ckks_c2s_generalized: transform scale must be positive");
    if (lv != P.getDepth())
        throw std::runtime_error("ckks_c2s_generalized: cached EvalMod C2S DFT plan requires input level " +
                                 std::to_string(P.getDepth()) + ", got level " + std::to_string(lv));
    const double cached = c.getEvalModScaledC2STransformScale();
    if (std::abs(transform_scale - cached) > 1e-12 * std::max({1.0, std::abs(transform_scale), std::abs(cached)}))
        throw std::runtime_error("ckks_c2s_generalized: cached EvalMod C2S DFT plan requires transform scale " +
                                 std::to_string(cached) + ", got " + std::to_string(transform_scale));
    return ckks_detail::apply_linear_transform_plan_from(c, ct, c.getEvalModScaledC2SPlan());
}

void ckks_evalmod_binary_inplace(const CKKSContext& c, CKKSCiphertext& ct) {
    const auto& P = c.getParams();
    CKKSEvalPolyOptions opt;
    opt.target_scale = P.getScale();
    ct.setScale(std::exp2((double)P.getEvalModLogScaleBits()));
    ct = ckks_evalchebyshev_preconditioned(c, ct, c.getEvalModPolynomial(), opt);
    ct.setScale(P.getScale());
    double_angle(c, ct, P.getBootstrapParams().evalmod_double_angle);
    ct.setScale(P.getScale());
}

std::pair<CKKSCiphertext, CKKSCiphertext> ckks_bootstrap_pre_evalmod(const CKKSContext& c, CKKSCiphertext packed) {
    const auto& P = c.getParams();
    const std::string l = "ckks_batch_bootstrap_evalmod_from_complex_packed";
    if (packed.getN() != P.getN()) throw std::invalid_argument(l + ": ciphertext/context size mismatch");
    if (packed.getNumPolys() != 2) throw std::invalid_argument(l + ": expected a 2-polynomial ciphertext");
    if (packed.getLevel() != 0) throw std::invalid_argument(l + ": expected a level-0 ciphertext");
    if (packed.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument(l + ": expected a dense-secret ciphertext");
    if (packed.getMessageEncodingState() != CKKSMessageEncodingState::Coefficients)
        throw std::invalid_argument(l + ": expected coefficient-encoded ciphertext");
    if (!(packed.getScale() > 0.0)) throw std::invalid_argument(l + ": invalid ciphertext scale");
    ckks_key_switch_dense_to_sparse_inplace(c, packed);
    ckks_modraise_inplace(c, packed, P.getDepth());
    ckks_key_switch_sparse_to_dense_inplace(c, packed);
    sparse_trace(c, packed);
    packed = ckks_c2s_generalized(c, packed, c.getEvalModScaledC2STransformScale());
    CKKSCiphertext conj = ckks_conjugate(c, packed);
    const long double scale = packed.getScale();
    CKKSCiphertext out0 = packed.clone();
    ckks_add_inplace(c, out0, conj);
    out0.setScale(scale);
    // the reference's second copy of packed is its last use: packed itself
    CKKSCiphertext out1 = std::move(packed);
    ckks_sub_inplace(c, out1, conj);
    ckks_mul_inplace(c, out1, Complex(0.0, -1.0), 1.0);
    out1.setScale(scale);
    return {std::move(out0), std::move(out1)};
}

std::pair<CKKSCiphertext, CKKSCiphertext> ckks_batch_bootstrap_evalmod_from_complex_packed(const CKKSContext& c,
                                                                                           CKKSCiphertext packed) {
    const size_t B = packed.getBatchSize();
    auto lanes = ckks_detail::bootstrap_lanes_stacked(c, std::move(packed)).unstack();
    MyVector<CKKSCiphertext> m0, m1;
    for (size_t m = 0; m < B; ++m) {
        m0.push_back(std::move(lanes[m]));
        m1.push_back(std::move(lanes[B + m]));
    }
    return {CKKSCiphertext::stack(std::move(m0)), CKKSCiphertext::stack(std::move(m1))};
}

CKKSCiphertext ckks_detail::bootstrap_lanes_stacked(const CKKSContext& c, CKKSCiphertext packed) {
    const char* label = "ckks_batch_bootstrap_evalmod_from_complex_packed";
    validate_packed_input(c, packed, label);
    const uint64_t ratio = input_scale_ratio(c, packed, label);
    packed = ckks_apply_linear_transform_plan(c, std::move(packed), c.getScaledS2CPlan(ratio));
    // the S2C output's residues are normalised to q0; its scale restarts at the default
    packed.setScale(c.getParams().getScale());
    auto [out0, out1] = ckks_bootstrap_pre_evalmod(c, std::move(packed));
    // the two lanes have one level and scale: their EvalMods (evalmod_binary_inplace
    // of out0, then of out1) run as one batch of twice the members
    MyVector<CKKSCiphertext> both;
    both.push_back(std::move(out0));
    both.push_back(std::move(out1));
    CKKSCiphertext lanes = CKKSCiphertext::stack(std::move(both));
    ckks_evalmod_binary_inplace(c, lanes);
    return lanes;
}

void ckks_batch_bootstrap_evalmod_inplace(const CKKSContext& c, CKKSCiphertext& ct0, CKKSCiphertext& ct1) {
    validate_packed_input(c, ct0, "ckks_batch_bootstrap_evalmod");
    validate_packed_input(c, ct1, "ckks_batch_bootstrap_evalmod");
    CKKSCiphertext packed = ct0.clone();
    CKKSCiphertext imag = ckks_mul(c, ct1, Complex(0.0, 1.0), 1.0);
    ckks_add_inplace(c, packed, imag);
    packed.setScale(ct0.getScale());
    auto [o0, o1] = ckks_batch_bootstrap_evalmod_from_complex_packed(c, std::move(packed));
    ct0 = std::move(o0);
    ct1 = std::move(o1);
}
