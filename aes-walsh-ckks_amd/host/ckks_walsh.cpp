// The Walsh S-box's output-pair sums (ckks_walsh_sbox.cpp): the AES S-box
// Walsh table, the sum over 4+4 sign products of one output pair (its linear
// combinations, i-combination and raw products fused on the device by
// mi_walsh_product_sum), the map of a Walsh sum to a pair of output bits, and
// the four pairs of one S-box.
#include <cmath>
#include <cstring>

#include "ckks.hpp"

namespace {

constexpr size_t kMasks = 16;  // nibble values (ckks_walsh_sbox.cpp:38)

// scales_close (ckks_walsh_sbox.cpp:164-172)
bool scales_close(long double lhs, long double rhs, long double rel_tol = 0x1p-45L) {
    if (!std::isfinite(lhs) || !std::isfinite(rhs)) return false;
    const long double scale = std::max<long double>({1.0L, std::abs(lhs), std::abs(rhs)});
    return std::abs(lhs - rhs) <= rel_tol * scale;
}

// GF(2^8) modulo x^8 + x^4 + x^3 + x + 1.
uint8_t gf_mul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}

// The AES S-box from its definition (FIPS-197 5.1.1): the multiplicative
// inverse (0 -> 0) followed by the affine map b_i ^ b_{i+4} ^ b_{i+5} ^
// b_{i+6} ^ b_{i+7} ^ 0x63_i.
std::array<uint8_t, 256> aes_sbox() {
    std::array<uint8_t, 256> s{};
    for (int x = 0; x < 256; ++x) {
        uint8_t inv = 0;
        for (int y = 1; y < 256 && x; ++y)
            if (gf_mul((uint8_t)x, (uint8_t)y) == 1) {
                inv = (uint8_t)y;
                break;
            }
        uint8_t out = 0x63;
        for (int i = 0; i < 8; ++i) {
            const int bit = ((inv >> i) ^ (inv >> ((i + 4) & 7)) ^ (inv >> ((i + 5) & 7)) ^ (inv >> ((i + 6) & 7)) ^
                             (inv >> ((i + 7) & 7))) & 1;
            out ^= (uint8_t)(bit << i);
        }
        s[x] = out;
    }
    return s;
}

int parity(unsigned v) { return __builtin_popcount(v) & 1; }

// ckks_add_integer_constant_inplace's rounded constant (ckks_addition.cpp:
// 17-31) as an exact integer: |V| in two words and its sign.
void rounded_constant(int64_t value, long double scale, uint64_t* abs_words, int32_t* neg) {
    const long double v = std::round((long double)value * scale);
    if (!std::isfinite(v) || std::abs(v) >= 0x1p127L)
        throw std::invalid_argument("ckks_walsh_sbox: scaled constant out of range");
    const long double a = std::abs(v);
    const uint64_t hi = (uint64_t)(a / 0x1p64L);
    const uint64_t lo = (uint64_t)(a - (long double)hi * 0x1p64L);
    abs_words[0] = lo;
    abs_words[1] = hi;
    *neg = v < 0 ? 1 : 0;
}

bool ciphertext_is_zero(const CKKSCiphertext& ct) {
    for (size_t m = 0; m < ct.getBatchSize(); ++m)
        for (size_t k = 0; k < ct.getNumPolys(); ++k)
            for (uint64_t w : ct.getPoly(k, m))
                if (w) return false;
    return true;
}

// scale_up_ciphertext_to_match_inplace with multiply_by_integer_scalar_inplace
// (ckks_walsh_sbox.cpp:228-270): the ciphertext's scale is raised to
// `target` by an integer factor.  The reference first tests for an all-zero
// ciphertext (which only takes the new scale); a zero ciphertext times the
// integer is zero too, so the test (a device-to-host copy here) is made only
// where the integer path would throw.
void scale_up_to(const CKKSContext& context, CKKSCiphertext& ct, long double target) {
    if (scales_close(ct.getScale(), target)) {
        ct.setScale(target);
        return;
    }
    const char* err = nullptr;
    const long double scalar = target / ct.getScale();
    const long double rounded = std::round(scalar);
    if (ct.getScale() > target)
        err = "ckks_walsh_sbox: cannot scale down ciphertext for addition";
    else if (rounded < 1.0L ||
             std::abs(scalar - rounded) > std::max<long double>(1.0L, std::abs(scalar)) * 0x1p-20L)
        err = "ckks_walsh_sbox: non-integral scalar scale-up";
    else if (rounded > (long double)std::numeric_limits<int64_t>::max())
        err = "ckks_walsh_sbox: scalar scale-up is too large";
    if (err) {
        if (!ciphertext_is_zero(ct)) throw std::invalid_argument(err);
        ct.setScale(target);
        return;
    }
    const int64_t s = (int64_t)rounded;
    if (s != 1) ckks_mul_inplace(context, ct, s);
    ct.setScale(target);
}

uint64_t bits_of(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return b;
}

// The all-slots constant `value` at `scale`, from the context's encoding
// cache (`own` holds it when the cache is full).
const CKKSEncoding& constant_encoding(const CKKSContext& context, Complex value, long double scale, size_t level,
                                      std::optional<CKKSEncoding>& own) {
    const double sc = (double)scale;
    return *context.cachedEncoding({(uint64_t{1} << 32) | level, bits_of(value.real()), bits_of(value.imag()),
                                    bits_of(sc)},
                                   [&] {
                                       MyVector<Complex> slots(context.getParams().getSlots(), value);
                                       return ckks_encode(context, slots, sc, level, /*eval=*/true);
                                   },
                                   own);
}

}  // namespace

const std::array<uint8_t, 256>& aes_sbox_table() {
    static const std::array<uint8_t, 256> table = aes_sbox();
    return table;
}

namespace ckks_detail {
void scale_up_to(const CKKSContext& context, CKKSCiphertext& ct, long double target) {
    ::scale_up_to(context, ct, target);
}
}  // namespace ckks_detail

const CKKSWalshTable& ckks_aes_walsh_table() {
    static const CKKSWalshTable table = [] {
        const auto& sbox = aes_sbox_table();
        CKKSWalshTable t{};
        for (size_t bit = 0; bit < 8; ++bit)
            for (size_t lo = 0; lo < kMasks; ++lo)
                for (size_t hi = 0; hi < kMasks; ++hi) {
                    const unsigned mask = (unsigned)(lo | (hi << 4));
                    int w = 0;
                    for (unsigned x = 0; x < 256; ++x)
                        w += ((sbox[x] >> bit) & 1) == parity(mask & x) ? 1 : -1;
                    if (w % 4 != 0)
                        throw std::runtime_error("ckks_walsh_sbox: AES Walsh coefficient is not divisible by four");
                    t[bit][lo][hi] = (int16_t)(w / 4);
                }
        return t;
    }();
    return table;
}

namespace {

// The device part of walsh_sum_for_output_pair: acc (2 polys, the h = 0 term)
// and raw (3 polys, the raw products' sum) of one output pair.
// One output pair's product sum, prepared: its tables and its (uninitialised)
// acc and raw ciphertexts; walsh_product_sum_launch runs it (one pair, or the
// four pairs of an S-box as one batch).
struct WalshPairPlan {
    int16_t cr[kMasks][kMasks], ci[kMasks][kMasks];
    uint64_t vabs[4 * kMasks];
    int32_t vneg[2 * kMasks];
    std::optional<CKKSCiphertext> acc, raw;
};

WalshPairPlan walsh_product_plan(const CKKSContext& context, const MyVector<CKKSCiphertext>& low_signs,
                                 const MyVector<CKKSCiphertext>& high_signs, size_t real_out_bit, size_t imag_out_bit,
                                 const CKKSEncoding& encoded_i) {
    if (low_signs.size() != kMasks || high_signs.size() != kMasks)
        throw std::invalid_argument("ckks_walsh_sbox: expected 16 low and 16 high sign ciphertexts");
    if (real_out_bit >= 8 || imag_out_bit >= 8) throw std::invalid_argument("ckks_walsh_sbox: invalid output bit");
    const CKKSCiphertext& ref = low_signs.front();
    const size_t level = ref.getLevel();
    if (level == 0) throw std::invalid_argument("ckks_walsh_sbox: insufficient level for 4+4 Walsh product");
    // linear_combination_with_constant's result takes signs.front()'s metadata;
    // mul_raw sets the low sum's scale to low_signs.front()'s (:297-321)
    const long double S = ref.getScale();
    // each product's scale; their sum keeps the largest (add_term_inplace)
    long double P = 0.0L;
    for (size_t m = 1; m < kMasks; ++m) {
        const CKKSCiphertext& y = high_signs[m];
        P = std::max(P, S * (scales_close(y.getScale(), S) ? y.getScale() : S));
    }
    for (size_t m = 1; m < kMasks; ++m) {
        const CKKSCiphertext& x = low_signs[m];
        const CKKSCiphertext& y = high_signs[m];
        if (x.getLevel() != level || y.getLevel() != level || x.getNumPolys() != 2 || y.getNumPolys() != 2)
            throw std::invalid_argument("ckks_walsh_sbox: sign ciphertexts must be 2-polynomial at one level");
        if (x.getSecretOwner() != ref.getSecretOwner() || y.getSecretOwner() != ref.getSecretOwner() ||
            x.getMessageEncodingState() != ref.getMessageEncodingState() ||
            y.getMessageEncodingState() != ref.getMessageEncodingState())
            throw std::invalid_argument("ckks_walsh_sbox: sign ciphertext metadata mismatch");
        const long double pm = S * (scales_close(y.getScale(), S) ? y.getScale() : S);
        if (!scales_close(pm, P))
            throw std::invalid_argument("ckks_walsh_sbox: high sign scales differ (raw products would need alignment)");
    }
    if (encoded_i.getLevel() != level || encoded_i.getN() != ref.getN())
        throw std::invalid_argument("ckks_walsh_sbox: encoded i level mismatch");

    const auto& table = ckks_aes_walsh_table();
    WalshPairPlan plan;
    for (size_t h = 0; h < kMasks; ++h) {
        for (size_t m = 0; m < kMasks; ++m) {
            plan.cr[h][m] = table[real_out_bit][m][h];
            plan.ci[h][m] = table[imag_out_bit][m][h];
        }
        rounded_constant(plan.cr[h][0], S, &plan.vabs[4 * h], &plan.vneg[2 * h]);
        rounded_constant(plan.ci[h][0], S, &plan.vabs[4 * h + 2], &plan.vneg[2 * h + 1]);
    }
    const size_t B = ref.getBatchSize();
    for (size_t m = 1; m < kMasks; ++m)
        if (low_signs[m].getBatchSize() != B || high_signs[m].getBatchSize() != B)
            throw std::invalid_argument("ckks_walsh_sbox: sign ciphertexts differ in batch size");
    plan.acc.emplace(context, 2, S, level, ref.getSecretOwner(), ref.getMessageEncodingState(), B);
    plan.raw.emplace(context, 3, P, level, ref.getSecretOwner(), ref.getMessageEncodingState(), B);
    return plan;
}

// The product sums of the planned pairs, every member of every pair as one
// mi_walsh_product_sum_batch call.
void walsh_product_launch(const CKKSContext& context, const MyVector<CKKSCiphertext>& low_signs,
                          const MyVector<CKKSCiphertext>& high_signs, const CKKSEncoding& encoded_i,
                          std::vector<WalshPairPlan>& plans) {
    const CKKSCiphertext& ref = low_signs.front();
    const size_t B = ref.getBatchSize();
    std::vector<const uint64_t*> lb, la, hb, ha, vabs;
    std::vector<const int16_t*> cr, ci;
    std::vector<const int32_t*> vneg;
    std::vector<uint64_t*> acc_b, acc_a, raw_b, raw_a, raw_c2;
    for (auto& plan : plans)
        for (size_t j = 0; j < B; ++j) {
            for (size_t m = 1; m < kMasks; ++m) {
                lb.push_back(low_signs[m].polyData(0, j));
                la.push_back(low_signs[m].polyData(1, j));
                hb.push_back(high_signs[m].polyData(0, j));
                ha.push_back(high_signs[m].polyData(1, j));
            }
            cr.push_back(&plan.cr[0][0]);
            ci.push_back(&plan.ci[0][0]);
            vabs.push_back(plan.vabs);
            vneg.push_back(plan.vneg);
            acc_b.push_back(plan.acc->polyData(0, j));
            acc_a.push_back(plan.acc->polyData(1, j));
            raw_b.push_back(plan.raw->polyData(0, j));
            raw_a.push_back(plan.raw->polyData(1, j));
            raw_c2.push_back(plan.raw->polyData(2, j));
        }
    mi_check(mi_walsh_product_sum_batch(context.handle(), (int)ref.getLevel(), (int)acc_b.size(), lb.data(), la.data(),
                                        hb.data(), ha.data(), cr.data(), ci.data(), vabs.data(), vneg.data(),
                                        encoded_i.data(), encoded_i.isMontgomeryForm() ? 1 : 0, acc_b.data(),
                                        acc_a.data(), raw_b.data(), raw_a.data(), raw_c2.data(), context.stream()),
             "ckks_walsh_sbox");
}

std::pair<CKKSCiphertext, CKKSCiphertext> walsh_product_sum(const CKKSContext& context,
                                                            const MyVector<CKKSCiphertext>& low_signs,
                                                            const MyVector<CKKSCiphertext>& high_signs,
                                                            size_t real_out_bit, size_t imag_out_bit,
                                                            const CKKSEncoding& encoded_i) {
    std::vector<WalshPairPlan> plans;
    plans.push_back(walsh_product_plan(context, low_signs, high_signs, real_out_bit, imag_out_bit, encoded_i));
    walsh_product_launch(context, low_signs, high_signs, encoded_i, plans);
    return {std::move(*plans[0].acc), std::move(*plans[0].raw)};
}

// add_term_inplace(acc, relinearised raw) (:288-295)
CKKSCiphertext add_raw_term(const CKKSContext& context, CKKSCiphertext acc, CKKSCiphertext raw) {
    const long double target = std::max(acc.getScale(), raw.getScale());
    scale_up_to(context, acc, target);
    scale_up_to(context, raw, target);
    ckks_add_inplace(context, acc, raw);
    return acc;
}

}  // namespace

CKKSCiphertext ckks_walsh_sum_for_output_pair(const CKKSContext& context, const MyVector<CKKSCiphertext>& low_signs,
                                              const MyVector<CKKSCiphertext>& high_signs, size_t real_out_bit,
                                              size_t imag_out_bit, const CKKSEncoding& encoded_i) {
    auto [acc, raw] = walsh_product_sum(context, low_signs, high_signs, real_out_bit, imag_out_bit, encoded_i);
    ckks_relin_hybrid_inplace(context, raw);  // relin_inplace (:323-328)
    return add_raw_term(context, std::move(acc), std::move(raw));
}

CKKSCiphertext ckks_walsh_sum_to_bit_pair(const CKKSContext& context, CKKSCiphertext walsh_sum) {
    if (walsh_sum.getLevel() == 0)
        throw std::invalid_argument("ckks_walsh_sbox: Walsh sum must have one level for final normalization");
    ckks_rescale(context, walsh_sum);
    const long double output_scale = walsh_sum.getScale() * 128.0L;
    walsh_sum.setScale(output_scale);
    ckks_mul_inplace(context, walsh_sum, int64_t{-1});
    std::optional<CKKSEncoding> own;
    const CKKSEncoding& half = constant_encoding(context, Complex(0.5, 0.5), output_scale, walsh_sum.getLevel(), own);
    ckks_add_inplace(context, walsh_sum, half);
    return walsh_sum;
}

MyVector<CKKSCiphertext> ckks_walsh_recombine_signs_to_paired_outputs(const CKKSContext& context,
                                                                      const MyVector<CKKSCiphertext>& low_signs,
                                                                      const MyVector<CKKSCiphertext>& high_signs) {
    if (low_signs.size() != kMasks || high_signs.size() != kMasks ||
        low_signs[1].getLevel() != high_signs[1].getLevel() ||
        !scales_close(low_signs[1].getScale(), high_signs[1].getScale()))
        throw std::runtime_error("ckks_walsh_sbox: low/high sign metadata mismatch");
    if (low_signs[1].getLevel() == 0)
        throw std::runtime_error("ckks_walsh_sbox: sign output level is too low for 4+4 products");
    std::optional<CKKSEncoding> own;
    const CKKSEncoding& encoded_i = constant_encoding(context, Complex(0.0, 1.0), 1.0L, low_signs[1].getLevel(), own);
    // the four pairs are independent: their product sums first, then their
    // relinearisations as one key-switch batch (one launch per phase)
    std::vector<WalshPairPlan> plans;
    plans.reserve(4);
    for (size_t pair = 0; pair < 4; ++pair)
        plans.push_back(walsh_product_plan(context, low_signs, high_signs, pair, pair + 4, encoded_i));
    walsh_product_launch(context, low_signs, high_signs, encoded_i, plans);
    std::vector<std::pair<CKKSCiphertext, CKKSCiphertext>> sums;
    sums.reserve(4);
    for (auto& plan : plans) sums.emplace_back(std::move(*plan.acc), std::move(*plan.raw));
    if (low_signs.front().getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_relin: sparse-secret relinearization keys are not available");
    std::vector<mi_keyswitch_item> items;
    for (size_t pair = 0; pair < 4; ++pair) {
        CKKSCiphertext& raw = sums[pair].second;
        for (size_t m = 0; m < raw.getBatchSize(); ++m)
            items.push_back(mi_keyswitch_item{raw.polyData(2, m), context.getRelinKeyDevice(), raw.polyData(0, m),
                                              raw.polyData(1, m), (int32_t)raw.getLevel(), 0, 0, 0});
    }
    mi_check(mi_keyswitch_batch(context.handle(), items.data(), (int)items.size(),
                                context.getParams().montgomery ? 1 : 0, context.stream()),
             "ckks_relin");
    MyVector<CKKSCiphertext> outputs;
    outputs.reserve(4);
    for (auto& [acc, raw] : sums) {
        raw.resizePolys(2);
        outputs.emplace_back(ckks_walsh_sum_to_bit_pair(context, add_raw_term(context, std::move(acc), std::move(raw))));
    }
    return outputs;
}

namespace {

constexpr size_t kSegments = 30;  // kPackedWalshSignSegments: 15 low + 15 high masks

// packed_walsh_segment (ckks_walsh_sbox.cpp:476-482)
size_t packed_segment(bool high_nibble, size_t mask) { return (high_nibble ? kMasks - 1 : 0) + (mask - 1); }

CKKSCiphertext constant_one_like(const CKKSContext& context, const CKKSCiphertext& ref) {
    auto out = ref.clone();  // zero_like (:216-220), then the constant 1
    ckks_mul_inplace(context, out, int64_t{0});
    ckks_add_integer_constant_inplace(context, out, int64_t{1});
    return out;
}

void check_segment_keys(const CKKSContext& context, size_t block_count) {
    const size_t slots = context.getParams().getSlots();
    if (block_count == 0) throw std::invalid_argument("ckks_walsh_sbox: mask-packed block count must be nonzero");
    if (block_count * kSegments > slots)
        throw std::invalid_argument("ckks_walsh_sbox: mask-packed block count exceeds slot capacity");
    for (size_t segment = 1; segment < kSegments; ++segment) {
        const int64_t off = (int64_t)(segment * block_count);
        if (!context.hasRotationKey(off) || !context.hasRotationKey(-off))
            throw std::invalid_argument("ckks_walsh_sbox: missing mask-packed segment rotation key");
    }
}

}  // namespace

CKKSCiphertext ckks_walsh_build_mask_packed_parity_input(const CKKSContext& context,
                                                         const MyVector<CKKSCiphertext>& input, size_t block_count) {
    if (input.size() != 8) throw std::invalid_argument("ckks_walsh_sbox: input must contain 8 bit ciphertexts");
    const CKKSCiphertext& ref = input.front();
    if (ref.getN() != context.getParams().getN())
        throw std::invalid_argument("ckks_walsh_sbox: ciphertext/context size mismatch");
    if (ref.getNumPolys() != 2 || ref.getSecretOwner() != CKKSSecretOwner::Dense ||
        ref.getMessageEncodingState() != CKKSMessageEncodingState::Slots)
        throw std::invalid_argument("ckks_walsh_sbox: expected dense slot-encoded 2-polynomial inputs");
    for (const auto& bit : input)
        if (bit.getN() != ref.getN() || bit.getLevel() != ref.getLevel() || bit.getNumPolys() != 2 ||
            bit.getSecretOwner() != ref.getSecretOwner() ||
            bit.getMessageEncodingState() != ref.getMessageEncodingState() ||
            !scales_close(bit.getScale(), ref.getScale()))
            throw std::invalid_argument("ckks_walsh_sbox: inconsistent input ciphertexts");
    check_segment_keys(context, block_count);
    const size_t N = context.getParams().getN();
    // the 30 parity sums in the reference's order (low mask m, then high mask
    // m), as one scaled sum: term t's poly k is the sum of the masked input
    // bits' poly k (parity_sum: the first set bit's copy plus the others)
    const size_t B = ref.getBatchSize();
    MyVector<CKKSCiphertext> terms;
    std::vector<size_t> segments;
    terms.reserve(kSegments);
    const size_t np = kSegments * 2 * B;
    std::vector<uint64_t*> out;
    std::vector<const uint64_t*> in(4 * np, nullptr);
    out.reserve(np);
    for (size_t mask = 1; mask < kMasks; ++mask)
        for (int high = 0; high < 2; ++high) {
            const size_t off = high ? 4 : 0;
            size_t first = 0;
            while (((mask >> first) & 1) == 0) ++first;
            const CKKSCiphertext& lead = input[off + first];
            terms.emplace_back(context, 2, lead.getScale(), lead.getLevel(), lead.getSecretOwner(),
                               lead.getMessageEncodingState(), B);
            segments.push_back(packed_segment(high != 0, mask));
            for (size_t k = 0; k < 2; ++k)
                for (size_t m = 0; m < B; ++m) {
                    const size_t row = out.size();
                    out.push_back(terms.back().polyData(k, m));
                    for (size_t bit = 0; bit < 4; ++bit)
                        if ((mask >> bit) & 1) in[bit * np + row] = input[off + bit].polyData(k, m);
                }
        }
    const size_t L = ref.getLevel() + 1;
    const std::vector<uint64_t> ones(4 * L, 1);
    mi_check(mi_scaled_sum(context.handle(), (int)ref.getLevel(), (int)np, out.data(), 4, in.data(), ones.data(), 0,
                           nullptr, context.stream()),
             "ckks_walsh_sbox");
    // rotate_segment_inplace (:484-495) of every term but segment 0, as one batch
    std::vector<uint64_t*> b, a;
    std::vector<uint64_t> galois;
    std::vector<const uint64_t*> keys;
    for (size_t t = 0; t < terms.size(); ++t) {
        if (segments[t] == 0) continue;
        const int64_t r = normalize_rotation_value(N, (int64_t)(segments[t] * block_count));
        for (size_t m = 0; m < terms[t].getBatchSize(); ++m) {
            b.push_back(terms[t].polyData(0, m));
            a.push_back(terms[t].polyData(1, m));
            galois.push_back(rotation_automorphism_index(N, r));
            keys.push_back(context.getRotationKeyDevice(r));
        }
    }
    mi_check(mi_rotate_batch(context.handle(), (int)ref.getLevel(), (int)b.size(), b.data(), a.data(), galois.data(),
                             keys.data(), context.getParams().montgomery ? 1 : 0, context.stream()),
             "ckks_rotate");
    // add_rotated_segment_term_inplace's accumulation (:497-519) as one sum of
    // the 30 terms (term 0 unrotated) into a new ciphertext with term 0's metadata
    CKKSCiphertext packed(context, 2, terms[0].getScale(), terms[0].getLevel(), terms[0].getSecretOwner(),
                          terms[0].getMessageEncodingState(), B);
    std::vector<uint64_t*> pout;
    std::vector<const uint64_t*> pin(terms.size() * 2 * B);
    for (size_t k = 0; k < 2; ++k)
        for (size_t m = 0; m < B; ++m) pout.push_back(packed.polyData(k, m));
    for (size_t t = 0; t < terms.size(); ++t)
        for (size_t k = 0; k < 2; ++k)
            for (size_t m = 0; m < B; ++m) pin[t * 2 * B + k * B + m] = terms[t].polyData(k, m);
    const std::vector<uint64_t> pones(terms.size() * L, 1);
    mi_check(mi_scaled_sum(context.handle(), (int)ref.getLevel(), (int)(2 * B), pout.data(), (int)terms.size(),
                           pin.data(), pones.data(), 0, nullptr, context.stream()),
             "ckks_walsh_sbox");
    return packed;
}

std::pair<MyVector<CKKSCiphertext>, MyVector<CKKSCiphertext>> ckks_walsh_unpack_mask_packed_signs(
    const CKKSContext& context, const CKKSCiphertext& packed, size_t block_count) {
    if (packed.getNumPolys() != 2 || packed.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_walsh_sbox: invalid hoisted rotation source");
    check_segment_keys(context, block_count);
    const auto& P = context.getParams();
    const size_t N = P.getN(), level = packed.getLevel();
    MyVector<CKKSCiphertext> low, high;
    low.reserve(kMasks);
    high.reserve(kMasks);
    low.push_back(constant_one_like(context, packed));
    high.push_back(constant_one_like(context, packed));
    // one hoist of a (ckks_hoist_ciphertext_a_qp_eval), then every segment's
    // rotation by -offset from it (unpack_segment_from_hoist, :498-509) in one batch
    int64_t info[6];
    mi_check(mi_context_info(context.handle(), info), "ckks_walsh_sbox");
    const size_t L = level + 1, K = P.getPModuli().size(), alpha = (size_t)info[4];
    const size_t parts = std::min<size_t>(P.getDnum(), (L + alpha - 1) / alpha);
    const size_t B = packed.getBatchSize();
    std::vector<size_t> rotated;  // indices into (low | high) of the hoisted outputs
    std::vector<uint64_t> galois;
    std::vector<const uint64_t*> keys;
    for (size_t mask = 1; mask < kMasks; ++mask)
        for (int hi = 0; hi < 2; ++hi) {
            auto& dst = hi ? high : low;
            const size_t segment = packed_segment(hi != 0, mask);
            if (segment == 0) {
                dst.push_back(packed.clone());
                continue;
            }
            const int64_t r = normalize_rotation_value(N, -(int64_t)(segment * block_count));
            if (!context.hasRotationKey(r))
                throw std::invalid_argument("ckks_walsh_sbox: missing hoisted unpack rotation key");
            dst.emplace_back(context, 2, packed.getScale(), level, packed.getSecretOwner(),
                             packed.getMessageEncodingState(), B);
            rotated.push_back((hi ? kMasks : 0) + dst.size() - 1);
            galois.push_back(rotation_automorphism_index(N, r));
            keys.push_back(context.getRotationKeyDevice(r));
        }
    // the hoists of every member as one batch, then each member's rotations
    const size_t dw = parts * (L + K) * N;
    DeviceBuffer digits(context.handle(), B * dw);
    std::vector<const uint64_t*> a_src;
    for (size_t m = 0; m < B; ++m) a_src.push_back(packed.polyData(1, m));
    mi_check(mi_hoist_batch(context.handle(), (int)level, (int)B, a_src.data(), digits.data(), (int64_t)dw,
                            context.stream()),
             "ckks_walsh_sbox");
    // every member's rotation by every segment offset in one call,
    // rotation-major (the members of a launch share the rotation key; the
    // library spreads the launches over its streams)
    std::vector<const uint64_t*> b_in, key;
    std::vector<uint64_t*> b_out, a_out;
    std::vector<uint64_t> gal;
    for (size_t r = 0; r < rotated.size(); ++r) {
        const size_t idx = rotated[r];
        CKKSCiphertext& dst = idx >= kMasks ? high[idx - kMasks] : low[idx];
        for (size_t m = 0; m < B; ++m) {
            b_in.push_back(packed.polyData(0, m));
            b_out.push_back(dst.polyData(0, m));
            a_out.push_back(dst.polyData(1, m));
            gal.push_back(galois[r]);
            key.push_back(keys[r]);
        }
    }
    // item r * B + m reads member m's digits (period B, stride dw)
    mi_check(mi_rotate_hoisted_items(context.handle(), (int)level, (int)b_in.size(), digits.data(), (int64_t)dw,
                                     (int)B, b_in.data(), gal.data(), key.data(), b_out.data(), a_out.data(),
                                     P.montgomery ? 1 : 0, context.stream()),
             "ckks_walsh_sbox");
    return {std::move(low), std::move(high)};
}

namespace {

// encode_active_slot_mask (ckks_aes_ctr_walsh_sbox.cpp:199-215): 1 in the
// first `count` slots, 0 elsewhere, at `scale`, evaluation domain.
const CKKSEncoding& active_slot_mask(const CKKSContext& context, size_t count, size_t level, long double scale,
                                     std::optional<CKKSEncoding>& own) {
    const double sc = (double)scale;
    return *context.cachedEncoding({(uint64_t{2} << 32) | level, (uint64_t)count, 0, bits_of(sc)},
                                   [&] {
                                       MyVector<Complex> slots(context.getParams().getSlots(), Complex(0.0, 0.0));
                                       for (size_t i = 0; i < count; ++i) slots[i] = Complex(1.0, 0.0);
                                       return ckks_encode(context, slots, sc, level, /*eval=*/true);
                                   },
                                   own);
}

// aes_scales_close (ckks_aes_ctr_walsh_sbox.cpp:42-49)
bool aes_scales_close(long double lhs, long double rhs) { return scales_close(lhs, rhs, 0x1p-45L); }

// Every ciphertext times the mask, then (rescale) one rescale of all their
// polys: one batched call instead of the reference's OpenMP loop.
void mask_all(const CKKSContext& context, MyVector<CKKSCiphertext>& cts, const CKKSEncoding& mask, bool rescale,
              const char* level_msg) {
    std::vector<uint64_t*> polys;
    for (auto& ct : cts) {
        if (ct.getLevel() != mask.getLevel()) throw std::invalid_argument(level_msg);
        for (uint64_t* p : ct.allPolyData()) polys.push_back(p);
    }
    const size_t level = mask.getLevel();
    mi_check(mi_mul_plain_rescale_list(context.handle(), (int)level, (int)polys.size(), polys.data(), mask.data(),
                                       mask.isMontgomeryForm() ? 1 : 0, rescale ? 1 : 0, context.stream()),
             "ckks_mul_inplace");
    const long double p = (long double)context.getParams().getModuli()[level];
    for (auto& ct : cts) {
        ct.setScale(ct.getScale() * mask.getScale());
        if (rescale) {
            ct.setScale(ct.getScale() / p);
            ct.setLevel(level - 1);
        }
    }
}

}  // namespace

void ckks_aes_mask_inactive_slots_inplace(const CKKSContext& context, MyVector<CKKSCiphertext>& state,
                                          size_t block_count) {
    if (state.empty()) return;
    const size_t level = state.front().getLevel();
    const size_t s2c_level = context.getParams().getBootstrapS2CDepth();
    if (level <= s2c_level)
        throw std::invalid_argument(
            "ckks_aes_walsh: active slot mask requires a selection limb above S2C input level");
    const long double target_scale = state.front().getScale();
    std::optional<CKKSEncoding> own;
    const CKKSEncoding& mask =
        active_slot_mask(context, block_count, level, (long double)context.getParams().getModuli()[level], own);
    for (const auto& ct : state)
        if (!aes_scales_close(ct.getScale(), target_scale))
            throw std::invalid_argument("ckks_aes_walsh: active slot mask scale mismatch");
    mask_all(context, state, mask, true, "ckks_aes_walsh: active slot mask level mismatch");
    for (auto& ct : state) ct.setScale(target_scale);
}

void ckks_aes_mask_final_refresh_slots_inplace(const CKKSContext& context, MyVector<CKKSCiphertext>& bits,
                                               size_t active_slot_count, size_t minimum_level_to_keep) {
    if (bits.empty()) return;
    const auto& params = context.getParams();
    if (active_slot_count > params.getSlots())
        throw std::invalid_argument("ckks_aes_walsh: final refresh mask active slot count exceeds CKKS slots");
    const size_t level = bits.front().getLevel();
    const bool consume_selection_limb = level > minimum_level_to_keep;
    const long double target_scale = bits.front().getScale();
    const long double mask_scale = consume_selection_limb ? (long double)params.getModuli()[level] : 1.0L;
    std::optional<CKKSEncoding> own;
    const CKKSEncoding& mask = active_slot_mask(context, active_slot_count, level, mask_scale, own);
    mask_all(context, bits, mask, consume_selection_limb, "ckks_aes_walsh: final refresh output mask level mismatch");
    if (consume_selection_limb)
        for (auto& ct : bits) ct.setScale(target_scale);
}

namespace {

// bit_to_sign_inplace (ckks_walsh_sbox.cpp:385-388): b -> 1 - 2b
void bit_to_sign(const CKKSContext& context, CKKSCiphertext& bit) {
    ckks_mul_inplace(context, bit, int64_t{-2});
    ckks_add_integer_constant_inplace(context, bit, int64_t{1});
}

// multiply_by_i (:913-924) and multiply_by_cached_i_inplace without a cache
// (:926-935): times the constant +-i at plaintext scale 1, scale kept
CKKSCiphertext times_i(const CKKSContext& context, const CKKSCiphertext& ct, double sign) {
    auto out = ckks_mul(context, ct, Complex(0.0, sign), 1.0);
    out.setScale(ct.getScale());
    return out;
}

// split_real_imag_pair (:1080-1095): (z + conj z) and -i (z - conj z), both
// read at twice the scale
std::pair<CKKSCiphertext, CKKSCiphertext> split_real_imag_pair(const CKKSContext& context, const CKKSCiphertext& pair,
                                                               const CKKSCiphertext& conjugated) {
    auto real = pair.clone();
    ckks_add_inplace(context, real, conjugated);
    real.setScale(pair.getScale() * 2.0L);
    auto imag = pair.clone();
    ckks_sub_inplace(context, imag, conjugated);
    imag = times_i(context, imag, -1.0);
    imag.setScale(pair.getScale() * 2.0L);
    return {std::move(real), std::move(imag)};
}

// combine_real_imag (:937-958) for two ciphertexts of the same metadata
CKKSCiphertext combine_real_imag(const CKKSContext& context, CKKSCiphertext real, const CKKSCiphertext& imag) {
    auto shifted = times_i(context, imag, 1.0);
    if (real.getLevel() != shifted.getLevel() || real.getNumPolys() != shifted.getNumPolys() ||
        !scales_close(real.getScale(), shifted.getScale()))
        throw std::invalid_argument("ckks_walsh_sbox: real/imaginary lanes do not align");
    ckks_add_inplace(context, real, shifted);
    return real;
}

}  // namespace

MyVector<CKKSCiphertext> ckks_walsh_sbox_complex_blocks_mask_packed(CKKSContext& context,
                                                                    MyVector<CKKSCiphertext> input,
                                                                    size_t block_count) {
    if (input.size() != 8) throw std::invalid_argument("ckks_walsh_sbox: input must contain 8 bit ciphertexts");
    if (input.front().getLevel() != context.getParams().getBootstrapS2CDepth())
        throw std::invalid_argument("ckks_walsh_sbox: input level must equal bootstrap S2C entry depth");
    auto packed = ckks_walsh_build_mask_packed_parity_input(context, input, block_count);
    const size_t B = packed.getBatchSize();
    // bootstrap_complex_packed_parity_lanes: the parity sums refresh to parity
    // bits.  The two lanes go on as one batch (lane 0's members, then lane
    // 1's): bit_to_sign, unpacking and recombination of lane 0 and of lane 1
    // as the same steps over twice the members.
    CKKSCiphertext lanes = ckks_detail::bootstrap_lanes_stacked(context, std::move(packed));
    bit_to_sign(context, lanes);
    auto unpacked = ckks_walsh_unpack_mask_packed_signs(context, lanes, block_count);
    auto paired = ckks_walsh_recombine_signs_to_paired_outputs(context, unpacked.first, unpacked.second);
    MyVector<CKKSCiphertext> paired0, paired1;
    for (auto& p : paired) {
        auto members = p.unstack();
        MyVector<CKKSCiphertext> m0, m1;
        for (size_t m = 0; m < B; ++m) {
            m0.push_back(std::move(members[m]));
            m1.push_back(std::move(members[B + m]));
        }
        paired0.push_back(CKKSCiphertext::stack(std::move(m0)));
        paired1.push_back(CKKSCiphertext::stack(std::move(m1)));
    }
    // combine_paired_lane_outputs_as_complex_bits (:1154-1193)
    std::array<std::optional<CKKSCiphertext>, 8> bits;
    // the eight conjugations of split_real_imag_pair as one key-switch batch
    MyVector<CKKSCiphertext> all;
    for (size_t pair = 0; pair < 4; ++pair) all.push_back(paired0[pair].clone());
    for (size_t pair = 0; pair < 4; ++pair) all.push_back(paired1[pair].clone());
    bool one_level = true;
    for (const auto& ct : all) one_level = one_level && ct.getLevel() == all[0].getLevel() && ct.getNumPolys() == 2;
    MyVector<CKKSCiphertext> conj;
    if (one_level) {
        conj = ckks_conjugate_batch(context, all);
    } else {
        for (const auto& ct : all) conj.push_back(ckks_conjugate(context, ct));
    }
    for (size_t pair = 0; pair < 4; ++pair) {
        auto [real_low, real_high] = split_real_imag_pair(context, paired0[pair], conj[pair]);
        auto [imag_low, imag_high] = split_real_imag_pair(context, paired1[pair], conj[4 + pair]);
        bits[pair] = combine_real_imag(context, std::move(real_low), imag_low);
        bits[pair + 4] = combine_real_imag(context, std::move(real_high), imag_high);
    }
    MyVector<CKKSCiphertext> out;
    out.reserve(8);
    for (auto& b : bits) out.push_back(std::move(*b));
    return out;
}
