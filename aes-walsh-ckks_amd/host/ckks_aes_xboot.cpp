// AES-128 CTR decryption over CKKS with the XBOOT S-box (ckks_aes_ctr.cpp):
// one AES block per slot, each state bit a ciphertext; the S-box is the
// multilinear (Mobius) expansion of each output bit over the 255 products of
// the 8 input bits, the products built layer by layer (pairs, quads, the
// full byte) with a relinearisation and rescale each; MixColumns and
// AddRoundKey are lazy additions; every round ends with a bootstrap of the
// 128 state bits in 64 complex-packed pairs (XBOOT's per-round refresh).
//
// The reference runs the 16 S-boxes and the 64 pair bootstraps on OpenMP
// threads.  Here (options.batch_independent_work) they are ciphertext
// batches: input bit k of the 16 bytes is one batch, so each of the 247
// products of a round's S-boxes is one tensor / key-switch / rescale launch
// sequence for all 16, and the 64 pairs go through one batched bootstrap.
// Every member's limbs equal the one-at-a-time evaluation, which runs when
// the option is off.
#include <algorithm>
#include <cmath>
#include <limits>
#include <optional>

#include "ckks.hpp"

namespace {

constexpr size_t kBytes = 16;
constexpr size_t kBits = 128;
constexpr size_t kRounds = 10;
constexpr size_t kExpandedKeyBits = 176 * 8;

using CKKSByte = MyVector<CKKSCiphertext>;

// aes_scales_close (:48-55)
bool scales_close(long double lhs, long double rhs) {
    if (!std::isfinite(lhs) || !std::isfinite(rhs)) return false;
    const long double m = std::max<long double>({1.0L, std::abs(lhs), std::abs(rhs)});
    return std::abs(lhs - rhs) <= 0x1p-45L * m;
}

bool is_zero(const CKKSCiphertext& ct) {
    for (size_t m = 0; m < ct.getBatchSize(); ++m)
        for (size_t k = 0; k < ct.getNumPolys(); ++k)
            for (uint64_t w : ct.getPoly(k, m))
                if (w) return false;
    return true;
}

// scale_up_ciphertext_to_match_inplace with multiply_by_integer_scalar_inplace
// (:92-137); an all-zero ciphertext only takes the new scale
void scale_up_to(const CKKSContext& context, CKKSCiphertext& ct, long double target) {
    if (scales_close(ct.getScale(), target)) {
        ct.setScale(target);
        return;
    }
    const char* err = nullptr;
    const long double scalar = target / ct.getScale(), rounded = std::round(scalar);
    if (ct.getScale() > target)
        err = "ckks_aes: cannot scale down ciphertext for addition";
    else if (!std::isfinite(scalar) || scalar < 0.0L)
        err = "ckks_aes: invalid scalar scale-up";
    else if (rounded < 1.0L || std::abs(scalar - rounded) > std::max<long double>(1.0L, std::abs(scalar)) * 1e-3L)
        err = "ckks_aes: non-integral scalar scale-up";
    else if (rounded > (long double)std::numeric_limits<int64_t>::max())
        err = "ckks_aes: scalar scale-up is too large";
    if (err) {
        if (!is_zero(ct)) throw std::invalid_argument(err);
        ct.setScale(target);
        return;
    }
    if ((int64_t)rounded != 1) ckks_mul_inplace(context, ct, (int64_t)rounded);
    ct.setScale(target);
}

void pad_with_zero_polys(const CKKSContext& context, CKKSCiphertext& ct, size_t target) {
    if (target < ct.getNumPolys()) throw std::invalid_argument("ckks_aes: cannot shrink ciphertext degree");
    const size_t had = ct.getNumPolys();
    ct.resizePolys(target);
    ckks_detail::zero_polys(context, ct, had, "ckks_aes");
}

// add_term_inplace (:139-158): both operands to the lower level, the higher
// scale and the larger degree, then one addition.
void add_term(const CKKSContext& context, CKKSCiphertext& acc, CKKSCiphertext term) {
    const size_t level = std::min(acc.getLevel(), term.getLevel());
    const long double scale = std::max(acc.getScale(), term.getScale());
    const size_t polys = std::max(acc.getNumPolys(), term.getNumPolys());
    for (CKKSCiphertext* ct : {&acc, &term}) {
        ckks_drop_to_level_inplace(*ct, level);
        pad_with_zero_polys(context, *ct, polys);
        scale_up_to(context, *ct, scale);
    }
    ckks_add_inplace(context, acc, term);
}

// zero_like (:86-90)
CKKSCiphertext zero_like(const CKKSContext& context, const CKKSCiphertext& ref) {
    CKKSCiphertext out(context, ref.getNumPolys(), ref.getScale(), ref.getLevel(), ref.getSecretOwner(),
                       ref.getMessageEncodingState(), ref.getBatchSize());
    ckks_detail::zero_polys(context, out, 0, "ckks_aes");
    return out;
}

// xboot_mul_relin_rescale (:160-183)
CKKSCiphertext mul_relin_rescale(const CKKSContext& context, const CKKSCiphertext& lhs, const CKKSCiphertext& rhs,
                                 long double target_scale) {
    if (lhs.getLevel() == 0 || rhs.getLevel() == 0)
        throw std::invalid_argument("ckks_aes: insufficient level for XBOOT S-box multiplication");
    const size_t level = std::min(lhs.getLevel(), rhs.getLevel());
    CKKSCiphertext out = lhs.cloneAtLevel(level);
    out.setScale(target_scale);
    CKKSCiphertext r = rhs.cloneAtLevel(level);
    r.setScale(target_scale);
    ckks_mul_inplace(context, out, r);
    ckks_relin_hybrid_inplace(context, out);
    ckks_rescale(context, out);
    out.setScale(target_scale);
    return out;
}

struct Monomial {
    size_t mask;
    CKKSCiphertext value;
};

// xboot_layered_monomials (:185-256): all products of the bits begin..end-1
// (a power-of-two range), in the reference's order: the cross products of
// the two halves' monomials, then the left half's, then the right half's.
std::vector<Monomial> layered_monomials(const CKKSContext& context, const CKKSByte& input, size_t begin, size_t end,
                                        long double target_scale) {
    const size_t count = end - begin;
    if (count == 0 || (count & (count - 1)) != 0)
        throw std::invalid_argument("ckks_aes: XBOOT S-box monomial range must be a power of two");
    std::vector<Monomial> out;
    if (count == 1) {
        CKKSCiphertext v = input[begin].clone();
        v.setScale(target_scale);
        out.push_back({size_t{1} << begin, std::move(v)});
        return out;
    }
    if (count == 2) {
        CKKSCiphertext lhs = input[begin].clone();
        lhs.setScale(target_scale);
        CKKSCiphertext rhs = input[begin + 1].clone();
        rhs.setScale(target_scale);
        CKKSCiphertext product = mul_relin_rescale(context, lhs, rhs, target_scale);
        out.push_back({size_t{1} << begin, std::move(lhs)});
        out.push_back({size_t{1} << (begin + 1), std::move(rhs)});
        out.push_back({(size_t{1} << begin) | (size_t{1} << (begin + 1)), std::move(product)});
        return out;
    }
    const size_t mid = begin + count / 2;
    auto left = layered_monomials(context, input, begin, mid, target_scale);
    auto right = layered_monomials(context, input, mid, end, target_scale);
    out.reserve(left.size() * right.size() + left.size() + right.size());
    for (const auto& l : left)
        for (const auto& r : right)
            out.push_back({l.mask | r.mask, mul_relin_rescale(context, l.value, r.value, target_scale)});
    for (auto& t : left) out.push_back(std::move(t));
    for (auto& t : right) out.push_back(std::move(t));
    return out;
}

// xboot_sbox_coefficients (:258-278): the Mobius transform of each S-box
// output bit, i.e. its integer coefficients over the monomials prod_{i in m} x_i.
const std::array<std::array<int64_t, 256>, 8>& sbox_coefficients() {
    static const std::array<std::array<int64_t, 256>, 8> coeffs = [] {
        const auto& sbox = aes_sbox_table();
        std::array<std::array<int64_t, 256>, 8> out{};
        for (size_t bit = 0; bit < 8; ++bit) {
            for (size_t x = 0; x < 256; ++x) out[bit][x] = (int64_t)((sbox[x] >> bit) & 1);
            for (size_t v = 0; v < 8; ++v)
                for (size_t m = 0; m < 256; ++m)
                    if (m & (size_t{1} << v)) out[bit][m] -= out[bit][m ^ (size_t{1} << v)];
        }
        return out;
    }();
    return coeffs;
}

// ckks_aes_xboot_sbox_impl (:280-335) over one byte (or one batch of bytes)
CKKSByte xboot_sbox(const CKKSContext& context, const CKKSByte& input) {
    if (input.size() != 8) throw std::invalid_argument("ckks_aes: XBOOT S-box input must have 8 bits");
    const long double target_scale = input.front().getScale();
    auto layered = layered_monomials(context, input, 0, 8, target_scale);
    std::array<std::optional<CKKSCiphertext>, 256> monomials;
    for (auto& t : layered) {
        if (t.mask == 0 || t.mask >= 256) throw std::runtime_error("ckks_aes: invalid XBOOT S-box monomial mask");
        monomials[t.mask] = std::move(t.value);
    }
    for (size_t m = 1; m < 256; ++m)
        if (!monomials[m]) throw std::runtime_error("ckks_aes: missing XBOOT S-box monomial");
    const auto& coeffs = sbox_coefficients();
    CKKSByte output;
    output.reserve(8);
    for (size_t bit = 0; bit < 8; ++bit) {
        std::optional<CKKSCiphertext> acc;
        if (coeffs[bit][0] != 0) {
            acc = zero_like(context, input.front());
            ckks_add_integer_constant_inplace(context, *acc, coeffs[bit][0]);
        }
        for (size_t m = 1; m < 256; ++m) {
            const int64_t c = coeffs[bit][m];
            if (c == 0) continue;
            CKKSCiphertext term = monomials[m]->clone();
            ckks_mul_inplace(context, term, c);
            if (acc)
                add_term(context, *acc, std::move(term));
            else
                acc = std::move(term);
        }
        if (!acc) acc = zero_like(context, input.front());
        output.push_back(std::move(*acc));
    }
    return output;
}

// sub_bytes_inplace (:474-509)
void sub_bytes(const CKKSContext& context, MyVector<CKKSCiphertext>& state, bool batched) {
    if (state.size() != kBits) throw std::invalid_argument("ckks_aes: SubBytes state must have 128 bits");
    if (batched) {
        CKKSByte input;
        for (size_t k = 0; k < 8; ++k) {
            MyVector<CKKSCiphertext> members;
            for (size_t byte = 0; byte < kBytes; ++byte) members.push_back(std::move(state[8 * byte + k]));
            input.push_back(CKKSCiphertext::stack(std::move(members)));
        }
        auto out = xboot_sbox(context, input);
        for (size_t k = 0; k < 8; ++k) {
            auto split = out[k].unstack();
            for (size_t byte = 0; byte < kBytes; ++byte) state[8 * byte + k] = std::move(split[byte]);
        }
        return;
    }
    for (size_t byte = 0; byte < kBytes; ++byte) {
        CKKSByte input;
        for (size_t k = 0; k < 8; ++k) input.push_back(state[8 * byte + k].clone());
        auto out = xboot_sbox(context, input);
        for (size_t k = 0; k < 8; ++k) state[8 * byte + k] = std::move(out[k]);
    }
}

// shift_rows_inplace (:511-526)
void shift_rows(MyVector<CKKSCiphertext>& state) {
    MyVector<CKKSCiphertext> old = std::move(state);
    state.clear();
    state.reserve(kBits);
    for (size_t c = 0; c < 4; ++c)
        for (size_t r = 0; r < 4; ++r)
            for (size_t k = 0; k < 8; ++k) state.push_back(std::move(old[(4 * ((c + r) % 4) + r) * 8 + k]));
}

CKKSCiphertext add_bits(const CKKSContext& context, const CKKSCiphertext& a, const CKKSCiphertext& b) {
    CKKSCiphertext out = a.clone();
    ckks_add_inplace(context, out, b);
    return out;
}

// fmix_column_byte (:415-443): xtime(b0 + b1) + b1 + b2 + b3, where
// xtime(x) = (x7, x0 + x7, x1, x2 + x7, x3 + x7, x4, x5, x6)
CKKSByte mix_column_byte(const CKKSContext& context, const MyVector<CKKSCiphertext>& st, size_t i0, size_t i1,
                         size_t i2, size_t i3) {
    CKKSByte sum;
    for (size_t k = 0; k < 8; ++k) sum.push_back(add_bits(context, st[8 * i0 + k], st[8 * i1 + k]));
    CKKSByte out;
    out.push_back(sum[7].clone());
    out.push_back(add_bits(context, sum[0], sum[7]));
    out.push_back(sum[1].clone());
    out.push_back(add_bits(context, sum[2], sum[7]));
    out.push_back(add_bits(context, sum[3], sum[7]));
    out.push_back(sum[4].clone());
    out.push_back(sum[5].clone());
    out.push_back(sum[6].clone());
    for (size_t idx : {i1, i2, i3})
        for (size_t k = 0; k < 8; ++k) ckks_add_inplace(context, out[k], st[8 * idx + k]);
    return out;
}

// fmix_columns_inplace (:528-554)
void mix_columns(const CKKSContext& context, MyVector<CKKSCiphertext>& state) {
    std::array<CKKSByte, kBytes> bytes;
    for (size_t c = 0; c < 4; ++c) {
        const size_t b0 = 4 * c, b1 = b0 + 1, b2 = b0 + 2, b3 = b0 + 3;
        bytes[b0] = mix_column_byte(context, state, b0, b1, b2, b3);
        bytes[b1] = mix_column_byte(context, state, b1, b2, b3, b0);
        bytes[b2] = mix_column_byte(context, state, b2, b3, b0, b1);
        bytes[b3] = mix_column_byte(context, state, b3, b0, b1, b2);
    }
    state.clear();
    for (auto& byte : bytes)
        for (auto& bit : byte) state.push_back(std::move(bit));
}

// add_round_key_lazy_inplace (:460-472)
void add_round_key(const CKKSContext& context, MyVector<CKKSCiphertext>& state, const MyVector<CKKSCiphertext>& keys,
                   size_t round) {
    if (state.size() != kBits || round > kRounds) throw std::invalid_argument("ckks_aes: invalid AddRoundKey input");
    for (size_t bit = 0; bit < kBits; ++bit) ckks_add_inplace(context, state[bit], keys[round * kBits + bit]);
}

// xboot_state_inplace (:556-574): the 128 bits refreshed as 64 complex-packed
// pairs (bit 2p in the real lane, 2p + 1 in the imaginary one)
void xboot_state(const CKKSContext& context, MyVector<CKKSCiphertext>& state, bool batched) {
    if (state.size() != kBits) throw std::invalid_argument("ckks_aes: XBOOT state must have 128 bits");
    if (batched) {
        MyVector<CKKSCiphertext> re, im;
        for (size_t p = 0; p < kBits / 2; ++p) {
            re.push_back(std::move(state[2 * p]));
            im.push_back(std::move(state[2 * p + 1]));
        }
        CKKSCiphertext ct0 = CKKSCiphertext::stack(std::move(re)), ct1 = CKKSCiphertext::stack(std::move(im));
        ckks_batch_bootstrap_evalmod_inplace(context, ct0, ct1);
        auto r = ct0.unstack(), i = ct1.unstack();
        for (size_t p = 0; p < kBits / 2; ++p) {
            state[2 * p] = std::move(r[p]);
            state[2 * p + 1] = std::move(i[p]);
        }
        return;
    }
    for (size_t p = 0; p < kBits / 2; ++p) ckks_batch_bootstrap_evalmod_inplace(context, state[2 * p], state[2 * p + 1]);
}

size_t logical_slot_count(const CKKSParams& P) {
    return P.hasBootstrapParams() ? P.getBootstrapActiveSlotCount() : P.getSlots();
}

// encode_message_block (:337-363): bit `bit` of block (slot mod active) in
// every slot, at the default scale
CKKSEncoding encode_message_block(const CKKSContext& context, std::span<const std::array<uint8_t, 16>> blocks,
                                  size_t bit, size_t level) {
    const auto& P = context.getParams();
    const size_t logical = logical_slot_count(P);
    if (blocks.empty() || blocks.size() > logical)
        throw std::invalid_argument("ckks_aes: public block count exceeds logical slot count");
    MyVector<Complex> v(P.getSlots());
    for (size_t s = 0; s < P.getSlots(); ++s) {
        const size_t ls = s % logical;
        const double b = ls < blocks.size() ? (double)((blocks[ls][bit / 8] >> (bit % 8)) & 1) : 0.0;
        v[s] = Complex(b, 0.0);
    }
    return ckks_encode(context, v, level, /*eval=*/true);
}

// make_batched_initial_counter_state (:576-683): round key 0 XOR the public
// counter bit plane; x XOR b = (1 - 2b) x + b, with the all-zero and all-one
// planes as a level drop and 1 - x
MyVector<CKKSCiphertext> initial_counter_state(const CKKSContext& context, const MyVector<CKKSCiphertext>& keys,
                                               std::span<const std::array<uint8_t, 16>> counters) {
    const auto& P = context.getParams();
    const size_t logical = logical_slot_count(P);
    if (counters.empty() || counters.size() > logical)
        throw std::invalid_argument("ckks_aes: public counter count exceeds logical slot count");
    MyVector<CKKSCiphertext> state;
    state.reserve(kBits);
    for (size_t bit = 0; bit < kBits; ++bit) {
        bool saw_zero = counters.size() < logical, saw_one = false;
        for (const auto& b : counters) {
            const bool v = ((b[bit / 8] >> (bit % 8)) & 1) != 0;
            saw_zero = saw_zero || !v;
            saw_one = saw_one || v;
        }
        CKKSCiphertext out = keys[bit].clone();
        if (out.getLevel() == 0) throw std::invalid_argument("ckks_aes: exact public XOR requires one level");
        const size_t in_level = out.getLevel();
        if (!saw_one) {
            ckks_drop_to_level_inplace(out, in_level - 1);
        } else if (!saw_zero) {
            ckks_drop_to_level_inplace(out, in_level - 1);
            ckks_mul_inplace(context, out, int64_t{-1});
            ckks_add_integer_constant_inplace(context, out, int64_t{1});
        } else {
            MyVector<Complex> signs(P.getSlots()), bits(P.getSlots());
            for (size_t s = 0; s < P.getSlots(); ++s) {
                const size_t ls = s % logical;
                const int v = ls < counters.size() ? (counters[ls][bit / 8] >> (bit % 8)) & 1 : 0;
                bits[s] = Complex((double)v, 0.0);
                signs[s] = Complex(v == 0 ? 1.0 : -1.0, 0.0);
            }
            ckks_mul_inplace(context, out,
                             ckks_encode(context, signs, (double)P.getModuli()[in_level], in_level, /*eval=*/true));
            ckks_rescale(context, out);
            ckks_add_inplace(context, out, ckks_encode(context, bits, out.getLevel(), /*eval=*/true));
        }
        state.push_back(std::move(out));
    }
    return state;
}

}  // namespace

MyVector<CKKSCiphertext> ckks_aes128_encrypt_expanded_round_key_bits(const CKKSContext& context,
                                                                     const std::array<uint8_t, 176>& rk,
                                                                     size_t first_round_key_level) {
    const auto& P = context.getParams();
    const size_t base = P.getBootstrapS2CDepth();
    if (base + first_round_key_level > P.getMaxLevel())
        throw std::invalid_argument("ckks_aes: invalid encryption level for first round key bits");
    MyVector<CKKSCiphertext> out;
    out.reserve(kExpandedKeyBits);
    for (size_t bit = 0; bit < kExpandedKeyBits; ++bit) {
        const size_t level = bit / kBits == 0 ? base + first_round_key_level : base;
        const double v = ((rk[bit / 8] >> (bit % 8)) & 1) ? 1.0 : 0.0;
        MyVector<Complex> slots(P.getSlots(), Complex(v, 0.0));
        out.push_back(ckks_encrypt(context, slots, level));
    }
    return out;
}

MyVector<CKKSCiphertext> ckks_aes_ctr(CKKSContext& context, const MyVector<CKKSCiphertext>& keys,
                                      std::span<const std::array<uint8_t, 16>> counters,
                                      std::span<const std::array<uint8_t, 16>> ciphertexts,
                                      const CKKSAes128CtrOptions& options) {
    if (keys.size() != kExpandedKeyBits) throw std::invalid_argument("ckks_aes: expanded round key must have 1408 bits");
    const bool batched = options.batch_independent_work;
    // round-key bits of another context of the same keys (a lane's owner)
    // are copied onto this context's stream and pool first
    MyVector<CKKSCiphertext> local;
    if (&keys.front().getContext() != &context) {
        const CKKSContext& src = keys.front().getContext();
        mi_check(mi_stream_synchronize(src.handle(), src.stream()), "ckks_aes");
        local.reserve(keys.size());
        for (const auto& k : keys) {
            CKKSCiphertext c(context, k.getNumPolys(), k.getScale(), k.getLevel(), k.getSecretOwner(),
                             k.getMessageEncodingState());
            std::vector<uint64_t*> to = c.allPolyData();
            std::vector<const uint64_t*> from = k.allPolyData();
            mi_check(mi_poly_list_op(context.handle(), MI_LIST_COPY, (int)k.getLevel(), (int)to.size(), to.data(),
                                     from.data(), nullptr, context.stream()),
                     "ckks_aes");
            local.push_back(std::move(c));
        }
    }
    const MyVector<CKKSCiphertext>& round_keys = local.empty() ? keys : local;
    // evaluate_aes128_rounds_from_state (:727-755)
    MyVector<CKKSCiphertext> state = initial_counter_state(context, round_keys, counters);
    for (size_t round = 1; round < kRounds; ++round) {
        sub_bytes(context, state, batched);
        shift_rows(state);
        mix_columns(context, state);
        add_round_key(context, state, round_keys, round);
        xboot_state(context, state, batched);
    }
    sub_bytes(context, state, batched);
    shift_rows(state);
    add_round_key(context, state, round_keys, kRounds);
    // ckks_aes_ctr (:757-779): plaintext = ciphertext XOR keystream, refreshed
    for (size_t bit = 0; bit < kBits; ++bit)
        ckks_add_inplace(context, state[bit], encode_message_block(context, ciphertexts, bit, state[bit].getLevel()));
    xboot_state(context, state, batched);
    return state;
}
