// AES-128 CTR decryption over CKKS with the mask-packed Walsh S-box
// (ckks_aes_ctr_walsh_sbox.cpp), and the plain AES-128 it is checked against
// (math/aes/aes.cpp).
//
// The state is 128 bit ciphertexts (byte b, bit k at 8b + k; byte b = column
// b / 4, row b % 4).  XOR is a lazy addition: bits are sums of 0/1 values
// and the S-box's parity packing only reads them mod 2, so AddRoundKey,
// MixColumns and the public counter / ciphertext bits are exact additions of
// ciphertexts (with the scale raised by an integer factor where the operands'
// scales differ).  Every step is exact arithmetic mod q in the same order of
// operations as the reference, so each limb equals the reference's.
#include <algorithm>
#include <cmath>
#include <exception>
#include <limits>
#include <thread>

#include "ckks.hpp"

namespace {

constexpr size_t kBlockBytes = 16;
constexpr size_t kBlockBits = 128;
constexpr size_t kRounds = 10;
constexpr size_t kExpandedKeyBits = 176 * 8;
constexpr size_t kMinimumPackedSegments = 30;  // kWalshMinimumPackedSegments

using Block = std::array<uint8_t, 16>;
using CKKSByte = MyVector<CKKSCiphertext>;

uint8_t xtime(uint8_t v) { return (uint8_t)((v << 1) ^ ((v & 0x80) ? 0x1b : 0)); }

}  // namespace

// FIPS-197 5.2: words w[0..43], w[i] = w[i-4] ^ (i % 4 == 0 ?
// SubWord(RotWord(w[i-1])) ^ Rcon[i/4] : w[i-1]); bytes in word order.
std::array<uint8_t, 176> aes128_expand_key(const std::array<uint8_t, 16>& key) {
    const auto& sbox = aes_sbox_table();
    std::array<uint8_t, 176> w{};
    std::copy(key.begin(), key.end(), w.begin());
    uint8_t rcon = 1;
    for (size_t i = 4; i < 44; ++i) {
        uint8_t t[4] = {w[4 * i - 4], w[4 * i - 3], w[4 * i - 2], w[4 * i - 1]};
        if (i % 4 == 0) {
            const uint8_t t0 = t[0];
            t[0] = (uint8_t)(sbox[t[1]] ^ rcon);
            t[1] = sbox[t[2]];
            t[2] = sbox[t[3]];
            t[3] = sbox[t0];
            rcon = xtime(rcon);
        }
        for (size_t k = 0; k < 4; ++k) w[4 * i + k] = (uint8_t)(w[4 * i - 16 + k] ^ t[k]);
    }
    return w;
}

// FIPS-197 5.1: state byte 4c + r is row r of column c.
std::array<uint8_t, 16> aes128_encrypt_block_expanded(const std::array<uint8_t, 176>& rk,
                                                      const std::array<uint8_t, 16>& in) {
    const auto& sbox = aes_sbox_table();
    Block s;
    for (size_t i = 0; i < 16; ++i) s[i] = (uint8_t)(in[i] ^ rk[i]);
    for (size_t round = 1; round <= kRounds; ++round) {
        Block t;
        for (size_t c = 0; c < 4; ++c)
            for (size_t r = 0; r < 4; ++r) t[4 * c + r] = sbox[s[4 * ((c + r) % 4) + r]];
        if (round != kRounds)
            for (size_t c = 0; c < 4; ++c) {
                const uint8_t a0 = t[4 * c], a1 = t[4 * c + 1], a2 = t[4 * c + 2], a3 = t[4 * c + 3];
                t[4 * c + 0] = (uint8_t)(xtime(a0 ^ a1) ^ a1 ^ a2 ^ a3);
                t[4 * c + 1] = (uint8_t)(xtime(a1 ^ a2) ^ a2 ^ a3 ^ a0);
                t[4 * c + 2] = (uint8_t)(xtime(a2 ^ a3) ^ a3 ^ a0 ^ a1);
                t[4 * c + 3] = (uint8_t)(xtime(a3 ^ a0) ^ a0 ^ a1 ^ a2);
            }
        for (size_t i = 0; i < 16; ++i) s[i] = (uint8_t)(t[i] ^ rk[16 * round + i]);
    }
    return s;
}

// aes128_encrypt_block / aes128_ctr_crypt_block (math/aes/aes.cpp:100-125)
std::array<uint8_t, 16> aes128_encrypt_block(const std::array<uint8_t, 16>& key, const std::array<uint8_t, 16>& in) {
    return aes128_encrypt_block_expanded(aes128_expand_key(key), in);
}

std::array<uint8_t, 16> aes128_ctr_crypt_block(const std::array<uint8_t, 16>& key, const std::array<uint8_t, 16>& counter,
                                               const std::array<uint8_t, 16>& input) {
    const auto keystream = aes128_encrypt_block(key, counter);
    std::array<uint8_t, 16> out;
    for (size_t i = 0; i < 16; ++i) out[i] = (uint8_t)(input[i] ^ keystream[i]);
    return out;
}

namespace ckks_detail {

// aes_scales_close (:42-49)
static bool aes_scales_close(long double lhs, long double rhs) {
    if (!std::isfinite(lhs) || !std::isfinite(rhs)) return false;
    const long double m = std::max<long double>({1.0L, std::abs(lhs), std::abs(rhs)});
    return std::abs(lhs - rhs) <= 0x1p-45L * m;
}

// The integer factor scale_up_ciphertext_to_match_inplace (:105-124) with
// multiply_by_integer_scalar_inplace (:80-103) multiplies a ciphertext of
// scale `from` by to reach `to`: 1 when the scales are close, otherwise the
// rounded ratio within the reference's 1e-3 integrality tolerance.  Returns
// the reference's error message where it would throw (for a non-zero
// ciphertext), nullptr otherwise.
const char* aes_scale_up_factor(long double from, long double to, int64_t* factor) {
    *factor = 1;
    if (aes_scales_close(from, to)) return nullptr;
    if (from > to) return "ckks_aes_walsh: cannot scale down ciphertext for addition";
    const long double scalar = to / from;
    if (!std::isfinite(scalar) || scalar < 0.0L) return "ckks_aes_walsh: invalid scalar scale-up";
    const long double rounded = std::round(scalar);
    if (rounded < 1.0L || std::abs(scalar - rounded) > std::max<long double>(1.0L, std::abs(scalar)) * 1e-3L)
        return "ckks_aes_walsh: non-integral scalar scale-up";
    if (rounded > (long double)std::numeric_limits<int64_t>::max())
        return "ckks_aes_walsh: scalar scale-up is too large";
    *factor = (int64_t)rounded;
    return nullptr;
}

}  // namespace ckks_detail

using ckks_detail::aes_scale_up_factor;

namespace {

// pad_ciphertext_with_zero_polys (ckks_aes_ctr_walsh_sbox.cpp:58-66)
void pad_with_zero_polys(const CKKSContext& context, CKKSCiphertext& ct, size_t target) {
    if (target < ct.getNumPolys()) throw std::invalid_argument("ckks_aes_walsh: cannot shrink ciphertext degree");
    while (ct.getNumPolys() < target) {
        ct.addPoly();
        mi_check(mi_memset_zero(context.handle(), ct.polyData(ct.getNumPolys() - 1),
                                (ct.getLevel() + 1) * ct.getN() * sizeof(uint64_t), context.stream()),
                 "ckks_aes_walsh");
    }
}

bool is_zero_ciphertext(const CKKSCiphertext& ct) {
    for (size_t m = 0; m < ct.getBatchSize(); ++m)
        for (size_t k = 0; k < ct.getNumPolys(); ++k)
            for (uint64_t w : ct.getPoly(k, m))
                if (w) return false;
    return true;
}

// scale_up_ciphertext_to_match_inplace (:105-124): an all-zero ciphertext
// only takes the new scale (tested here only where the integer path throws:
// a zero times the factor is zero too).
void aes_scale_up_to(const CKKSContext& context, CKKSCiphertext& ct, long double target) {
    int64_t factor;
    if (const char* err = aes_scale_up_factor(ct.getScale(), target, &factor)) {
        if (!is_zero_ciphertext(ct)) throw std::invalid_argument(err);
        ct.setScale(target);
        return;
    }
    if (factor != 1) ckks_mul_inplace(context, ct, factor);
    ct.setScale(target);
}

// align_for_add_exact / add_term_inplace (:126-147): both operands to the
// lower level, the higher scale (an integer scale-up) and the larger degree,
// then one addition.
void add_term(const CKKSContext& context, CKKSCiphertext& acc, CKKSCiphertext term) {
    const size_t level = std::min(acc.getLevel(), term.getLevel());
    const long double scale = std::max(acc.getScale(), term.getScale());
    const size_t polys = std::max(acc.getNumPolys(), term.getNumPolys());
    for (CKKSCiphertext* ct : {&acc, &term}) {
        ckks_drop_to_level_inplace(*ct, level);
        pad_with_zero_polys(context, *ct, polys);
        aes_scale_up_to(context, *ct, scale);
    }
    ckks_add_inplace(context, acc, term);
}

// add_term on an operand the caller keeps: when it already has acc's level,
// degree and (within 2^-45) scale, the alignment changes nothing but acc's
// scale label, so it is added in place of a copy.
void add_term_ref(const CKKSContext& context, CKKSCiphertext& acc, const CKKSCiphertext& term) {
    const long double sa = acc.getScale(), st = term.getScale();
    const long double m = std::max<long double>({1.0L, std::abs(sa), std::abs(st)});
    if (term.getLevel() == acc.getLevel() && term.getNumPolys() == acc.getNumPolys() && std::isfinite(sa) &&
        std::isfinite(st) && std::abs(sa - st) <= 0x1p-45L * m) {
        acc.setScale(std::max(sa, st));
        ckks_add_inplace(context, acc, term);
        return;
    }
    add_term(context, acc, term.cloneAtLevel(std::min(acc.getLevel(), term.getLevel())));
}

CKKSCiphertext add_bits(const CKKSContext& context, const CKKSCiphertext& lhs, const CKKSCiphertext& rhs) {
    auto out = lhs.clone();
    add_term_ref(context, out, rhs);
    return out;
}

// walsh_packing_block_count (:154-171)
size_t packing_block_count(const CKKSContext& context, size_t active) {
    const auto& P = context.getParams();
    const size_t pack = P.usesAesWalsh() ? P.getAesWalshBlockCount() : P.getSlots() / kMinimumPackedSegments;
    const size_t segments = P.usesAesWalsh() ? P.getAesWalshPackedSegmentCount() : kMinimumPackedSegments;
    if (active == 0 || pack == 0 || segments < kMinimumPackedSegments || active > pack ||
        segments > P.getSlots() / pack)
        throw std::invalid_argument("ckks_aes_walsh: active blocks exceed padded Walsh packing capacity");
    return pack;
}

// aes_walsh_state_entry_level (:173-175): one selection limb above S2C's input
size_t state_entry_level(const CKKSContext& context) { return context.getParams().getBootstrapS2CDepth() + 1; }

// encode_active_complex_block_public_bit (:177-197): bit `bit` of block s in
// the real lane and of block n + s in the imaginary lane of slot s.
CKKSEncoding encode_public_bit(const CKKSContext& context, std::span<const Block> blocks, size_t bit, size_t level,
                               long double scale) {
    const size_t n = blocks.size() / 2;
    MyVector<Complex> slots(context.getParams().getSlots(), Complex(0.0, 0.0));
    for (size_t s = 0; s < n; ++s)
        slots[s] = Complex((double)((blocks[s][bit / 8] >> (bit % 8)) & 1),
                           (double)((blocks[n + s][bit / 8] >> (bit % 8)) & 1));
    return ckks_encode(context, slots, (double)scale, level, /*eval=*/true);
}

// encode_active_complex_block_public_bit + ckks_add_inplace(ct, pt) for all
// 128 bits (add_public_complex_block_bits_lazy_inplace, :492-532, and the
// counter half of make_initial_counter_complex_block_state, :534-559): one
// batched device encoding of the 128 slot vectors at the bits' common level
// and scale, then one launch adding each to its bit's b poly (the encodings
// are not in Montgomery form, so the addition is the plain one
// ckks_add_inplace makes).  Bits of differing level or scale take the
// per-bit path.
void add_public_bits(const CKKSContext& context, MyVector<CKKSCiphertext>& bits, std::span<const Block> blocks,
                     bool batched) {
    const size_t level = bits[0].getLevel();
    const long double scale = bits[0].getScale();
    bool uniform = batched;
    for (const auto& b : bits) uniform = uniform && b.getLevel() == level && b.getScale() == scale;
    if (!uniform || !(scale > 0)) {
        for (size_t bit = 0; bit < bits.size(); ++bit)
            ckks_add_inplace(context, bits[bit],
                             encode_public_bit(context, blocks, bit, bits[bit].getLevel(), bits[bit].getScale()));
        return;
    }
    const auto& P = context.getParams();
    const size_t slots = P.getSlots(), N = P.getN(), words = (level + 1) * N, n = blocks.size() / 2;
    std::vector<Complex> vals(bits.size() * slots, Complex(0.0, 0.0));
    for (size_t bit = 0; bit < bits.size(); ++bit)
        for (size_t s = 0; s < n; ++s)
            vals[bit * slots + s] = Complex((double)((blocks[s][bit / 8] >> (bit % 8)) & 1),
                                            (double)((blocks[n + s][bit / 8] >> (bit % 8)) & 1));
    DeviceBuffer enc(context.handle(), bits.size() * words);
    mi_check(mi_encode_slots(context.handle(), reinterpret_cast<const double*>(vals.data()), (int)bits.size(),
                             (double)scale, (int)(level + 1), 1, enc.data(), (int64_t)words, context.stream()),
             "ckks_encode");
    std::vector<uint64_t*> x(bits.size());
    std::vector<const uint64_t*> y(bits.size());
    for (size_t bit = 0; bit < bits.size(); ++bit) {
        x[bit] = bits[bit].polyData(0);
        y[bit] = enc.data() + bit * words;
    }
    mi_check(mi_poly_list_op(context.handle(), MI_LIST_ADD, (int)level, (int)bits.size(), x.data(), y.data(), nullptr,
                             context.stream()),
             "ckks_add_inplace(ct,pt)");
}

// In-place rotations of many ciphertexts of one level as one batch
// (each equal to ckks_rotate_inplace).
void rotate_batch(const CKKSContext& context, std::vector<CKKSCiphertext*>& cts, const std::vector<int64_t>& rot) {
    if (cts.empty()) return;
    const size_t N = context.getParams().getN();
    std::vector<uint64_t*> b, a;
    std::vector<uint64_t> galois;
    std::vector<const uint64_t*> keys;
    for (size_t i = 0; i < cts.size(); ++i) {
        const int64_t r = normalize_rotation_value(N, rot[i]);
        if (!context.hasRotationKey(r)) throw std::invalid_argument("ckks_aes_walsh: missing refresh rotation key");
        if (cts[i]->getNumPolys() != 2 || cts[i]->getLevel() != cts[0]->getLevel())
            throw std::invalid_argument("ckks_aes_walsh: refresh rotation inputs differ in shape");
        b.push_back(cts[i]->polyData(0));
        a.push_back(cts[i]->polyData(1));
        galois.push_back(rotation_automorphism_index(N, r));
        keys.push_back(context.getRotationKeyDevice(r));
    }
    mi_check(mi_rotate_batch(context.handle(), (int)cts[0]->getLevel(), (int)b.size(), b.data(), a.data(),
                             galois.data(), keys.data(), context.getParams().montgomery ? 1 : 0, context.stream()),
             "ckks_rotate");
}

CKKSCiphertext copy_to_context(const CKKSContext& dst, const CKKSCiphertext& src);

// Jobs 0..n-1 on contexts[j mod count], one host thread per context (the
// calling thread's stream is synchronised first, each worker synchronises
// its own stream at the end); `job(j, context)` runs on its context only.
template <class Job>
void run_on_contexts(CKKSContext& main, const std::vector<CKKSContext*>& ctxs, size_t n, Job job) {
    auto run = [&](size_t w) {
        for (size_t j = w; j < n; j += ctxs.size()) job(j, *ctxs[w]);
        mi_check(mi_stream_synchronize(ctxs[w]->handle(), ctxs[w]->stream()), "ckks_aes_walsh");
    };
    if (ctxs.size() == 1) {
        run(0);
        return;
    }
    mi_check(mi_stream_synchronize(main.handle(), main.stream()), "ckks_aes_walsh");
    std::vector<std::exception_ptr> errors(ctxs.size());
    std::vector<std::thread> threads;
    for (size_t w = 0; w < ctxs.size(); ++w)
        threads.emplace_back([&, w] {
            try {
                run(w);
            } catch (...) {
                errors[w] = std::current_exception();
            }
        });
    for (auto& t : threads) t.join();
    for (auto& e : errors)
        if (e) std::rethrow_exception(e);
}

// final_binary_refresh_complex_block_output_bits_inplace (:406-490): mask,
// pack slots / n bits per ciphertext (bit first + s rotated by s n), refresh
// the packed real and imaginary lanes with one complex-packed bootstrap,
// recombine them, rotate each bit back and mask again.  The packs are
// independent (the reference runs them on OpenMP threads): they run over
// the main context and `extra` as the S-boxes do.
void final_binary_refresh(CKKSContext& context, MyVector<CKKSCiphertext>& bits, size_t active,
                          const std::vector<CKKSContext*>& extra, bool batched) {
    if (bits.size() != kBlockBits)
        throw std::invalid_argument("ckks_aes_walsh: expected 128 output bits for final refresh");
    const auto& P = context.getParams();
    const size_t stride = active;
    const size_t per_batch = stride == 0 ? 0 : P.getSlots() / stride;
    if (stride == 0 || per_batch == 0 || stride * per_batch > P.getSlots())
        throw std::invalid_argument("ckks_aes_walsh: invalid final refresh packing layout");
    ckks_aes_mask_final_refresh_slots_inplace(context, bits, stride, P.getBootstrapS2CDepth());
    const size_t batches = (kBlockBits + per_batch - 1) / per_batch;
    if (batched) {
        // the packs as one batch: all their rotations in one rotate batch,
        // the packed ciphertexts through one batched bootstrap
        MyVector<CKKSCiphertext> terms;
        terms.reserve(kBlockBits);
        for (auto& b : bits) terms.push_back(std::move(b));
        std::vector<CKKSCiphertext*> to_rotate;
        std::vector<int64_t> rot;
        for (size_t i = 0; i < kBlockBits; ++i)
            if (i % per_batch != 0) {
                to_rotate.push_back(&terms[i]);
                rot.push_back((int64_t)((i % per_batch) * stride));
            }
        rotate_batch(context, to_rotate, rot);
        MyVector<CKKSCiphertext> packs;
        for (size_t batch = 0; batch < batches; ++batch) {
            const size_t first = batch * per_batch, count = std::min(per_batch, kBlockBits - first);
            CKKSCiphertext packed = std::move(terms[first]);
            for (size_t s = 1; s < count; ++s) ckks_add_inplace(context, packed, terms[first + s]);
            packs.push_back(std::move(packed));
        }
        terms.clear();
        auto lanes = ckks_batch_bootstrap_evalmod_from_complex_packed(context, CKKSCiphertext::stack(std::move(packs)));
        auto imag_i = ckks_mul(context, lanes.second, Complex(0.0, 1.0), 1.0);
        imag_i.setScale(lanes.second.getScale());
        auto combined = std::move(lanes.first);
        ckks_add_inplace(context, combined, imag_i);
        auto per_pack = combined.unstack();
        MyVector<CKKSCiphertext> refreshed;
        refreshed.reserve(kBlockBits);
        to_rotate.clear();
        rot.clear();
        for (size_t batch = 0; batch < batches; ++batch) {
            const size_t first = batch * per_batch, count = std::min(per_batch, kBlockBits - first);
            for (size_t s = 0; s < count; ++s)
                refreshed.push_back(s + 1 == count ? std::move(per_pack[batch]) : per_pack[batch].clone());
        }
        for (size_t i = 0; i < kBlockBits; ++i)
            if (i % per_batch != 0) {
                to_rotate.push_back(&refreshed[i]);
                rot.push_back(-(int64_t)((i % per_batch) * stride));
            }
        rotate_batch(context, to_rotate, rot);
        bits = std::move(refreshed);
        ckks_aes_mask_final_refresh_slots_inplace(context, bits, stride, /*minimum_level_to_keep=*/0);
        return;
    }
    std::vector<CKKSContext*> ctxs{&context};
    ctxs.insert(ctxs.end(), extra.begin(), extra.end());
    std::vector<MyVector<CKKSCiphertext>> done(batches);
    run_on_contexts(context, ctxs, batches, [&](size_t batch, CKKSContext& wc) {
        const size_t first = batch * per_batch, count = std::min(per_batch, kBlockBits - first);
        // pack_final_refresh_lane (:347-367): zero_like + each rotated term
        MyVector<CKKSCiphertext> terms;
        terms.reserve(count);
        std::vector<CKKSCiphertext*> to_rotate;
        std::vector<int64_t> rot;
        for (size_t s = 0; s < count; ++s) terms.push_back(copy_to_context(wc, bits[first + s]));
        for (size_t s = 1; s < count; ++s) {
            to_rotate.push_back(&terms[s]);
            rot.push_back((int64_t)(s * stride));
        }
        rotate_batch(wc, to_rotate, rot);
        auto packed = std::move(terms[0]);
        for (size_t s = 1; s < count; ++s) ckks_add_inplace(wc, packed, terms[s]);
        auto lanes = ckks_batch_bootstrap_evalmod_from_complex_packed(wc, std::move(packed));
        // combine_real_imag_for_final_refresh (:392-404)
        auto imag_i = ckks_mul(wc, lanes.second, Complex(0.0, 1.0), 1.0);
        imag_i.setScale(lanes.second.getScale());
        auto combined = std::move(lanes.first);
        ckks_add_inplace(wc, combined, imag_i);
        // unpack_final_refresh_lane_inplace (:369-390)
        auto& out = done[batch];
        to_rotate.clear();
        rot.clear();
        for (size_t s = 0; s < count; ++s) out.push_back(combined.clone());
        for (size_t s = 1; s < count; ++s) {
            to_rotate.push_back(&out[s]);
            rot.push_back(-(int64_t)(s * stride));
        }
        rotate_batch(wc, to_rotate, rot);
    });
    MyVector<CKKSCiphertext> refreshed;
    refreshed.reserve(kBlockBits);
    for (auto& batch : done)
        for (auto& ct : batch)
            refreshed.push_back(&ct.getContext() == &context ? std::move(ct) : copy_to_context(context, ct));
    if (ctxs.size() > 1) {
        // the copies finish before the workers' buffers return to their pools
        mi_check(mi_stream_synchronize(context.handle(), context.stream()), "ckks_aes_walsh");
        done.clear();
    }
    bits = std::move(refreshed);
    ckks_aes_mask_final_refresh_slots_inplace(context, bits, stride, /*minimum_level_to_keep=*/0);
}

// shift_rows_bits_inplace (:723-739): new byte 4c + r <- old byte 4((c + r) % 4) + r
void shift_rows(MyVector<CKKSCiphertext>& state) {
    MyVector<CKKSCiphertext> old = std::move(state);
    state.clear();
    state.reserve(kBlockBits);
    for (size_t c = 0; c < 4; ++c)
        for (size_t r = 0; r < 4; ++r) {
            const size_t src = 4 * ((c + r) % 4) + r;
            for (size_t k = 0; k < 8; ++k) state.push_back(std::move(old[src * 8 + k]));
        }
}

// fmix_column_byte (:611-641): xtime(b0 + b1) + b1 + b2 + b3, where
// xtime(x) = (x7, x0 + x7, x1, x2 + x7, x3 + x7, x4, x5, x6).
CKKSByte mix_column_byte(const CKKSContext& context, const MyVector<CKKSCiphertext>& st, size_t i0, size_t i1,
                         size_t i2, size_t i3) {
    CKKSByte sum;
    sum.reserve(8);
    for (size_t k = 0; k < 8; ++k) sum.push_back(add_bits(context, st[8 * i0 + k], st[8 * i1 + k]));
    CKKSByte out;
    out.reserve(8);
    out.push_back(sum[7].clone());
    out.push_back(add_bits(context, sum[0], sum[7]));
    out.push_back(sum[1].clone());
    out.push_back(add_bits(context, sum[2], sum[7]));
    out.push_back(add_bits(context, sum[3], sum[7]));
    out.push_back(sum[4].clone());
    out.push_back(sum[5].clone());
    out.push_back(sum[6].clone());
    for (size_t idx : {i1, i2, i3})
        for (size_t k = 0; k < 8; ++k) add_term_ref(context, out[k], st[8 * idx + k]);
    return out;
}

// fmix_columns_bits_inplace (:741-783)
void mix_columns(const CKKSContext& context, MyVector<CKKSCiphertext>& state) {
    std::array<CKKSByte, kBlockBytes> bytes;
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

// add_round_key_lazy_inplace (:785-817)
void add_round_key(const CKKSContext& context, MyVector<CKKSCiphertext>& state, const MyVector<CKKSCiphertext>& keys,
                   size_t round) {
    for (size_t bit = 0; bit < kBlockBits; ++bit) add_term_ref(context, state[bit], keys[round * kBlockBits + bit]);
}

// A copy of `src` owned by `dst` (its memory and stream).  The copy is
// ordered on dst's stream only: src's stream must have finished writing it.
CKKSCiphertext copy_to_context(const CKKSContext& dst, const CKKSCiphertext& src) {
    if (&src.getContext() == &dst) return src.clone();
    CKKSCiphertext out(dst, src.getNumPolys(), src.getScale(), src.getLevel(), src.getSecretOwner(),
                       src.getMessageEncodingState());
    std::vector<uint64_t*> to(src.getNumPolys());
    std::vector<const uint64_t*> from(src.getNumPolys());
    for (size_t k = 0; k < to.size(); ++k) {
        to[k] = out.polyData(k);
        from[k] = src.polyData(k);
    }
    if (!to.empty())
        mi_check(mi_poly_list_op(dst.handle(), MI_LIST_COPY, (int)src.getLevel(), (int)to.size(), to.data(),
                                 from.data(), nullptr, dst.stream()),
                 "ckks_aes_walsh");
    return out;
}

// The worker contexts must hold the same keys as the main one: the same
// parameters and seed (hence the same secret and switching keys) and the same
// EvalMod polynomial, so that every S-box result is identical whichever
// context evaluates it.
void check_workers(const CKKSContext& context, const std::vector<CKKSContext*>& workers) {
    const auto& P = context.getParams();
    for (const CKKSContext* w : workers) {
        if (!w) throw std::invalid_argument("ckks_aes_walsh: null S-box context");
        const auto& Q = w->getParams();
        if (Q.getN() != P.getN() || Q.getSeed() != P.getSeed() || Q.getModuli() != P.getModuli() ||
            Q.getPModuli() != P.getPModuli() || Q.getDnum() != P.getDnum() ||
            Q.getSecretKeyHammingWeight() != P.getSecretKeyHammingWeight() ||
            Q.getAesWalshBlockCount() != P.getAesWalshBlockCount() ||
            Q.getAesWalshPackedSegmentCount() != P.getAesWalshPackedSegmentCount() ||
            w->getEvalModPolynomial() != context.getEvalModPolynomial())
            throw std::invalid_argument("ckks_aes_walsh: S-box contexts must share the main context's parameters and keys");
    }
}

// The factor aes_scale_up_to multiplies by, or 0 where it would throw or
// the factor exceeds 32 bits (the per-operation path then runs and raises
// the reference's error where there is one).
uint64_t scale_up_factor(long double from, long double to) {
    int64_t factor;
    if (aes_scale_up_factor(from, to, &factor) || factor < 1 || factor > 4294967295LL) return 0;
    return (uint64_t)factor;
}

// ShiftRows has run; MixColumns (unless the last round) and AddRoundKey as
// one linear map: every new bit is a sum of S-box output bits with small
// multiplicities plus its round-key bit times the key's scale-up factor,
// derived by running the per-operation sequence (mix_column_byte, add_term)
// on coefficient vectors; one mi_linear_combine evaluates all 128 bits.
// Returns false (state untouched) where the per-operation path must run:
// bits of differing level, scale or degree, or a scale-up the reference
// would refuse.
bool fused_linear_layer(const CKKSContext& context, MyVector<CKKSCiphertext>& state,
                        const MyVector<CKKSCiphertext>& keys, size_t round, bool mix) {
    const CKKSCiphertext& s0 = state[0];
    // same level, degree, scale, secret owner and encoding state for every
    // state and key bit (the additions of the per-operation path would
    // otherwise refuse mismatched metadata)
    auto same_meta = [&](const CKKSCiphertext& ct, const CKKSCiphertext& ref) {
        return ct.getLevel() == ref.getLevel() && ct.getNumPolys() == 2 && ct.getScale() == ref.getScale() &&
               ct.getBatchSize() == 1 && ct.getSecretOwner() == s0.getSecretOwner() &&
               ct.getMessageEncodingState() == s0.getMessageEncodingState();
    };
    for (const auto& ct : state)
        if (!same_meta(ct, s0)) return false;
    const CKKSCiphertext& k0 = keys[round * kBlockBits];
    for (size_t b = 0; b < kBlockBits; ++b)
        if (!same_meta(keys[round * kBlockBits + b], k0)) return false;
    using Vec = std::array<uint32_t, 2 * kBlockBits>;  // [S-box outputs after ShiftRows | round-key bits]
    std::array<Vec, kBlockBits> v{};
    for (size_t b = 0; b < kBlockBits; ++b) v[b][b] = 1;
    auto add = [](Vec a, const Vec& b) {
        for (size_t i = 0; i < a.size(); ++i) a[i] += b[i];
        return a;
    };
    if (mix) {
        std::array<Vec, kBlockBits> m{};
        for (size_t c = 0; c < 4; ++c)
            for (size_t o = 0; o < 4; ++o) {
                const size_t i0 = 4 * c + o, i1 = 4 * c + (o + 1) % 4, i2 = 4 * c + (o + 2) % 4,
                             i3 = 4 * c + (o + 3) % 4;
                std::array<Vec, 8> sum, out;
                for (size_t k = 0; k < 8; ++k) sum[k] = add(v[8 * i0 + k], v[8 * i1 + k]);
                out = {sum[7], add(sum[0], sum[7]), sum[1], add(sum[2], sum[7]), add(sum[3], sum[7]),
                       sum[4], sum[5], sum[6]};
                for (size_t k = 0; k < 8; ++k) m[8 * i0 + k] = add(add(add(out[k], v[8 * i1 + k]), v[8 * i2 + k]), v[8 * i3 + k]);
            }
        v = m;
    }
    const long double target = std::max(s0.getScale(), k0.getScale());
    const uint64_t fs = scale_up_factor(s0.getScale(), target), fk = scale_up_factor(k0.getScale(), target);
    if (!fs || !fk) return false;
    const size_t level = std::min(s0.getLevel(), k0.getLevel());
    int nterms = 0;
    for (size_t b = 0; b < kBlockBits; ++b) {
        v[b][kBlockBits + b] += 1;
        int n = 0;
        for (size_t i = 0; i < v[b].size(); ++i) {
            const uint64_t c = (uint64_t)v[b][i] * (i < kBlockBits ? fs : fk);
            if (c > 0xffffffffull) return false;
            n += v[b][i] != 0;
        }
        nterms = std::max(nterms, n);
    }
    if (nterms > 16) return false;
    MyVector<CKKSCiphertext> out;
    out.reserve(kBlockBits);
    std::vector<uint64_t*> dst;
    std::vector<const uint64_t*> src;
    std::vector<uint32_t> coef;
    for (size_t b = 0; b < kBlockBits; ++b) {
        out.emplace_back(context, 2, target, level, s0.getSecretOwner(), s0.getMessageEncodingState());
        for (size_t p = 0; p < 2; ++p) {
            dst.push_back(out.back().polyData(p));
            int n = 0;
            for (size_t i = 0; i < v[b].size(); ++i) {
                if (!v[b][i]) continue;
                const CKKSCiphertext& term = i < kBlockBits ? state[i] : keys[round * kBlockBits + (i - kBlockBits)];
                src.push_back(term.polyData(p));
                coef.push_back((uint32_t)((uint64_t)v[b][i] * (i < kBlockBits ? fs : fk)));
                ++n;
            }
            for (; n < nterms; ++n) {
                src.push_back(src.back());
                coef.push_back(0);
            }
        }
    }
    mi_check(mi_linear_combine(context.handle(), (int)level, (int)dst.size(), dst.data(), nterms, src.data(),
                               coef.data(), context.stream()),
             "ckks_aes_walsh");
    state = std::move(out);
    return true;
}

// sub_bytes_walsh_complex_blocks_mask_packed (:659-721) as one batch: input
// bit k of all 16 bytes forms one batched ciphertext (the masked state bits
// share level and scale), the mask-packed S-box runs once over the 8
// batches, so every step of the 16 S-boxes (parity packing, bootstrap,
// unpacking, product sums, conjugations) is one launch sequence for all of
// them; member b of output bit k is byte b's bit k, limb for limb the
// one-S-box result.
MyVector<CKKSCiphertext> sub_bytes_batched(CKKSContext& context, MyVector<CKKSCiphertext> state, size_t pack) {
    MyVector<CKKSCiphertext> input;
    input.reserve(8);
    for (size_t k = 0; k < 8; ++k) {
        MyVector<CKKSCiphertext> members;
        members.reserve(kBlockBytes);
        for (size_t byte = 0; byte < kBlockBytes; ++byte) members.push_back(std::move(state[8 * byte + k]));
        input.push_back(CKKSCiphertext::stack(std::move(members)));
    }
    auto bits = ckks_walsh_sbox_complex_blocks_mask_packed(context, std::move(input), pack);
    if (bits.size() != 8) throw std::runtime_error("ckks_aes_walsh: complex-block Walsh S-box returned wrong size");
    std::array<MyVector<CKKSCiphertext>, 8> split;
    for (size_t k = 0; k < 8; ++k) {
        split[k] = bits[k].unstack();
        if (split[k].size() != kBlockBytes)
            throw std::runtime_error("ckks_aes_walsh: batched S-box returned a wrong batch size");
    }
    MyVector<CKKSCiphertext> out;
    out.reserve(kBlockBits);
    for (size_t byte = 0; byte < kBlockBytes; ++byte)
        for (size_t k = 0; k < 8; ++k) out.push_back(std::move(split[k][byte]));
    return out;
}

// sub_bytes_walsh_complex_blocks_mask_packed (:659-721): the 16 S-boxes.
// The reference runs them in parallel over OpenMP threads; here each host
// thread drives one context (its own stream and memory pool), byte b going
// to context b mod the number of contexts, so the S-boxes' small kernels
// overlap on the device.  Without extra contexts they run one after another
// on the main stream.
MyVector<CKKSCiphertext> sub_bytes(CKKSContext& context, const MyVector<CKKSCiphertext>& state, size_t pack,
                                   const std::vector<CKKSContext*>& extra) {
    std::vector<CKKSContext*> ctxs{&context};
    ctxs.insert(ctxs.end(), extra.begin(), extra.end());
    std::array<MyVector<CKKSCiphertext>, kBlockBytes> bytes;
    auto run = [&](size_t w) {
        CKKSContext& wc = *ctxs[w];
        for (size_t byte = w; byte < kBlockBytes; byte += ctxs.size()) {
            MyVector<CKKSCiphertext> input;
            input.reserve(8);
            for (size_t k = 0; k < 8; ++k) input.push_back(copy_to_context(wc, state[8 * byte + k]));
            bytes[byte] = ckks_walsh_sbox_complex_blocks_mask_packed(wc, std::move(input), pack);
            if (bytes[byte].size() != 8)
                throw std::runtime_error("ckks_aes_walsh: complex-block Walsh S-box returned wrong size");
        }
        mi_check(mi_stream_synchronize(wc.handle(), wc.stream()), "ckks_aes_walsh");
    };
    if (ctxs.size() == 1) {
        run(0);
    } else {
        // the state's producers on the main stream finish before other streams read it
        mi_check(mi_stream_synchronize(context.handle(), context.stream()), "ckks_aes_walsh");
        std::vector<std::exception_ptr> errors(ctxs.size());
        std::vector<std::thread> threads;
        for (size_t w = 0; w < ctxs.size(); ++w)
            threads.emplace_back([&, w] {
                try {
                    run(w);
                } catch (...) {
                    errors[w] = std::current_exception();
                }
            });
        for (auto& t : threads) t.join();
        for (auto& e : errors)
            if (e) std::rethrow_exception(e);
    }
    MyVector<CKKSCiphertext> out;
    out.reserve(kBlockBits);
    for (auto& byte : bytes)
        for (auto& bit : byte)
            out.push_back(&bit.getContext() == &context ? std::move(bit) : copy_to_context(context, bit));
    if (ctxs.size() > 1) {
        // the copies finish before the workers' buffers return to their pools
        mi_check(mi_stream_synchronize(context.handle(), context.stream()), "ckks_aes_walsh");
        for (auto& byte : bytes) byte.clear();
    }
    return out;
}

}  // namespace

MyVector<CKKSCiphertext> ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(
    const CKKSContext& context, const std::array<uint8_t, 176>& rk, size_t active) {
    const auto& P = context.getParams();
    packing_block_count(context, active);
    const size_t level = state_entry_level(context);
    if (level > P.getMaxLevel())
        throw std::invalid_argument("ckks_aes_walsh: selection key level exceeds parameter chain");
    MyVector<CKKSCiphertext> out;
    out.reserve(kExpandedKeyBits);
    for (size_t bit = 0; bit < kExpandedKeyBits; ++bit) {
        const double v = ((rk[bit / 8] >> (bit % 8)) & 1) ? 1.0 : 0.0;
        MyVector<Complex> slots(P.getSlots(), Complex(0.0, 0.0));
        for (size_t s = 0; s < active; ++s) slots[s] = Complex(v, v);
        const double scale = (bit < kBlockBits || bit >= kRounds * kBlockBits) ? P.getScale() : P.getScale() / 4.0;
        out.push_back(ckks_encrypt(context, ckks_encode(context, slots, scale, level, /*eval=*/true), /*eval=*/true));
    }
    return out;
}

MyVector<CKKSCiphertext> ckks_aes_ctr_walsh_sbox_complex_blocks(CKKSContext& context,
                                                                const MyVector<CKKSCiphertext>& keys,
                                                                std::span<const Block> counters,
                                                                std::span<const Block> ciphertexts,
                                                                const CKKSAesCtrWalshSboxOptions& options) {
    const std::vector<CKKSContext*> no_workers;
    const auto& workers =
        options.parallelize_sbox_pairs && !options.batch_sboxes ? options.sbox_contexts : no_workers;
    check_workers(context, workers);
    // validate_complex_block_inputs (:905-934)
    if (keys.size() != kExpandedKeyBits)
        throw std::invalid_argument("ckks_aes_walsh: expanded round key must have 1408 bits");
    if (counters.empty() || counters.size() != ciphertexts.size() || counters.size() % 2 != 0)
        throw std::invalid_argument(
            "ckks_aes_walsh: complex-block counter/ciphertext spans must be non-empty, equal, and even");
    const size_t active = counters.size() / 2;
    const size_t pack = packing_block_count(context, active);
    const size_t key_level = state_entry_level(context);
    if (key_level > context.getParams().getMaxLevel())
        throw std::invalid_argument("ckks_aes_walsh: Walsh path requires a selection limb above S2C input level");
    for (const auto& ct : keys)
        if (ct.getN() != context.getParams().getN() || ct.getNumPolys() != 2 || ct.getLevel() != key_level)
            throw std::invalid_argument(
                "ckks_aes_walsh: complex-block path received a round key bit at the wrong level");
    // Round-key bits encrypted under another context of the same keys (the
    // key owner of a lane) are copied onto this context's stream and pool
    // first: every operation below then stays on one stream.
    MyVector<CKKSCiphertext> local_keys;
    if (&keys.front().getContext() != &context) {
        mi_check(mi_stream_synchronize(keys.front().getContext().handle(), keys.front().getContext().stream()),
                 "ckks_aes_walsh");
        local_keys.reserve(keys.size());
        for (const auto& k : keys) local_keys.push_back(copy_to_context(context, k));
    }
    const MyVector<CKKSCiphertext>& round_keys = local_keys.empty() ? keys : local_keys;
    // make_initial_counter_complex_block_state (:534-559): round key 0 + counter
    MyVector<CKKSCiphertext> state;
    state.reserve(kBlockBits);
    for (size_t bit = 0; bit < kBlockBits; ++bit) state.push_back(round_keys[bit].clone());
    add_public_bits(context, state, counters, options.fuse_linear_layer);
    // evaluate_aes128_rounds_walsh_complex_blocks_impl (:819-903)
    for (size_t round = 1; round <= kRounds; ++round) {
        ckks_aes_mask_inactive_slots_inplace(context, state, pack);
        if (options.batch_sboxes)
            state = sub_bytes_batched(context, std::move(state), pack);
        else
            state = sub_bytes(context, state, pack, workers);
        shift_rows(state);
        if (!options.fuse_linear_layer || !fused_linear_layer(context, state, round_keys, round, round != kRounds)) {
            if (round != kRounds) mix_columns(context, state);
            add_round_key(context, state, round_keys, round);
        }
    }
    // CTR: plaintext = ciphertext + AES(counter) (add_public_complex_block_bits_lazy_inplace, :492-532)
    add_public_bits(context, state, ciphertexts, options.fuse_linear_layer);
    final_binary_refresh(context, state, active, workers, options.batch_sboxes);
    return state;
}
