#include <thread>
#include <exception>
#include <algorithm>
// Encode / encrypt / decrypt (ckks_encoding.cpp:75-580, fft.cpp:82-117).
//
// Slot encoding runs on the device (mi_encode_slots: the special FFT in double
// with the reference's operation order, llrint of value * scale / N, signed
// residues, NTT).  Decoding's floating-point parts stay on the host (centred
// CRT of the first limbs whose product fits 128 bits, division by the scale
// in long double, the inverse FFT), as does encode_slots_custom_basis for
// callers that want host residues; built with -ffp-contract=off so no
// product is fused.  The ring parts run on the device: NTTs, the mask product
// a*s and the sums.
#include <bit>
#include <cmath>

#include <map>

#include "ckks.hpp"
#include "ckks_rng.hpp"

namespace {

typedef unsigned __int128 u128;

void fft_dif(MyVector<Complex>& x, const MyVector<Complex>& tw, size_t N) {
    for (size_t m = N >> 1, k = 1; m > 0; m >>= 1, k <<= 1)
        for (size_t i = 0; i < m; ++i) {
            const Complex w = tw[m + i];
            Complex* x0 = x.data() + ((i << 1) * k);
            Complex* x1 = x0 + k;
            for (size_t j = 0; j < k; ++j) {
                const Complex u = x0[j], v = x1[j];
                x0[j] = u + v;
                x1[j] = (u - v) * w;
            }
        }
}

void fft_dit(MyVector<Complex>& x, const MyVector<Complex>& tw, size_t N) {
    for (size_t m = 1, k = N >> 1; m < N; m <<= 1, k >>= 1)
        for (size_t i = 0; i < m; ++i) {
            const Complex w = tw[m + i];
            Complex* x0 = x.data() + ((i << 1) * k);
            Complex* x1 = x0 + k;
            for (size_t j = 0; j < k; ++j) {
                const Complex t = x0[j], u = x1[j] * w;
                x0[j] = t + u;
                x1[j] = t - u;
            }
        }
}

uint64_t signed_to_residue(int64_t v, uint64_t q) {
    const int64_t qi = (int64_t)q;
    const int64_t r = v % qi;
    return (uint64_t)(r + (qi & (r >> 63)));
}

size_t reverse_bits(size_t v, size_t width) {
    size_t r = 0;
    for (size_t i = 0; i < width; ++i) {
        r = (r << 1) | (v & 1);
        v >>= 1;
    }
    return r;
}

void check_level(const CKKSContext& c, size_t level, const char* fn) {
    if (level > c.getParams().getMaxLevel()) throw std::invalid_argument(std::string(fn) + ": invalid level");
}

DeviceBuffer upload_plaintext(const CKKSContext& c, const MyVector<uint64_t>& h, size_t limbs, bool eval) {
    const size_t N = c.getParams().getN();
    DeviceBuffer d(c.handle(), limbs * N);
    mi_check(mi_memcpy(c.handle(), d.data(), h.data(), h.size() * 8, MI_COPY_H2D, c.stream()), "ckks_encode");
    if (eval) mi_check(mi_ntt_forward(c.handle(), d.data(), 1, (int)limbs, (int64_t)(limbs * N), 0, c.stream()), "ckks_encode");
    return d;
}

}  // namespace

MyVector<uint64_t> encode_slots_custom_basis(size_t N, const MyVector<Complex>& values, double scale,
                                             const MyVector<uint64_t>& moduli,
                                             const MyVector<Complex>& twiddle_fft,
                                             const MyVector<size_t>& perm_table) {
    const size_t half = N >> 1;
    if (values.size() != half)
        throw std::invalid_argument("encode_slots_custom_basis: expected exactly N/2 slot values");
    if (!(scale > 0.0)) throw std::invalid_argument("encode_slots_custom_basis: scale must be positive");
    if (moduli.empty()) throw std::invalid_argument("encode_slots_custom_basis: invalid modulus basis");
    MyVector<Complex> ev(N);
    for (size_t i = 0; i < half; ++i) {
        const size_t d = perm_table[i];
        ev[d] = values[i];
        ev[N - 1 - d] = std::conj(values[i]);
    }
    fft_dif(ev, twiddle_fft, N);
    MyVector<uint64_t> out(moduli.size() * N);
    const double k = scale / (double)N;
    for (size_t l = 0; l < moduli.size(); ++l)
        for (size_t i = 0; i < N; ++i) out[l * N + i] = signed_to_residue(std::llrint(ev[i].real() * k), moduli[l]);
    return out;
}

// ckks_encode (ckks_encoding.cpp:332-366): the encoding itself runs on the
// device (mi_encode_slots: placement, FFT, rounding, residues, NTT).
CKKSEncoding ckks_encode(const CKKSContext& c, const MyVector<Complex>& values, double scale, size_t level,
                         bool eval) {
    if (!(scale > 0.0)) throw std::invalid_argument("ckks_encode: scale must be positive");
    check_level(c, level, "ckks_encode");
    const size_t N = c.getParams().getN();
    if (values.size() != N / 2)
        throw std::invalid_argument("encode_slots_custom_basis: expected exactly N/2 slot values");
    const size_t words = (level + 1) * N;
    DeviceBuffer d(c.handle(), words);
    mi_check(mi_encode_slots(c.handle(), reinterpret_cast<const double*>(values.data()), 1, scale, (int)(level + 1),
                             eval ? 1 : 0, d.data(), (int64_t)words, c.stream()),
             "ckks_encode");
    return CKKSEncoding(c, std::move(d), scale, level, false);
}

CKKSEncoding ckks_encode(const CKKSContext& c, const MyVector<Complex>& values, size_t level, bool eval) {
    return ckks_encode(c, values, c.getParams().getScale(), level, eval);
}

CKKSEncoding ckks_encode_coefficients(const CKKSContext& c, const MyVector<Complex>& values, double scale,
                                      size_t level, bool eval) {
    const size_t N = c.getParams().getN(), slots = N >> 1;
    if (values.size() != slots)
        throw std::invalid_argument("ckks_encode_coefficients: expected exactly N/2 slot values");
    if (!(scale > 0.0)) throw std::invalid_argument("ckks_encode_coefficients: scale must be positive");
    check_level(c, level, "ckks_encode_coefficients");
    const auto& q = c.getParams().getModuli();
    const size_t log_slots = (size_t)std::countr_zero(slots);
    MyVector<uint64_t> h((level + 1) * N, 0);
    for (size_t l = 0; l <= level; ++l)
        for (size_t i = 0; i < slots; ++i) {
            const size_t ci = reverse_bits(i, log_slots);
            h[l * N + ci] = signed_to_residue(std::llrint(values[i].real() * scale), q[l]);
            h[l * N + slots + ci] = signed_to_residue(std::llrint(values[i].imag() * scale), q[l]);
        }
    return CKKSEncoding(c, upload_plaintext(c, h, level + 1, eval), scale, level, false);
}

CKKSCiphertext ckks_encrypt(const CKKSContext& c, const MyVector<Complex>& values, size_t level, bool eval,
                            CKKSMessageEncodingState state) {
    CKKSEncoding e = state == CKKSMessageEncodingState::Coefficients
                         ? ckks_encode_coefficients(c, values, c.getParams().getScale(), level, eval)
                         : ckks_encode(c, values, level, eval);
    return ckks_encrypt(c, e, eval, state);
}

// ckks_encrypt_custom_eval (ckks_encoding.cpp:464-510): a uniform in the
// evaluation domain, e one Gaussian integer per coefficient reduced per limb,
// b = m + e - a*s.
CKKSCiphertext ckks_encrypt(const CKKSContext& c, const CKKSEncoding& encd, bool eval,
                            CKKSMessageEncodingState state) {
    const auto& P = c.getParams();
    const size_t N = P.getN(), level = encd.getLevel(), L = level + 1;
    if (encd.getN() != N) throw std::invalid_argument("ckks_encrypt(ptxt): plaintext/context size mismatch");
    check_level(c, level, "ckks_encrypt(ptxt)");
    mi_context* mi = c.handle();
    const int64_t W = (int64_t)(L * N);
    CKKSCiphertext ct(c, 2, encd.getScale(), level, CKKSSecretOwner::Dense, state);
    uint64_t* b = ct.polyData(0);
    uint64_t* a = ct.polyData(1);
    mi_check(mi_memcpy(mi, b, encd.data(), L * N * 8, MI_COPY_D2D, c.stream()), "ckks_encrypt(ptxt)");
    if (encd.isMontgomeryForm()) {
        // x * 1 with the plaintext in Montgomery form gives the normal value
        MyVector<uint64_t> ones(L * N, 1);
        DeviceBuffer one(mi, L * N);
        mi_check(mi_memcpy(mi, one.data(), ones.data(), ones.size() * 8, MI_COPY_H2D, c.stream()), "ckks_encrypt(ptxt)");
        mi_check(mi_poly_op(mi, MI_OP_MUL_MONT, one.data(), b, (int)L, 0, 1, W, c.stream()), "ckks_encrypt(ptxt)");
        mi_check(mi_memcpy(mi, b, one.data(), L * N * 8, MI_COPY_D2D, c.stream()), "ckks_encrypt(ptxt)");
        mi_check(mi_stream_synchronize(mi, c.stream()), "ckks_encrypt(ptxt)");
    }
    if (!eval) mi_check(mi_ntt_forward(mi, b, 1, (int)L, W, 0, c.stream()), "ckks_encrypt(ptxt)");
    // a uniform (evaluation domain) and e Gaussian, drawn on the device
    const uint64_t seed = c.nextRandomnessSeed();
    mi_check(mi_sample_uniform(mi, a, (int)L, 0, seed, 0, c.stream()), "ckks_encrypt(ptxt)");
    DeviceBuffer tmp(mi, L * N);
    mi_check(mi_sample_gaussian(mi, tmp.data(), (int)L, P.getSigma(), seed, 1, c.stream()), "ckks_encrypt(ptxt)");
    mi_check(mi_ntt_forward(mi, tmp.data(), 1, (int)L, W, 0, c.stream()), "ckks_encrypt(ptxt)");
    mi_check(mi_poly_op(mi, MI_OP_ADD, b, tmp.data(), (int)L, 0, 1, W, c.stream()), "ckks_encrypt(ptxt)");
    mi_check(mi_memcpy(mi, tmp.data(), a, L * N * 8, MI_COPY_D2D, c.stream()), "ckks_encrypt(ptxt)");
    mi_check(mi_poly_op(mi, MI_OP_MUL, tmp.data(), c.secretKeyDevice(), (int)L, 0, 1, W, c.stream()), "ckks_encrypt(ptxt)");
    mi_check(mi_poly_op(mi, MI_OP_SUB, b, tmp.data(), (int)L, 0, 1, W, c.stream()), "ckks_encrypt(ptxt)");
    mi_check(mi_stream_synchronize(mi, c.stream()), "ckks_encrypt(ptxt)");
    return ct;
}

// crt_interpolate + the scale division + the inverse special FFT
// (ckks_encoding.cpp:96-170, 240-329).
MyVector<Complex> ckks_decode_plaintext(const CKKSContext& c, const MyVector<uint64_t>& coeffs, size_t limb_count,
                                        long double scale, CKKSDecryptDomain domain) {
    const auto& q = c.getParams().getModuli();
    if (limb_count == 0 || limb_count > q.size()) throw std::invalid_argument("ckks_decrypt: invalid level");
    return decode_plaintext_custom_basis(c.getParams().getN(), coeffs, MyVector<uint64_t>(q.begin(), q.begin() + limb_count),
                                         scale, c.getTwiddleIfft(), c.getPermTable(), domain);
}

MyVector<Complex> decode_plaintext_custom_basis(size_t N, const MyVector<uint64_t>& coeffs,
                                                const MyVector<uint64_t>& q, long double scale,
                                                const MyVector<Complex>& twiddle_ifft,
                                                const MyVector<size_t>& perm, CKKSDecryptDomain domain) {
    const size_t half = N >> 1, limb_count = q.size();
    if (coeffs.size() < limb_count * N) throw std::invalid_argument("ckks_decrypt: plaintext size mismatch");
    MyVector<Complex> m(N);
    {
        const u128 max = ~(u128)0;
        size_t use = 0;
        u128 Pr = 1;
        for (size_t i = 0; i < limb_count; ++i) {
            if (Pr > max / q[i]) break;
            Pr *= q[i];
            ++use;
        }
        const u128 half_p = Pr / 2;
        std::vector<u128> phat(use);
        std::vector<uint64_t> phat_inv(use);
        for (size_t i = 0; i < use; ++i) {
            phat[i] = Pr / q[i];
            const uint64_t r = (uint64_t)(phat[i] % q[i]);
            // inverse by Fermat (q prime)
            uint64_t e = q[i] - 2, base = r, acc = 1;
            while (e) {
                if (e & 1) acc = (uint64_t)((u128)acc * base % q[i]);
                base = (uint64_t)((u128)base * base % q[i]);
                e >>= 1;
            }
            phat_inv[i] = acc;
        }
        for (size_t j = 0; j < N; ++j) {
            u128 acc = 0;
            for (size_t k = 0; k < use; ++k) {
                const u128 al = ((u128)coeffs[k * N + j] * phat_inv[k]) % q[k];
                const u128 term = al * phat[k];
                const u128 rem = Pr - acc;
                acc = term >= rem ? term - rem : acc + term;
            }
            const double v = acc > half_p ? -1.0 * (double)(Pr - acc) : (double)acc;
            m[j] = Complex(v, 0.0);
        }
    }
    for (size_t i = 0; i < N; ++i) m[i] = Complex((double)((long double)m[i].real() / scale), 0.0);
    if (domain == CKKSDecryptDomain::Coefficients) {
        const size_t log_slots = (size_t)std::countr_zero(half);
        MyVector<Complex> out(half);
        for (size_t i = 0; i < half; ++i) {
            const size_t ci = reverse_bits(i, log_slots);
            out[i] = Complex(m[ci].real(), m[half + ci].real());
        }
        return out;
    }
    fft_dit(m, twiddle_ifft, N);
    MyVector<Complex> slots(half);
    for (size_t i = 0; i < half; ++i) slots[i] = m[perm[i]];
    return slots;
}

// ckks_decrypt_with_secret_eval (ckks_encoding.cpp:96-170): m = b + a*s (+ c2*s^2)
// and the inverse NTT on the device, then the decode (CRT, scale division,
// inverse special FFT) on the device too (mi_decode_batch, equal to the host
// decode above bit for bit); only the slots come back to the host.
namespace {
// The device part of ckks_decrypt: b + a s (+ c2 s^2) and the inverse NTT.
DeviceBuffer decrypt_coefficients(const CKKSContext& c, const CKKSCiphertext& ct, bool use_low_secret);

// scale = mant 2^exp with the top bit of mant set (the x87 extended value)
void split_scale(long double scale, uint64_t& mant, int32_t& exp) {
    if (!(scale > 0.0L) || !std::isfinite(scale)) throw std::invalid_argument("ckks_decrypt: invalid ciphertext scale");
    int e = 0;
    const long double f = std::frexp(scale, &e);
    mant = (uint64_t)std::ldexp(f, 64);
    exp = e - 64;
}

// Decodes plaintexts of one limb count already on the device.
std::vector<MyVector<Complex>> decode_on_device(const CKKSContext& c, const std::vector<const uint64_t*>& coeffs,
                                               size_t limbs, const std::vector<long double>& scales,
                                               CKKSDecryptDomain domain) {
    const size_t n = coeffs.size(), N = c.getParams().getN(), half = N / 2;
    std::vector<uint64_t> mant(n);
    std::vector<int32_t> ex(n);
    for (size_t i = 0; i < n; ++i) split_scale(scales[i], mant[i], ex[i]);
    DeviceBuffer out(c.handle(), n * N);  // n x N/2 complex
    mi_check(mi_decode_batch(c.handle(), (int)n, coeffs.data(), (int)limbs, mant.data(), ex.data(),
                             domain == CKKSDecryptDomain::Coefficients ? 1 : 0, reinterpret_cast<double*>(out.data()),
                             c.stream()),
             "ckks_decrypt");
    std::vector<double> h(n * N);
    mi_check(mi_memcpy(c.handle(), h.data(), out.data(), h.size() * 8, MI_COPY_D2H, c.stream()), "ckks_decrypt");
    std::vector<MyVector<Complex>> res(n, MyVector<Complex>(half));
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < half; ++j) res[i][j] = Complex(h[i * N + 2 * j], h[i * N + 2 * j + 1]);
    return res;
}
}  // namespace

MyVector<Complex> ckks_decrypt(const CKKSContext& c, const CKKSCiphertext& ct, CKKSDecryptDomain domain,
                               bool use_low_secret) {
    DeviceBuffer t = decrypt_coefficients(c, ct, use_low_secret);
    return std::move(decode_on_device(c, {t.data()}, ct.getLevel() + 1, {ct.getScale()}, domain)[0]);
}

std::vector<MyVector<Complex>> ckks_decrypt_batch(const CKKSContext& c, const std::vector<const CKKSCiphertext*>& cts,
                                                  CKKSDecryptDomain domain) {
    std::vector<DeviceBuffer> coeffs;
    coeffs.reserve(cts.size());
    for (const CKKSCiphertext* ct : cts) {
        if (!ct) throw std::invalid_argument("ckks_decrypt: null ciphertext");
        coeffs.push_back(decrypt_coefficients(c, *ct, false));
    }
    // one decode launch sequence per limb count
    std::vector<MyVector<Complex>> out(cts.size());
    std::map<size_t, std::vector<size_t>> by_limbs;
    for (size_t i = 0; i < cts.size(); ++i) by_limbs[cts[i]->getLevel() + 1].push_back(i);
    for (const auto& [limbs, idx] : by_limbs) {
        std::vector<const uint64_t*> ptrs;
        std::vector<long double> scales;
        for (size_t i : idx) {
            ptrs.push_back(coeffs[i].data());
            scales.push_back(cts[i]->getScale());
        }
        auto dec = decode_on_device(c, ptrs, limbs, scales, domain);
        for (size_t k = 0; k < idx.size(); ++k) out[idx[k]] = std::move(dec[k]);
    }
    return out;
}

MyVector<uint64_t> ckks_decrypt_coefficients(const CKKSContext& c, const CKKSCiphertext& ct) {
    DeviceBuffer t = decrypt_coefficients(c, ct, false);
    MyVector<uint64_t> h((ct.getLevel() + 1) * c.getParams().getN());
    mi_check(mi_memcpy(c.handle(), h.data(), t.data(), h.size() * 8, MI_COPY_D2H, c.stream()), "ckks_decrypt");
    return h;
}

namespace {
DeviceBuffer decrypt_coefficients(const CKKSContext& c, const CKKSCiphertext& ct, bool use_low_secret) {
    if (use_low_secret || ct.getSecretOwner() == CKKSSecretOwner::BootstrapSparse)
        throw std::runtime_error("ckks_context: sparse bootstrap secret key is unavailable");
    const size_t polys = ct.getNumPolys();
    if (polys != 2 && polys != 3) throw std::invalid_argument("ckks_decrypt: expected 2- or 3-component ciphertext");
    const size_t N = c.getParams().getN(), level = ct.getLevel(), L = level + 1;
    check_level(c, level, "ckks_decrypt");
    mi_context* mi = c.handle();
    const int64_t W = (int64_t)(L * N);
    DeviceBuffer t(mi, L * N);
    mi_check(mi_memcpy(mi, t.data(), ct.polyData(1), L * N * 8, MI_COPY_D2D, c.stream()), "ckks_decrypt");
    mi_check(mi_poly_op(mi, MI_OP_MUL, t.data(), c.secretKeyDevice(), (int)L, 0, 1, W, c.stream()), "ckks_decrypt");
    mi_check(mi_poly_op(mi, MI_OP_ADD, t.data(), ct.polyData(0), (int)L, 0, 1, W, c.stream()), "ckks_decrypt");
    if (polys == 3) {
        DeviceBuffer s2(mi, L * N);
        mi_check(mi_memcpy(mi, s2.data(), c.secretKeyDevice(), L * N * 8, MI_COPY_D2D, c.stream()), "ckks_decrypt");
        mi_check(mi_poly_op(mi, MI_OP_MUL, s2.data(), c.secretKeyDevice(), (int)L, 0, 1, W, c.stream()), "ckks_decrypt");
        mi_check(mi_poly_op(mi, MI_OP_MUL, s2.data(), ct.polyData(2), (int)L, 0, 1, W, c.stream()), "ckks_decrypt");
        mi_check(mi_poly_op(mi, MI_OP_ADD, t.data(), s2.data(), (int)L, 0, 1, W, c.stream()), "ckks_decrypt");
        mi_check(mi_stream_synchronize(mi, c.stream()), "ckks_decrypt");
    }
    mi_check(mi_ntt_inverse(mi, t.data(), 1, (int)L, W, 0, c.stream()), "ckks_decrypt");
    return t;
}
}  // namespace
