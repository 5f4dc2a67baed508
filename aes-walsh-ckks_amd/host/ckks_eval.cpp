#include <cstring>
// Evaluator free functions over device-resident ciphertexts: metadata checks
// and scale bookkeeping of the reference (ckks_mul.cpp:27-288,
// ckks_rescale.cpp:42-148, ckks_addition.cpp:17-139, ckks_relin.cpp:8-44,
// ckks_rotation.cpp:34-96, ckks_conjugate.cpp:10-61) around the C-ABI kernels.
#include <cmath>

#include "ckks.hpp"

namespace {

void check_ct(const CKKSContext& c, const CKKSCiphertext& ct, const char* fn) {
    const auto& P = c.getParams();
    if (ct.getN() != P.getN()) throw std::invalid_argument(std::string(fn) + ": ciphertext/context size mismatch");
    if (ct.getLevel() > P.getMaxLevel()) throw std::invalid_argument(std::string(fn) + ": invalid level");
    if (&ct.getContext() != &c) throw std::invalid_argument(std::string(fn) + ": ciphertext belongs to another context");
}

uint64_t reduce_i64(int64_t v, uint64_t q) {
    const int64_t qi = (int64_t)q;
    const int64_t r = v % qi;
    return (uint64_t)(r < 0 ? r + qi : r);
}

// rounded_scaled_integer_mod_q (ckks_addition.cpp:17-31): round(v * scale)
// mod q in long double.
uint64_t rounded_scaled_integer_mod_q(int64_t value, long double scale, uint64_t q) {
    const long double scaled = (long double)value * scale;
    if (!std::isfinite(scaled))
        throw std::invalid_argument("ckks_add_integer_constant_inplace: non-finite scaled constant");
    long double r = std::fmod(std::round(scaled), (long double)q);
    if (r < 0.0L) r += (long double)q;
    uint64_t out = (uint64_t)r;
    if (out >= q) out %= q;
    return out;
}

}  // namespace

void ckks_mul_inplace(const CKKSContext& c, CKKSCiphertext& ct, const CKKSEncoding& pt) {
    check_ct(c, ct, "ckks_mul_inplace");
    if (pt.getN() != ct.getN() || pt.getLevel() != ct.getLevel())
        throw std::invalid_argument("ckks_mul_inplace: ciphertext/plaintext size mismatch");
    if (ct.getNumPolys() == 0) throw std::invalid_argument("ckks_mul_inplace: empty ciphertext");
    const size_t lv = ct.getLevel();
    std::vector<uint64_t*> polys = ct.allPolyData();
    mi_check(mi_mul_plain_rescale_list(c.handle(), (int)lv, (int)polys.size(), polys.data(), pt.data(),
                                       pt.isMontgomeryForm() ? 1 : 0, 0, c.stream()),
             "ckks_mul_inplace");
    ct.setScale(ct.getScale() * pt.getScale());
}

void ckks_mul_inplace(const CKKSContext& c, CKKSCiphertext& ct1, const CKKSCiphertext& ct2) {
    if (ct1.getN() != ct2.getN() || ct1.getLevel() != ct2.getLevel())
        throw std::invalid_argument("ckks_mul_inplace(ct,ct): ciphertext size mismatch");
    ckks_detail::mul_at_level(c, ct1, ct2);
}

void ckks_detail::mul_at_level(const CKKSContext& c, CKKSCiphertext& ct1, const CKKSCiphertext& ct2) {
    const size_t lv = ct1.getLevel();
    if (ct1.getN() != ct2.getN() || lv > ct2.getLevel())
        throw std::invalid_argument("ckks_mul_inplace(ct,ct): ciphertext size mismatch");
    check_ct(c, ct1, "ckks_mul_inplace(ct,ct)");
    if (ct1.getNumPolys() != 2 || ct2.getNumPolys() != 2)
        throw std::invalid_argument("ckks_mul_inplace(ct,ct): expected 2-component inputs");
    if (ct1.getSecretOwner() != ct2.getSecretOwner() ||
        ct1.getMessageEncodingState() != ct2.getMessageEncodingState())
        throw std::invalid_argument("ckks_mul_inplace(ct,ct): ciphertext metadata mismatch");
    if (ct1.getBatchSize() != ct2.getBatchSize())
        throw std::invalid_argument("ckks_mul_inplace(ct,ct): batch size mismatch");
    ct1.addPoly();
    if (ct1.getBatchSize() == 1) {
        mi_check(mi_tensor(c.handle(), (int)lv, ct1.polyData(0), ct1.polyData(1), ct2.polyData(0), ct2.polyData(1),
                           ct1.polyData(2), c.stream()),
                 "ckks_mul_inplace(ct,ct)");
    } else {
        const size_t B = ct1.getBatchSize();
        std::vector<uint64_t*> b1(B), a1(B), c2(B);
        std::vector<const uint64_t*> b2(B), a2(B);
        for (size_t m = 0; m < B; ++m) {
            b1[m] = ct1.polyData(0, m);
            a1[m] = ct1.polyData(1, m);
            c2[m] = ct1.polyData(2, m);
            b2[m] = ct2.polyData(0, m);
            a2[m] = ct2.polyData(1, m);
        }
        mi_check(mi_tensor_list(c.handle(), (int)lv, (int)B, b1.data(), a1.data(), b2.data(), a2.data(), c2.data(),
                                c.stream()),
                 "ckks_mul_inplace(ct,ct)");
    }
    ct1.setScale(ct1.getScale() * ct2.getScale());
}

void ckks_square_inplace(const CKKSContext& c, CKKSCiphertext& ct) {
    check_ct(c, ct, "ckks_square_inplace");
    if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_square_inplace: expected 2-component input");
    ct.addPoly();
    if (ct.getBatchSize() == 1) {
        mi_check(mi_square(c.handle(), (int)ct.getLevel(), ct.polyData(0), ct.polyData(1), ct.polyData(2),
                           c.stream()),
                 "ckks_square_inplace");
    } else {
        const size_t B = ct.getBatchSize();
        std::vector<uint64_t*> b(B), a(B), c2(B);
        for (size_t m = 0; m < B; ++m) {
            b[m] = ct.polyData(0, m);
            a[m] = ct.polyData(1, m);
            c2[m] = ct.polyData(2, m);
        }
        mi_check(mi_tensor_list(c.handle(), (int)ct.getLevel(), (int)B, b.data(), a.data(), nullptr, nullptr,
                                c2.data(), c.stream()),
                 "ckks_square_inplace");
    }
    ct.setScale(ct.getScale() * ct.getScale());
}

CKKSCiphertext ckks_detail::mul_into_at_level(const CKKSContext& c, const CKKSCiphertext& lhs,
                                              const CKKSCiphertext* rhs, size_t level) {
    const char* fn = rhs ? "ckks_mul_inplace(ct,ct)" : "ckks_square_inplace";
    check_ct(c, lhs, fn);
    const CKKSCiphertext& r = rhs ? *rhs : lhs;
    if (r.getN() != lhs.getN() || level > lhs.getLevel() || level > r.getLevel())
        throw std::invalid_argument(std::string(fn) + ": ciphertext size mismatch");
    if (lhs.getNumPolys() != 2 || r.getNumPolys() != 2)
        throw std::invalid_argument(std::string(fn) + ": expected 2-component inputs");
    if (lhs.getSecretOwner() != r.getSecretOwner() || lhs.getMessageEncodingState() != r.getMessageEncodingState())
        throw std::invalid_argument(std::string(fn) + ": ciphertext metadata mismatch");
    if (lhs.getBatchSize() != r.getBatchSize()) throw std::invalid_argument(std::string(fn) + ": batch size mismatch");
    const size_t B = lhs.getBatchSize();
    CKKSCiphertext out(c, 3, lhs.getScale() * r.getScale(), level, lhs.getSecretOwner(),
                       lhs.getMessageEncodingState(), B);
    std::vector<const uint64_t*> b1(B), a1(B), b2(B), a2(B);
    std::vector<uint64_t*> bo(B), ao(B), c2(B);
    for (size_t m = 0; m < B; ++m) {
        b1[m] = lhs.polyData(0, m);
        a1[m] = lhs.polyData(1, m);
        b2[m] = r.polyData(0, m);
        a2[m] = r.polyData(1, m);
        bo[m] = out.polyData(0, m);
        ao[m] = out.polyData(1, m);
        c2[m] = out.polyData(2, m);
    }
    mi_check(mi_tensor_list_into(c.handle(), (int)level, (int)B, b1.data(), a1.data(), rhs ? b2.data() : nullptr,
                                 rhs ? a2.data() : nullptr, bo.data(), ao.data(), c2.data(), c.stream()),
             fn);
    return out;
}

void ckks_mul_inplace(const CKKSContext& c, CKKSCiphertext& ct, int64_t scalar) {
    check_ct(c, ct, "ckks_mul_inplace(ct,int)");
    if (ct.getNumPolys() == 0) throw std::invalid_argument("ckks_mul_inplace(ct,int): empty ciphertext");
    if (scalar == 1) return;
    const size_t lv = ct.getLevel();
    const auto& q = c.getParams().getModuli();
    std::vector<uint64_t> r(lv + 1);
    for (size_t l = 0; l <= lv; ++l) r[l] = reduce_i64(scalar, q[l]);
    ckks_detail::poly_list_op(c, MI_LIST_MUL_SCALAR, ct, nullptr, r.data(), "ckks_mul_inplace(ct,int)");
}

void ckks_mul_inplace(const CKKSContext& c, CKKSCiphertext& ct, Complex value, double plaintext_scale) {
    // the same constant encoding every time: made once per (value, scale, level)
    const double re = value.real(), im = value.imag();
    uint64_t kr, ki, ks;
    std::memcpy(&kr, &re, 8);
    std::memcpy(&ki, &im, 8);
    std::memcpy(&ks, &plaintext_scale, 8);
    std::optional<CKKSEncoding> own;
    const CKKSEncoding* pt = c.cachedEncoding({(uint64_t{1} << 32) | ct.getLevel(), kr, ki, ks}, [&] {
        MyVector<Complex> slots(c.getParams().getSlots(), value);
        return ckks_encode(c, slots, plaintext_scale, ct.getLevel(), /*eval=*/true);
    }, own);
    ckks_mul_inplace(c, ct, *pt);
}

CKKSCiphertext ckks_mul(const CKKSContext& c, const CKKSCiphertext& ct, Complex value, double plaintext_scale) {
    CKKSCiphertext out = ct.clone();
    ckks_mul_inplace(c, out, value, plaintext_scale);
    return out;
}

void ckks_rescale(const CKKSContext& c, CKKSCiphertext& ct) {
    const size_t lv = ct.getLevel();
    if (lv == 0 || lv > c.getParams().getMaxLevel()) throw std::invalid_argument("ckks_rescale: invalid level");
    check_ct(c, ct, "ckks_rescale");
    if (ct.getNumPolys() == 0) throw std::invalid_argument("ckks_rescale: empty ciphertext");
    std::vector<uint64_t*> polys = ct.allPolyData();
    mi_check(mi_rescale_list(c.handle(), (int)lv, (int)polys.size(), polys.data(), c.stream()), "ckks_rescale");
    const uint64_t p = c.getParams().getModuli()[lv];
    ct.setScale(ct.getScale() / (long double)p);
    ct.setLevel(lv - 1);
}

void ckks_drop_to_level_inplace(CKKSCiphertext& ct, size_t target_level) {
    if (target_level > ct.getLevel())
        throw std::invalid_argument("ckks_drop_to_level_inplace: cannot raise ciphertext level");
    if (ct.getNumPolys() == 0) throw std::invalid_argument("ckks_drop_to_level_inplace: empty ciphertext");
    ct.setLevel(target_level);
}

void ckks_add_inplace(const CKKSContext& c, CKKSCiphertext& ct, const CKKSEncoding& pt) {
    check_ct(c, ct, "ckks_add_inplace(ct,pt)");
    if (pt.getN() != ct.getN() || pt.getLevel() != ct.getLevel())
        throw std::invalid_argument("ckks_add_inplace(ct,pt): ciphertext/plaintext size mismatch");
    if (ct.getNumPolys() == 0) throw std::invalid_argument("ckks_add_inplace(ct,pt): empty ciphertext");
    if (ct.getBatchSize() == 1) {
        mi_check(mi_add_plain(c.handle(), (int)ct.getLevel(), ct.polyData(0), pt.data(),
                              pt.isMontgomeryForm() ? 1 : 0, c.stream()),
                 "ckks_add_inplace(ct,pt)");
        return;
    }
    const size_t B = ct.getBatchSize();
    std::vector<uint64_t*> x(B);
    std::vector<const uint64_t*> y(B, pt.data());
    for (size_t m = 0; m < B; ++m) x[m] = ct.polyData(0, m);
    mi_check(mi_poly_list_op(c.handle(), pt.isMontgomeryForm() ? MI_LIST_ADD_PLAIN_MONT : MI_LIST_ADD,
                             (int)ct.getLevel(), (int)B, x.data(), y.data(), nullptr, c.stream()),
             "ckks_add_inplace(ct,pt)");
}

static void add_or_sub(const CKKSContext& c, CKKSCiphertext& ct1, const CKKSCiphertext& ct2, bool sub,
                       const char* fn) {
    check_ct(c, ct1, fn);
    if (ct1.getN() != ct2.getN() || ct1.getLevel() != ct2.getLevel() || ct1.getNumPolys() != ct2.getNumPolys())
        throw std::invalid_argument(std::string(fn) + ": ciphertext size mismatch");
    if (ct1.getSecretOwner() != ct2.getSecretOwner() ||
        ct1.getMessageEncodingState() != ct2.getMessageEncodingState())
        throw std::invalid_argument(std::string(fn) + ": ciphertext metadata mismatch");
    if (ct1.getBatchSize() != ct2.getBatchSize()) throw std::invalid_argument(std::string(fn) + ": batch size mismatch");
    ckks_detail::poly_list_op(c, sub ? MI_LIST_SUB : MI_LIST_ADD, ct1, &ct2, nullptr, fn);
}

void ckks_detail::poly_list_op(const CKKSContext& c, int op, CKKSCiphertext& ct, const CKKSCiphertext* other,
                               const uint64_t* residues, const char* fn) {
    if (ct.getNumPolys() == 0) return;
    if (other && (other->getBatchSize() != ct.getBatchSize() || other->getNumPolys() < ct.getNumPolys()))
        throw std::invalid_argument(std::string(fn) + ": ciphertext batch or degree mismatch");
    std::vector<uint64_t*> x = ct.allPolyData();
    std::vector<const uint64_t*> y(x.size(), nullptr);
    if (other) y = other->allPolyData();
    mi_check(mi_poly_list_op(c.handle(), op, (int)ct.getLevel(), (int)x.size(), x.data(), y.data(), residues,
                             c.stream()),
             fn);
}

void ckks_detail::b_list_op(const CKKSContext& c, int op, CKKSCiphertext& ct, const uint64_t* residues,
                            const char* fn) {
    if (ct.getNumPolys() == 0) throw std::invalid_argument(std::string(fn) + ": empty ciphertext");
    std::vector<uint64_t*> x(ct.getBatchSize());
    for (size_t m = 0; m < x.size(); ++m) x[m] = ct.polyData(0, m);
    mi_check(mi_poly_list_op(c.handle(), op, (int)ct.getLevel(), (int)x.size(), x.data(), nullptr, residues,
                             c.stream()),
             fn);
}

void ckks_detail::zero_polys(const CKKSContext& c, CKKSCiphertext& ct, size_t from, const char* fn) {
    std::vector<uint64_t*> x;
    for (size_t k = from; k < ct.getNumPolys(); ++k)
        for (size_t m = 0; m < ct.getBatchSize(); ++m) x.push_back(ct.polyData(k, m));
    if (x.empty()) return;
    mi_check(mi_poly_list_op(c.handle(), MI_LIST_ZERO, (int)ct.getLevel(), (int)x.size(), x.data(), nullptr, nullptr,
                             c.stream()),
             fn);
}

void ckks_add_inplace(const CKKSContext& c, CKKSCiphertext& ct1, const CKKSCiphertext& ct2) {
    add_or_sub(c, ct1, ct2, false, "ckks_add_inplace(ct,ct)");
}

void ckks_sub_inplace(const CKKSContext& c, CKKSCiphertext& ct1, const CKKSCiphertext& ct2) {
    add_or_sub(c, ct1, ct2, true, "ckks_sub_inplace(ct,ct)");
}

void ckks_add_integer_constant_inplace(const CKKSContext& c, CKKSCiphertext& ct, int64_t constant) {
    if (constant == 0) return;
    check_ct(c, ct, "ckks_add_integer_constant_inplace");
    const size_t lv = ct.getLevel();
    const auto& q = c.getParams().getModuli();
    std::vector<uint64_t> r(lv + 1);
    for (size_t l = 0; l <= lv; ++l) r[l] = rounded_scaled_integer_mod_q(constant, ct.getScale(), q[l]);
    if (ct.getBatchSize() == 1)
        mi_check(mi_add_scalar(c.handle(), (int)lv, ct.polyData(0), r.data(), c.stream()),
                 "ckks_add_integer_constant_inplace");
    else
        ckks_detail::b_list_op(c, MI_LIST_ADD_SCALAR, ct, r.data(), "ckks_add_integer_constant_inplace");
}

void ckks_relin_hybrid_inplace(const CKKSContext& c, CKKSCiphertext& ct) {
    check_ct(c, ct, "ckks_relin");
    if (ct.getNumPolys() != 3) throw std::invalid_argument("ckks_relin: expected 3-polynomial ciphertext");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_relin: sparse-secret relinearization keys are not available");
    if (ct.getBatchSize() == 1) {
        mi_check(mi_relin(c.handle(), (int)ct.getLevel(), ct.polyData(0), ct.polyData(1), ct.polyData(2),
                          c.getRelinKeyDevice(), c.getParams().montgomery ? 1 : 0, c.stream()),
                 "ckks_relin");
    } else {
        // (b, a) += KS_rlk(c2) of every member as one key-switch batch
        std::vector<mi_keyswitch_item> items(ct.getBatchSize());
        for (size_t m = 0; m < items.size(); ++m)
            items[m] = mi_keyswitch_item{ct.polyData(2, m), c.getRelinKeyDevice(), ct.polyData(0, m),
                                         ct.polyData(1, m),  (int32_t)ct.getLevel(), 0, 0, 0};
        mi_check(mi_keyswitch_batch(c.handle(), items.data(), (int)items.size(), c.getParams().montgomery ? 1 : 0,
                                    c.stream()),
                 "ckks_relin");
    }
    ct.resizePolys(2);
}

void ckks_rotate_inplace(const CKKSContext& c, CKKSCiphertext& ct, int64_t rotation) {
    const size_t N = c.getParams().getN();
    const int64_t r = normalize_rotation_value(N, rotation);
    if (r == 0) return;
    if (!c.hasRotationKey(r))
        throw std::invalid_argument("ckks_rotate: rotation key is unavailable for the requested rotation");
    if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_rotate: expected a 2-polynomial ciphertext");
    check_ct(c, ct, "ckks_rotate");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_rotate: sparse-secret rotation keys are not available");
    if (ct.getBatchSize() == 1) {
        mi_check(mi_rotate(c.handle(), (int)ct.getLevel(), ct.polyData(0), ct.polyData(1),
                           rotation_automorphism_index(N, r), c.getRotationKeyDevice(r),
                           c.getParams().montgomery ? 1 : 0, c.stream()),
                 "ckks_rotate");
        return;
    }
    const size_t B = ct.getBatchSize();
    std::vector<uint64_t*> b(B), a(B);
    std::vector<uint64_t> galois(B, rotation_automorphism_index(N, r));
    std::vector<const uint64_t*> keys(B, c.getRotationKeyDevice(r));
    for (size_t m = 0; m < B; ++m) {
        b[m] = ct.polyData(0, m);
        a[m] = ct.polyData(1, m);
    }
    mi_check(mi_rotate_batch(c.handle(), (int)ct.getLevel(), (int)B, b.data(), a.data(), galois.data(), keys.data(),
                             c.getParams().montgomery ? 1 : 0, c.stream()),
             "ckks_rotate");
}

CKKSCiphertext ckks_rotate(const CKKSContext& c, const CKKSCiphertext& ct, int64_t rotation) {
    CKKSCiphertext out = ct.clone();
    ckks_rotate_inplace(c, out, rotation);
    return out;
}

void ckks_conjugate_inplace(const CKKSContext& c, CKKSCiphertext& ct) {
    if (!c.hasConjugationKey()) throw std::invalid_argument("ckks_conjugate: conjugation key is unavailable");
    if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_conjugate: expected a 2-polynomial ciphertext");
    check_ct(c, ct, "ckks_conjugate");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_conjugate: sparse-secret conjugation keys are not available");
    if (ct.getBatchSize() == 1) {
        mi_check(mi_conjugate(c.handle(), (int)ct.getLevel(), ct.polyData(0), ct.polyData(1),
                              c.getConjugationKeyDevice(), c.getParams().montgomery ? 1 : 0, c.stream()),
                 "ckks_conjugate");
        return;
    }
    std::vector<uint64_t*> b(ct.getBatchSize()), a(ct.getBatchSize());
    for (size_t m = 0; m < b.size(); ++m) {
        b[m] = ct.polyData(0, m);
        a[m] = ct.polyData(1, m);
    }
    mi_check(mi_conjugate_batch(c.handle(), (int)ct.getLevel(), (int)b.size(), b.data(), a.data(),
                                c.getConjugationKeyDevice(), c.getParams().montgomery ? 1 : 0, c.stream()),
             "ckks_conjugate");
}

MyVector<CKKSCiphertext> ckks_conjugate_batch(const CKKSContext& c, const MyVector<CKKSCiphertext>& cts) {
    MyVector<CKKSCiphertext> out;
    if (cts.empty()) return out;
    // each element passes ckks_conjugate_inplace's checks
    if (!c.hasConjugationKey()) throw std::invalid_argument("ckks_conjugate: conjugation key is unavailable");
    out.reserve(cts.size());
    std::vector<uint64_t*> b, a;
    for (const auto& ct : cts) {
        if (ct.getNumPolys() != 2) throw std::invalid_argument("ckks_conjugate: expected a 2-polynomial ciphertext");
        check_ct(c, ct, "ckks_conjugate");
        if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
            throw std::invalid_argument("ckks_conjugate: sparse-secret conjugation keys are not available");
        if (ct.getLevel() != cts[0].getLevel())
            throw std::invalid_argument("ckks_conjugate: batch ciphertexts must have one level");
        out.push_back(ct.clone());
        for (size_t m = 0; m < ct.getBatchSize(); ++m) {
            b.push_back(out.back().polyData(0, m));
            a.push_back(out.back().polyData(1, m));
        }
    }
    mi_check(mi_conjugate_batch(c.handle(), (int)cts[0].getLevel(), (int)b.size(), b.data(), a.data(),
                                c.getConjugationKeyDevice(), c.getParams().montgomery ? 1 : 0, c.stream()),
             "ckks_conjugate");
    return out;
}

CKKSCiphertext ckks_conjugate(const CKKSContext& c, const CKKSCiphertext& ct) {
    CKKSCiphertext out = ct.clone();
    ckks_conjugate_inplace(c, out);
    return out;
}

void CKKSGraph::launch() const {
    if (!g_ || !ctx_) throw std::invalid_argument("ckks_graph: empty graph");
    mi_check(mi_graph_launch(ctx_->handle(), g_, ctx_->stream()), "ckks_graph");
}

void CKKSGraph::reset() {
    if (g_ && ctx_) (void)mi_graph_destroy(ctx_->handle(), g_);
    g_ = nullptr;
    ctx_ = nullptr;
}

CKKSGraph ckks_capture_graph(const CKKSContext& c, const std::function<void()>& body) {
    mi_check(mi_graph_begin(c.handle(), c.stream()), "ckks_graph");
    try {
        body();
    } catch (...) {
        mi_graph* g = nullptr;
        (void)mi_graph_end(c.handle(), c.stream(), &g);  // ends the capture; the partial graph is dropped
        if (g) (void)mi_graph_destroy(c.handle(), g);
        throw;
    }
    CKKSGraph out;
    mi_check(mi_graph_end(c.handle(), c.stream(), &out.g_), "ckks_graph");
    out.ctx_ = &c;
    return out;
}

void ckks_copy_into(const CKKSContext& c, CKKSCiphertext& dst, const CKKSCiphertext& src) {
    if (dst.getN() != src.getN() || dst.getLevel() != src.getLevel() || dst.getNumPolys() != src.getNumPolys() ||
        dst.getBatchSize() != src.getBatchSize())
        throw std::invalid_argument("ckks_copy_into: ciphertext shape mismatch");
    ckks_detail::poly_list_op(c, MI_LIST_COPY, dst, &src, nullptr, "ckks_copy_into");
    dst.setScale(src.getScale());
}
