// The keyswitch-level functions of the reference (ckks_keyswitch.hpp:59-149,
// ckks_keyswitch.cpp:335-1816) over the device pipeline.  Switching keys stay
// in HBM (CKKSSwitchingKey, from CKKSContext::getRotationKey and friends);
// polynomials given as host vectors are uploaded, switched on the device and
// downloaded, so each call returns what the reference's host code returns.
// The batched, device-resident forms of the same operations are the C-ABI
// entry points (mi_keyswitch*, mi_hoist, mi_rotate*) the evaluator uses.
#include <bit>
#include <cstring>

#include "ckks.hpp"

namespace {

int mode_of(CKKSHybridKeySwitchCombine m) { return m == CKKSHybridKeySwitchCombine::Assign ? 1 : 0; }

size_t brev(size_t x, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

// The Galois element of an evaluation-domain automorphism map
// (build_ntt_automorphism_map, rotation_key_generation.cpp:80-90: dst index
// brev(d) reads src index brev(((2d + 1) g mod 2N - 1) / 2)); 0 for an empty
// map.  The whole map is checked, since the device applies the automorphism
// of g rather than the table.
uint64_t galois_of_map(std::span<const size_t> map, size_t N) {
    if (map.empty()) return 0;
    if (map.size() != N) throw std::invalid_argument("ckks_hybrid_keyswitch: invalid NTT map size");
    const int logN = std::countr_zero(N);
    const uint64_t g = 2 * (uint64_t)brev(map[0], logN) + 1;
    for (size_t d = 0; d < N; ++d) {
        const size_t src = (size_t)((((2 * (uint64_t)d + 1) * g) % (2 * (uint64_t)N) - 1) >> 1);
        if (map[brev(d, logN)] != brev(src, logN))
            throw std::invalid_argument("ckks_hybrid_keyswitch: NTT map is not a ring automorphism");
    }
    return g;
}

void check_level(const CKKSContext& c, size_t level, const char* fn) {
    if (level > c.getParams().getMaxLevel()) throw std::invalid_argument(std::string(fn) + ": invalid level");
}

void check_key(const CKKSSwitchingKey& key) {
    if (!key.device) throw std::invalid_argument("ckks_hybrid_keyswitch: empty switching key");
}

DeviceBuffer to_device(const CKKSContext& c, const MyVector<uint64_t>& h) {
    DeviceBuffer d(c.handle(), h.size());
    if (!h.empty())
        mi_check(mi_memcpy(c.handle(), d.data(), h.data(), h.size() * 8, MI_COPY_H2D, c.stream()), "upload");
    return d;
}

void to_host(const CKKSContext& c, MyVector<uint64_t>& h, const DeviceBuffer& d) {
    if (!h.empty())
        mi_check(mi_memcpy(c.handle(), h.data(), d.data(), h.size() * 8, MI_COPY_D2H, c.stream()), "download");
}

MyVector<size_t> automorphism_map(size_t N, uint64_t g) {
    const int logN = std::countr_zero(N);
    MyVector<size_t> map(N);
    for (size_t d = 0; d < N; ++d)
        map[brev(d, logN)] = brev((size_t)((((2 * (uint64_t)d + 1) * g) % (2 * (uint64_t)N) - 1) >> 1), logN);
    return map;
}

}  // namespace

MyVector<size_t> CKKSContext::getRotationNttMap(int64_t rotation) const {
    if (!hasRotationKey(rotation)) throw std::invalid_argument("ckks_context: rotation key is unavailable");
    const size_t N = getParams().getN();
    return automorphism_map(N, rotation_automorphism_index(N, rotation));
}

MyVector<size_t> CKKSContext::getConjugationNttMap() const {
    const size_t N = getParams().getN();
    return automorphism_map(N, conjugation_automorphism_index(N));
}

void ckks_apply_ntt_automorphism_to_flat_rns_poly_inplace(MyVector<uint64_t>& poly, size_t N, size_t limb_count,
                                                          std::span<const size_t> dst_to_src_map) {
    if (dst_to_src_map.size() != N || poly.size() < N * limb_count)
        throw std::invalid_argument("ckks_apply_ntt_automorphism: invalid map or polynomial size");
    MyVector<uint64_t> limb(N);
    for (size_t l = 0; l < limb_count; ++l) {
        uint64_t* p = poly.data() + l * N;
        for (size_t i = 0; i < N; ++i) limb[i] = p[dst_to_src_map[i]];
        std::memcpy(p, limb.data(), N * 8);
    }
}

void ckks_hoist_ciphertext_a_qp_eval(const CKKSContext& context, const CKKSCiphertext& ct, CKKSHoistWorkspace&,
                                     CKKSEvalHoistedCiphertextQP& out) {
    const auto& P = context.getParams();
    if (ct.getNumPolys() != 2 || ct.getBatchSize() != 1)
        throw std::invalid_argument("ckks_hoist_ciphertext_a_qp_eval: expected one 2-polynomial ciphertext");
    const size_t level = ct.getLevel(), L = level + 1, N = P.getN(), K = P.getPModuli().size();
    check_level(context, level, "ckks_hoist_ciphertext_a_qp_eval");
    const size_t parts = (L + P.getAlpha() - 1) / P.getAlpha();
    out.N = N;
    out.level = level;
    out.q_limb_count = L;
    out.p_limb_count = K;
    out.scale = ct.getScale();
    out.moduli.assign(P.getModuli().begin(), P.getModuli().begin() + L);
    out.moduli.insert(out.moduli.end(), P.getPModuli().begin(), P.getPModuli().end());
    out.digits = DeviceBuffer(context.handle(), parts * (L + K) * N);
    mi_check(mi_hoist(context.handle(), (int)level, ct.polyData(1), out.digits.data(), context.stream()),
             "ckks_hoist_ciphertext_a_qp_eval");
}

void ckks_hybrid_key_switch_inplace(const CKKSContext& context, size_t level, MyVector<uint64_t>& key_term,
                                    CKKSSwitchingKey switching_key, MyVector<uint64_t>& b_out,
                                    MyVector<uint64_t>& a_out, CKKSHybridKeySwitchCombine b_mode,
                                    CKKSHybridKeySwitchCombine a_mode, bool switching_key_montgomery) {
    check_level(context, level, "ckks_hybrid_keyswitch");
    check_key(switching_key);
    const size_t W = (level + 1) * context.getParams().getN();
    if (key_term.size() != W || b_out.size() != W || a_out.size() != W)
        throw std::invalid_argument("ckks_hybrid_keyswitch: invalid polynomial size");
    DeviceBuffer kt = to_device(context, key_term), b = to_device(context, b_out), a = to_device(context, a_out);
    mi_check(mi_keyswitch(context.handle(), (int)level, kt.data(), switching_key.device, b.data(), a.data(),
                          mode_of(b_mode), mode_of(a_mode), switching_key_montgomery ? 1 : 0, context.stream()),
             "ckks_hybrid_keyswitch");
    to_host(context, b_out, b);
    to_host(context, a_out, a);
}

void ckks_hybrid_key_switch_hoisted_inplace(const CKKSContext& context,
                                            const CKKSEvalHoistedCiphertextQP& hoisted_key_term,
                                            CKKSSwitchingKey switching_key, MyVector<uint64_t>& b_out,
                                            MyVector<uint64_t>& a_out, CKKSHybridKeySwitchCombine b_mode,
                                            CKKSHybridKeySwitchCombine a_mode, bool switching_key_montgomery,
                                            std::span<const size_t> ntt_map) {
    check_key(switching_key);
    if (!hoisted_key_term.digits.data()) throw std::invalid_argument("ckks_hybrid_keyswitch: empty hoisted key term");
    const size_t level = hoisted_key_term.level, N = hoisted_key_term.N, W = (level + 1) * N;
    if (N != context.getParams().getN()) throw std::invalid_argument("ckks_hybrid_keyswitch: hoisted ring mismatch");
    if (b_out.size() != W || a_out.size() != W)
        throw std::invalid_argument("ckks_hybrid_keyswitch: invalid hoisted output polynomial size");
    const uint64_t g = galois_of_map(ntt_map, N);
    DeviceBuffer b = to_device(context, b_out), a = to_device(context, a_out);
    mi_check(mi_keyswitch_hoisted(context.handle(), (int)level, hoisted_key_term.digits.data(), switching_key.device,
                                  b.data(), a.data(), mode_of(b_mode), mode_of(a_mode), g,
                                  switching_key_montgomery ? 1 : 0, context.stream()),
             "ckks_hybrid_keyswitch");
    to_host(context, b_out, b);
    to_host(context, a_out, a);
}

void ckks_hybrid_key_switch_hoisted_qp_rotated(const CKKSContext& context,
                                               const CKKSEvalHoistedCiphertextQP& hoisted_key_term,
                                               CKKSSwitchingKey switching_key, const MyVector<uint64_t>& b_q_eval,
                                               MyVector<uint64_t>& b_qp_out, MyVector<uint64_t>& a_qp_out,
                                               bool switching_key_montgomery, std::span<const size_t> ntt_map) {
    check_key(switching_key);
    if (!hoisted_key_term.digits.data()) throw std::invalid_argument("ckks_hybrid_keyswitch: empty hoisted key term");
    const size_t level = hoisted_key_term.level, N = hoisted_key_term.N, L = level + 1;
    const size_t K = context.getParams().getPModuli().size();
    if (b_q_eval.size() != L * N) throw std::invalid_argument("ckks_hybrid_keyswitch: invalid b polynomial size");
    b_qp_out.assign((L + K) * N, 0);
    a_qp_out.assign((L + K) * N, 0);
    const uint64_t g = galois_of_map(ntt_map, N);
    DeviceBuffer bq = to_device(context, b_q_eval), b = to_device(context, b_qp_out), a = to_device(context, a_qp_out);
    mi_check(mi_keyswitch_hoisted_qp(context.handle(), (int)level, hoisted_key_term.digits.data(),
                                     switching_key.device, bq.data(), b.data(), a.data(), g,
                                     switching_key_montgomery ? 1 : 0, context.stream()),
             "ckks_hybrid_keyswitch");
    to_host(context, b_qp_out, b);
    to_host(context, a_qp_out, a);
}

void ckks_hybrid_key_switch_qp_from_key_term(const CKKSContext& context, size_t level,
                                             const MyVector<uint64_t>& key_term_q_eval,
                                             CKKSSwitchingKey switching_key, MyVector<uint64_t>& b_qp_out,
                                             MyVector<uint64_t>& a_qp_out, bool switching_key_montgomery) {
    check_level(context, level, "ckks_hybrid_keyswitch");
    check_key(switching_key);
    const size_t N = context.getParams().getN(), L = level + 1, K = context.getParams().getPModuli().size();
    if (key_term_q_eval.size() != L * N) throw std::invalid_argument("ckks_hybrid_keyswitch: invalid key term size");
    b_qp_out.assign((L + K) * N, 0);
    a_qp_out.assign((L + K) * N, 0);
    DeviceBuffer kt = to_device(context, key_term_q_eval), b = to_device(context, b_qp_out),
                 a = to_device(context, a_qp_out);
    mi_check(mi_keyswitch_qp(context.handle(), (int)level, kt.data(), switching_key.device, b.data(), a.data(),
                             switching_key_montgomery ? 1 : 0, context.stream()),
             "ckks_hybrid_keyswitch");
    to_host(context, b_qp_out, b);
    to_host(context, a_qp_out, a);
}

void ckks_key_switch_inplace(const CKKSContext& context, CKKSCiphertext& ct, CKKSSwitchingKey switching_key,
                             bool switching_key_montgomery) {
    check_key(switching_key);
    if (ct.getNumPolys() != 2 || ct.getBatchSize() != 1)
        throw std::invalid_argument("ckks_key_switch: expected one 2-polynomial ciphertext");
    check_level(context, ct.getLevel(), "ckks_key_switch");
    // (b, a) -> (b + KS_b(a), KS_a(a)): the switch of a, b added, a assigned
    mi_check(mi_keyswitch(context.handle(), (int)ct.getLevel(), ct.polyData(1), switching_key.device, ct.polyData(0),
                          ct.polyData(1), 0, 1, switching_key_montgomery ? 1 : 0, context.stream()),
             "ckks_key_switch");
}
