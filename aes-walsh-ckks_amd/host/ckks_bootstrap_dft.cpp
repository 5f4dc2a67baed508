// Factorised homomorphic DFT plans for bootstrapping
// (ckks_generate_bootstrap_dft_plan, ckks_bootstrap_dft.cpp:698-1079) and
// their application (ckks_apply_linear_transform_plan,
// ckks_linear_transform.cpp:1174-1427).
//
// Plan generation is host precompute in complex double with the reference's
// operation order: radix-2 butterfly factors of the (inverse) special FFT
// (:288-385), merged per stage into sparse diagonal maps (:500-690), scaled
// per the plaintext-scale schedule (:240-259) and encoded, pre-rotated by the
// giant-step offset (:52-96), at the stage's Q u P basis.  The residues are
// NTT'd and put in Montgomery form on the device and stay there; applying a
// stage runs the BSGS kernels behind mi_linear_transform_qp (QP evaluator) or
// mi_linear_transform_q (Q evaluator).
#include <algorithm>
#include <bit>
#include <span>
#include <cmath>
#include <limits>
#include <map>

#include "ckks.hpp"

namespace {

using Diagonals = std::map<size_t, MyVector<Complex>>;

size_t active_slot_count(const CKKSContext& c, const CKKSBootstrapDftMatrixLiteral& lit) {
    const size_t full = c.getParams().getSlots();
    const size_t a = lit.active_slots == 0 ? full : lit.active_slots;
    if (a == 0 || a > full || full % a != 0 || !std::has_single_bit(a))
        throw std::invalid_argument("ckks_bootstrap_dft: invalid active DFT slot count");
    return a;
}

size_t factor_count(const CKKSBootstrapDftMatrixLiteral& lit) {
    if (lit.log_plaintext_scale_groups.empty()) return lit.stage_count;
    size_t n = 0;
    for (const auto& g : lit.log_plaintext_scale_groups) n += g.size();
    return n;
}

struct FactorSlot {
    size_t group = 0;
    size_t log_scale = 0;
    bool ends_group = true;
};

std::vector<FactorSlot> factor_schedule(const CKKSBootstrapDftMatrixLiteral& lit) {
    std::vector<FactorSlot> out;
    if (lit.log_plaintext_scale_groups.empty()) {
        for (size_t i = 0; i < lit.stage_count; ++i) out.push_back({i, 0, true});
        return out;
    }
    for (size_t g = 0; g < lit.log_plaintext_scale_groups.size(); ++g) {
        const auto& f = lit.log_plaintext_scale_groups[g];
        if (f.empty()) throw std::invalid_argument("ckks_bootstrap_dft: DFT scale schedule groups must not be empty");
        for (size_t k = 0; k < f.size(); ++k) out.push_back({g, f[k], k + 1 == f.size()});
    }
    return out;
}

double stage_plaintext_scale(const CKKSContext& c, const CKKSBootstrapDftMatrixLiteral& lit, const FactorSlot& f,
                             size_t stage_level, double max_abs) {
    if (!(max_abs > 0.0)) throw std::invalid_argument("ckks_bootstrap_dft: invalid zero DFT diagonal");
    const double q_drop = (double)c.getParams().getModuli()[stage_level];
    if (lit.log_plaintext_scale_groups.empty()) return q_drop / (2.01 * max_abs);
    size_t sum = 0;
    for (size_t v : lit.log_plaintext_scale_groups.at(f.group)) {
        if (v == 0) throw std::invalid_argument("ckks_bootstrap_dft: DFT scale schedule entries must be positive");
        sum += v;
    }
    return std::pow(q_drop, (double)f.log_scale / (double)sum);
}

// one butterfly factor of the special FFT, as three diagonals a (0), b (+r), c (-r)
struct Factor {
    MyVector<Complex> a, b, c;
};

std::vector<Factor> fft_factors(size_t log_slots, bool inverse) {
    const size_t slots = size_t{1} << log_slots;
    const size_t nroots = slots << 2;
    constexpr double pi = 3.141592653589793238462643383279502884;
    MyVector<Complex> roots(nroots + 1);
    for (size_t i = 0; i <= nroots; ++i) {
        const double ang = 2.0 * pi * (double)i / (double)nroots;
        roots[i] = Complex(std::cos(ang), std::sin(ang));
    }
    std::vector<size_t> pow5((slots << 1) + 1);
    pow5[0] = 1;
    for (size_t i = 1; i < pow5.size(); ++i) pow5[i] = (pow5[i - 1] * 5) & ((slots << 2) - 1);
    std::vector<Factor> out;
    auto build = [&](size_t m) {
        Factor f{MyVector<Complex>(slots, Complex(0.0, 0.0)), MyVector<Complex>(slots, Complex(0.0, 0.0)),
                 MyVector<Complex>(slots, Complex(0.0, 0.0))};
        const size_t h = m >> 1, gap = slots / m, mask = (m << 2) - 1;
        for (size_t i = 0; i < slots; i += m)
            for (size_t j = 0; j < h; ++j) {
                const size_t i1 = i + j, i2 = i1 + h;
                if (!inverse) {
                    const size_t ri = (pow5[j] & mask) * gap;
                    f.a[i1] = Complex(1.0, 0.0);
                    f.a[i2] = -roots[ri];
                    f.b[i1] = roots[ri];
                    f.c[i2] = Complex(1.0, 0.0);
                } else {
                    const size_t ri = ((m << 2) - (pow5[j] & mask)) * gap;
                    f.a[i1] = Complex(1.0, 0.0);
                    f.a[i2] = -roots[ri];
                    f.b[i1] = Complex(1.0, 0.0);
                    f.c[i2] = roots[ri];
                }
            }
        out.push_back(std::move(f));
    };
    if (!inverse)
        for (size_t m = 2; m <= slots; m <<= 1) build(m);
    else
        for (size_t m = slots; m >= 2; m >>= 1) build(m);
    return out;
}

size_t factor_rotation(size_t log_slots, size_t modulo, size_t level, CKKSBootstrapDftType t) {
    const size_t r = t == CKKSBootstrapDftType::CoeffsToSlots ? (size_t{1} << (level - 1))
                                                              : (size_t{1} << (log_slots - level));
    return r & (modulo - 1);
}

void accumulate(Diagonals& m, size_t idx, const MyVector<Complex>& v) {
    auto [it, fresh] = m.emplace(idx, v);
    if (!fresh) {
        if (it->second.size() != v.size())
            throw std::invalid_argument("ckks_bootstrap_dft: inconsistent sparse diagonal vector size");
        for (size_t i = 0; i < v.size(); ++i) it->second[i] += v[i];
    }
}

MyVector<Complex> rotated_times(const MyVector<Complex>& v, int64_t rot, const MyVector<Complex>& mul) {
    const size_t n = v.size(), mask = n - 1;
    const size_t off = (size_t)((rot % (int64_t)n + (int64_t)n) % (int64_t)n);
    MyVector<Complex> o(n);
    for (size_t i = 0; i < n; ++i) o[i] = mul[i] * v[(i + off) & mask];
    return o;
}

std::vector<size_t> merge_depths(size_t log_slots, size_t nfactors, CKKSBootstrapDftType t,
                                 const std::vector<size_t>& override_depths) {
    if (!override_depths.empty()) {
        if (override_depths.size() != nfactors)
            throw std::invalid_argument("ckks_bootstrap_dft: DFT merge-depth override size mismatch");
        size_t tot = 0;
        for (size_t d : override_depths) {
            if (d == 0) throw std::invalid_argument("ckks_bootstrap_dft: DFT merge depths must be positive");
            tot += d;
        }
        if (tot != log_slots) throw std::invalid_argument("ckks_bootstrap_dft: DFT merge-depth override must sum to log_slots");
        return override_depths;
    }
    if (nfactors == 0 || nfactors > log_slots) throw std::invalid_argument("ckks_bootstrap_dft: invalid DFT factor count");
    std::vector<size_t> m(nfactors);
    size_t level = log_slots;
    for (size_t i = 0; i < nfactors; ++i) {
        const size_t rest = nfactors - i;
        const size_t d = (level + rest - 1) / rest;
        if (t == CKKSBootstrapDftType::CoeffsToSlots) m[i] = d;
        else m[nfactors - i - 1] = d;
        level -= d;
    }
    return m;
}

Diagonals stage_diagonals(size_t log_slots, const std::vector<Factor>& f, const std::vector<size_t>& merge,
                          CKKSBootstrapDftType t, size_t stage, double scale) {
    const size_t slots = size_t{1} << log_slots;
    size_t level = log_slots;
    for (size_t i = 0; i < stage; ++i) level -= merge[i];
    Diagonals m;
    {
        const Factor& g = f[log_slots - level];
        const size_t r = factor_rotation(log_slots, slots, level, t);
        accumulate(m, 0, g.a);
        accumulate(m, r, g.b);
        accumulate(m, (slots - r) & (slots - 1), g.c);
    }
    size_t next = level - 1;
    for (size_t j = 0; j + 1 < merge[stage]; ++j, --next) {
        const Factor& g = f[log_slots - next];
        const size_t r = factor_rotation(log_slots, slots, next, t);
        Diagonals merged;
        for (const auto& [idx, v] : m) {
            accumulate(merged, idx, rotated_times(v, 0, g.a));
            accumulate(merged, (idx + r) & (slots - 1), rotated_times(v, (int64_t)r, g.b));
            accumulate(merged, (idx + slots - r) & (slots - 1), rotated_times(v, -(int64_t)r, g.c));
        }
        m = std::move(merged);
    }
    for (auto& [idx, v] : m) {
        if (idx >= slots || v.size() != slots) throw std::invalid_argument("ckks_bootstrap_dft: invalid sparse DFT diagonal");
        for (auto& z : v) z *= scale;
    }
    return m;
}

void select_inputs(Diagonals& m, size_t slots, size_t start, size_t size) {
    if (size == 0) return;
    if (start >= slots || size > slots) throw std::invalid_argument("ckks_bootstrap_dft: S2C input selector exceeds slot count");
    for (auto& [idx, v] : m)
        for (size_t row = 0; row < slots; ++row) {
            const size_t col = (row + idx) % slots;
            if ((col + slots - start) % slots >= size) v[row] = Complex(0.0, 0.0);
        }
}

double max_abs(const Diagonals& m) {
    double x = 0.0;
    for (const auto& [idx, v] : m)
        for (const auto& z : v) x = std::max(x, std::abs(z));
    return x;
}

bool all_zero(const MyVector<Complex>& v) {
    for (const auto& z : v)
        if (z.real() != 0.0 || z.imag() != 0.0) return false;
    return true;
}

size_t auto_baby_steps(size_t slots, const std::vector<size_t>& idx, int log_ratio) {
    if (slots == 0 || idx.empty()) throw std::invalid_argument("ckks_bootstrap_dft: invalid empty diagonal index set");
    if (idx.size() <= 1) return 1;
    size_t best = 1, best_cost = std::numeric_limits<size_t>::max(), best_giants = std::numeric_limits<size_t>::max();
    const size_t w = size_t{1} << std::max(0, log_ratio);
    for (size_t n1 = 1; n1 < slots; n1 <<= 1) {
        std::vector<char> g(slots, 0), b(slots, 0);
        size_t ng = 0, nb = 0;
        for (size_t d : idx) {
            if (d >= slots) throw std::invalid_argument("ckks_bootstrap_dft: diagonal index exceeds slot count");
            const size_t r = d & (slots - 1);
            const size_t gi = ((r / n1) * n1) & (slots - 1), bi = r & (n1 - 1);
            if (!g[gi]) { g[gi] = 1; ++ng; }
            if (!b[bi]) { b[bi] = 1; ++nb; }
        }
        const size_t nz_g = ng > 0 && g[0] ? ng - 1 : ng;
        const size_t nz_b = nb > 0 && b[0] ? nb - 1 : nb;
        const size_t cost = w * nz_g + nz_b;
        if (cost < best_cost || (cost == best_cost && nz_g < best_giants)) {
            best_cost = cost;
            best_giants = nz_g;
            best = n1;
        }
    }
    return best;
}

// Slot values rotated right by `rot`, Hermitian-embedded, special FFT,
// llrint(value * scale / N) per modulus of the Q u P basis.
std::vector<uint64_t> encode_rotated(const CKKSContext& c, const MyVector<Complex>& values, size_t rot, double scale,
                                     const std::vector<uint64_t>& moduli) {
    const size_t N = c.getParams().getN(), slots = N >> 1, mask = slots - 1;
    const auto& perm = c.getPermTable();
    MyVector<Complex> ev(N, Complex(0.0, 0.0));
    rot &= mask;
    for (size_t i = 0; i < slots; ++i) {
        const Complex v = values[(i + slots - rot) & mask];
        ev[perm[i]] = v;
        ev[N - 1 - perm[i]] = std::conj(v);
    }
    const auto& tw = c.getTwiddleFft();
    for (size_t m = N >> 1, k = 1; m > 0; m >>= 1, k <<= 1)
        for (size_t i = 0; i < m; ++i) {
            const Complex w = tw[m + i];
            Complex* x0 = ev.data() + ((i << 1) * k);
            Complex* x1 = x0 + k;
            for (size_t j = 0; j < k; ++j) {
                const Complex u = x0[j], v = x1[j];
                x0[j] = u + v;
                x1[j] = (u - v) * w;
            }
        }
    std::vector<uint64_t> out(moduli.size() * N);
    const double kf = scale / (double)N;
    for (size_t l = 0; l < moduli.size(); ++l) {
        const int64_t q = (int64_t)moduli[l];
        for (size_t i = 0; i < N; ++i) {
            const int64_t r = std::llrint(ev[i].real() * kf) % q;
            out[l * N + i] = (uint64_t)(r + (q & (r >> 63)));
        }
    }
    return out;
}

CKKSLinearTransformOwnedStage encode_stage(const CKKSContext& c, const CKKSBootstrapDftMatrixLiteral& lit,
                                           const Diagonals& m, bool final_stage) {
    const auto& P = c.getParams();
    const size_t N = P.getN(), full = P.getSlots(), active = active_slot_count(c, lit);
    const size_t level = lit.level, L = level + 1, K = P.getPModuli().size(), nq = P.getModuli().size();
    if (m.empty()) throw std::invalid_argument("ckks_bootstrap_dft: empty sparse diagonal map");
    if (level > P.getMaxLevel()) throw std::invalid_argument("ckks_bootstrap_dft: invalid sparse matrix level");
    if (!(lit.plaintext_scale > 0.0)) throw std::invalid_argument("ckks_bootstrap_dft: sparse plaintext scale must be positive");
    std::vector<uint64_t> qp(P.getModuli().begin(), P.getModuli().begin() + L);
    qp.insert(qp.end(), P.getPModuli().begin(), P.getPModuli().end());
    std::vector<size_t> idx;
    std::vector<const MyVector<Complex>*> vals;
    for (const auto& [i, v] : m) {
        if (i >= active || (v.size() != active && v.size() != full))
            throw std::invalid_argument("ckks_bootstrap_dft: invalid sparse DFT diagonal");
        if (!all_zero(v)) {
            idx.push_back(i);
            vals.push_back(&v);
        }
    }
    if (idx.empty()) throw std::runtime_error("ckks_bootstrap_dft: generated sparse stage has no nonzero diagonals");
    const size_t bs = lit.baby_step_count == 0 ? auto_baby_steps(active, idx, lit.log_bsgs_ratio) : lit.baby_step_count;
    if (bs == 0 || bs > active) throw std::invalid_argument("ckks_bootstrap_dft: invalid sparse baby-step count");

    CKKSLinearTransformOwnedStage st;
    st.diagonal_indices = idx;
    st.slot_count = active;
    st.baby_step_count = bs;
    st.qp_limb_count = qp.size();
    st.qp_p_limb_count = K;
    st.level = level;
    st.diagonal_qp_plaintexts_montgomery = P.montgomery;
    st.plaintext_scale = lit.plaintext_scale;
    st.rescale_after = lit.rescale_after_each_stage;
    st.output_encoding_state = final_stage && lit.type == CKKSBootstrapDftType::SlotsToCoeffs
                                   ? CKKSMessageEncodingState::Coefficients
                                   : CKKSMessageEncodingState::Slots;
    mi_context* mi = c.handle();
    std::vector<uint64_t> rq(L), rp(K);
    for (size_t l = 0; l < L; ++l) rq[l] = (uint64_t)(((unsigned __int128)1 << 64) % qp[l]);
    for (size_t j = 0; j < K; ++j) rp[j] = (uint64_t)(((unsigned __int128)1 << 64) % qp[L + j]);
    for (size_t t = 0; t < idx.size(); ++t) {
        const size_t giant_off = (idx[t] / bs) * bs;
        MyVector<Complex> diag;
        if (vals[t]->size() == full) {
            diag = *vals[t];
        } else {
            diag.resize(full);
            for (size_t s = 0; s < full; ++s) diag[s] = (*vals[t])[s % active];
        }
        const auto h = encode_rotated(c, diag, giant_off, lit.plaintext_scale, qp);
        DeviceBuffer d(mi, h.size());
        mi_check(mi_memcpy(mi, d.data(), h.data(), h.size() * 8, MI_COPY_H2D, c.stream()), "ckks_bootstrap_dft");
        uint64_t* pp = d.data() + L * N;
        mi_check(mi_ntt_forward(mi, d.data(), 1, (int)L, (int64_t)(L * N), 0, c.stream()), "ckks_bootstrap_dft");
        mi_check(mi_ntt_forward(mi, pp, 1, (int)K, (int64_t)(K * N), (int)nq, c.stream()), "ckks_bootstrap_dft");
        if (P.montgomery) {
            mi_check(mi_poly_scalar(mi, MI_OP_MUL, d.data(), rq.data(), (int)L, 0, 1, (int64_t)(L * N), c.stream()),
                     "ckks_bootstrap_dft");
            mi_check(mi_poly_scalar(mi, MI_OP_MUL, pp, rp.data(), (int)K, (int)nq, 1, (int64_t)(K * N), c.stream()),
                     "ckks_bootstrap_dft");
        }
        st.diagonal_qp_plaintexts.push_back(std::move(d));
    }
    mi_check(mi_stream_synchronize(mi, c.stream()), "ckks_bootstrap_dft");
    return st;
}

}  // namespace

MyVector<uint64_t> CKKSLinearTransformOwnedStage::downloadDiagonalQp(const CKKSContext& c, size_t term) const {
    const auto& d = diagonal_qp_plaintexts.at(term);
    MyVector<uint64_t> h(qp_limb_count * c.getParams().getN());
    mi_check(mi_memcpy(c.handle(), h.data(), d.data(), h.size() * 8, MI_COPY_D2H, c.stream()), "downloadDiagonalQp");
    return h;
}

CKKSLinearTransformPlan ckks_encode_linear_transform_plan(const CKKSContext& c,
                                                          const std::map<size_t, MyVector<Complex>>& diagonals,
                                                          size_t level, size_t baby_step_count, double plaintext_scale,
                                                          bool use_qp_evaluator, bool rescale_after) {
    CKKSBootstrapDftMatrixLiteral lit;
    lit.type = CKKSBootstrapDftType::CoeffsToSlots;  // output stays slot-encoded
    lit.level = level;
    lit.baby_step_count = baby_step_count;
    lit.plaintext_scale = plaintext_scale;
    lit.rescale_after_each_stage = rescale_after;
    CKKSLinearTransformPlan plan;
    plan.use_qp_evaluator = use_qp_evaluator;
    plan.stages.push_back(encode_stage(c, lit, diagonals, true));
    return plan;
}

CKKSLinearTransformPlan ckks_generate_bootstrap_dft_plan(const CKKSContext& c, const CKKSBootstrapDftMatrixLiteral& lit) {
    const auto& P = c.getParams();
    const size_t slots = active_slot_count(c, lit);
    const size_t nf = factor_count(lit);
    const size_t groups = lit.log_plaintext_scale_groups.empty() ? lit.stage_count : lit.log_plaintext_scale_groups.size();
    if (nf == 0) throw std::invalid_argument("ckks_bootstrap_dft: stage count must be positive");
    if (lit.level > P.getMaxLevel()) throw std::invalid_argument("ckks_bootstrap_dft: invalid matrix level");
    if (!(lit.plaintext_scale > 0.0)) throw std::invalid_argument("ckks_bootstrap_dft: plaintext scale must be positive");
    if (!(lit.transform_scale > 0.0)) throw std::invalid_argument("ckks_bootstrap_dft: transform scale must be positive");
    if (lit.input_selector_size != 0 && lit.type != CKKSBootstrapDftType::SlotsToCoeffs)
        throw std::invalid_argument("ckks_bootstrap_dft: input selector is only valid for S2C");
    if (lit.input_selector_size != 0 && (lit.input_selector_start >= slots || lit.input_selector_size > slots))
        throw std::invalid_argument("ckks_bootstrap_dft: input selector exceeds slot count");
    const size_t log_slots = (size_t)std::countr_zero(slots);
    if (nf > log_slots) throw std::invalid_argument("ckks_bootstrap_dft: invalid factorized DFT stage count");
    if (lit.rescale_after_each_stage && groups > lit.level)
        throw std::invalid_argument("ckks_bootstrap_dft: factorized DFT consumes more levels than available");

    const auto sched = factor_schedule(lit);
    const bool s2c = lit.type == CKKSBootstrapDftType::SlotsToCoeffs;
    CKKSLinearTransformPlan plan;
    plan.use_qp_evaluator = s2c ? P.getBootstrapS2CUseQpLinearTransform() : P.getBootstrapC2SUseQpLinearTransform();
    const auto factors = fft_factors(log_slots, !s2c);
    const auto merge = merge_depths(log_slots, nf, lit.type, lit.merge_depths);
    const double stage_scale = s2c ? std::pow(lit.transform_scale, 1.0 / (double)nf)
                                   : std::pow(lit.transform_scale / (double)slots, 1.0 / (double)nf);
    size_t level = lit.level;
    for (size_t s = 0; s < nf; ++s) {
        Diagonals m = stage_diagonals(log_slots, factors, merge, lit.type, s, stage_scale);
        if (s2c && s == 0) select_inputs(m, slots, lit.input_selector_start, lit.input_selector_size);
        const double mx = max_abs(m);
        if (!(mx > 0.0))
            throw std::runtime_error(s2c ? "ckks_bootstrap_dft: generated empty S2C stage"
                                         : "ckks_bootstrap_dft: generated empty C2S stage");
        CKKSBootstrapDftMatrixLiteral sl = lit;
        sl.level = level;
        sl.plaintext_scale = stage_plaintext_scale(c, lit, sched[s], level, mx);
        sl.rescale_after_each_stage = sched[s].ends_group;
        plan.stages.push_back(encode_stage(c, sl, m, s + 1 == nf));
        if (sched[s].ends_group && s + 1 < nf) {
            if (level == 0) throw std::runtime_error("ckks_bootstrap_dft: invalid exhausted stage level");
            --level;
        }
    }
    return plan;
}

namespace {
// linear_transform_stage_slot_count (ckks_linear_transform.cpp:52-65)
size_t stage_slot_count(const CKKSParams& P, size_t slot_count) {
    const size_t full = P.getSlots(), s = slot_count == 0 ? full : slot_count;
    if (s == 0 || s > full || full % s != 0 || !std::has_single_bit(s))
        throw std::invalid_argument("ckks_linear_transform: invalid logical stage slot count");
    return s;
}
}  // namespace

CKKSCiphertext ckks_apply_linear_transform_owned_stage_qp(const CKKSContext& c, const CKKSCiphertext& ct,
                                                          const CKKSLinearTransformOwnedStage& st) {
    const auto& P = c.getParams();
    const size_t N = P.getN(), level = ct.getLevel(), K = P.getPModuli().size();
    if (st.qp_limb_count != level + 1 + K || st.qp_p_limb_count != K)
        throw std::invalid_argument("ckks_linear_transform: stage plaintexts do not match the ciphertext level");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense || ct.getN() != N || ct.getNumPolys() != 2)
        throw std::invalid_argument("ckks_linear_transform: invalid QP evaluator ciphertext input");
    const size_t n = st.diagonal_indices.size();
    if (n == 0 || st.diagonal_qp_plaintexts.size() != n) throw std::invalid_argument("ckks_linear_transform: empty stage");
    if (st.baby_step_count == 0) throw std::invalid_argument("ckks_linear_transform: baby-step count must be positive");
    const size_t slots = stage_slot_count(P, st.slot_count);
    for (size_t d : st.diagonal_indices)
        if (d >= slots) throw std::invalid_argument("ckks_linear_transform: diagonal index exceeds slot count");
    // the rotation keys the stage needs
    std::vector<int64_t> rots;
    std::vector<const uint64_t*> keys;
    std::vector<char> seen(std::max(slots, st.baby_step_count) + 1, 0);
    auto need = [&](size_t r) {
        if (r == 0 || seen[r]) return;
        seen[r] = 1;
        if (!c.hasRotationKey((int64_t)r))
            throw std::invalid_argument("ckks_linear_transform: missing rotation key for " + std::to_string(r));
        rots.push_back((int64_t)r);
        keys.push_back(c.getRotationKeyDevice((int64_t)r));
    };
    for (size_t d : st.diagonal_indices) {
        need(d % st.baby_step_count);
        need((d / st.baby_step_count) * st.baby_step_count);
    }
    std::vector<int64_t> didx(st.diagonal_indices.begin(), st.diagonal_indices.end());
    std::vector<const uint64_t*> dptr;
    for (const auto& d : st.diagonal_qp_plaintexts) dptr.push_back(d.data());
    CKKSCiphertext out(c, 2, 0.0L, level, ct.getSecretOwner(), ct.getMessageEncodingState(), ct.getBatchSize());
    if (ct.getBatchSize() == 1) {
        mi_check(mi_linear_transform_qp(c.handle(), (int)level, ct.polyData(1), ct.polyData(0), (int)n, didx.data(),
                                        dptr.data(), st.diagonal_qp_plaintexts_montgomery ? 1 : 0,
                                        (int)st.baby_step_count, (int)slots, (int)rots.size(), rots.data(),
                                        keys.data(), P.montgomery ? 1 : 0, out.polyData(1), out.polyData(0),
                                        c.stream()),
                 "ckks_linear_transform");
    } else {
        const size_t B = ct.getBatchSize();
        std::vector<const uint64_t*> ia(B), ib(B);
        std::vector<uint64_t*> oa(B), ob(B);
        for (size_t m = 0; m < B; ++m) {
            ia[m] = ct.polyData(1, m);
            ib[m] = ct.polyData(0, m);
            oa[m] = out.polyData(1, m);
            ob[m] = out.polyData(0, m);
        }
        mi_check(mi_linear_transform_qp_batch(c.handle(), (int)level, (int)B, ia.data(), ib.data(), (int)n,
                                              didx.data(), dptr.data(), st.diagonal_qp_plaintexts_montgomery ? 1 : 0,
                                              (int)st.baby_step_count, (int)slots, (int)rots.size(), rots.data(),
                                              keys.data(), P.montgomery ? 1 : 0, oa.data(), ob.data(), c.stream()),
                 "ckks_linear_transform");
    }
    out.setScale((double)(ct.getScale() * (long double)st.plaintext_scale));
    if (st.rescale_after) ckks_rescale(c, out);
    out.setMessageEncodingState(st.output_encoding_state);
    return out;
}

namespace {

// The Q evaluator (ckks_apply_linear_transform_bsgs, ckks_linear_transform.cpp:
// 818-996, and ckks_apply_linear_transform_owned_stage_q, :998-1172) with the
// checks of validate_stage_plaintexts (:758-814) and
// ensure_linear_transform_rotation_keys (:76-118).  diags[t] points to a
// level+1 limb plaintext; diag_scale[t] is its scale.
CKKSCiphertext apply_q_stage(const CKKSContext& c, const CKKSCiphertext& ct, const std::vector<const uint64_t*>& diags,
                             const std::vector<long double>& diag_scale, const std::vector<bool>& diag_mont,
                             std::span<const size_t> indices, size_t slot_count, size_t bs, bool rescale_after,
                             CKKSMessageEncodingState out_state) {
    const auto& P = c.getParams();
    const size_t N = P.getN(), slots = stage_slot_count(P, slot_count), level = ct.getLevel();
    const size_t n = diags.size();
    if (n == 0) throw std::invalid_argument("ckks_linear_transform: empty diagonal set");
    if (bs == 0) throw std::invalid_argument("ckks_linear_transform: baby-step count must be positive");
    if (!indices.empty() && indices.size() != n)
        throw std::invalid_argument("ckks_linear_transform: diagonal index count mismatch");
    if (ct.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_linear_transform: expected a dense-secret ciphertext");
    if (ct.getN() != N || ct.getNumPolys() != 2)
        throw std::invalid_argument("ckks_linear_transform: invalid ciphertext shape");
    if (level > P.getMaxLevel()) throw std::invalid_argument("ckks_linear_transform: invalid ciphertext level");
    std::vector<int64_t> idx(n);
    std::vector<char> seen(indices.empty() ? 0 : slots, 0);
    for (size_t t = 0; t < n; ++t) {
        if (!indices.empty()) {
            if (indices[t] >= slots)
                throw std::invalid_argument("ckks_linear_transform: diagonal index exceeds slot count");
            if (seen[indices[t]]) throw std::invalid_argument("ckks_linear_transform: duplicate diagonal index");
            seen[indices[t]] = 1;
        }
        idx[t] = (int64_t)(indices.empty() ? t : indices[t]);
    }
    for (size_t t = 1; t < n; ++t)
        if (diag_mont[t] != diag_mont[0])
            throw std::invalid_argument("ckks_linear_transform: diagonals mix Montgomery and normal form");
    // rotation keys: giant count as compile_linear_transform_stage_metadata (:720-756)
    const size_t giants = indices.empty() ? (n + bs - 1) / bs : (slots + bs - 1) / bs;
    std::vector<char> baby(bs, 0), giant(giants, 0);
    for (size_t t = 0; t < n; ++t) {
        baby[(size_t)idx[t] % bs] = 1;
        giant[(size_t)idx[t] / bs] = 1;
    }
    std::vector<int64_t> rots;
    std::vector<const uint64_t*> keys;
    for (size_t r = 1; r < bs; ++r)
        if (baby[r]) {
            if (!c.hasRotationKey((int64_t)r))
                throw std::invalid_argument("ckks_linear_transform: missing baby-step rotation key");
            rots.push_back((int64_t)r);
            keys.push_back(c.getRotationKeyDevice((int64_t)r));
        }
    for (size_t g = 1; g < giants; ++g)
        if (giant[g]) {
            const int64_t r = (int64_t)(g * bs);
            if (!c.hasRotationKey(r))
                throw std::invalid_argument("ckks_linear_transform: missing giant-step rotation key");
            rots.push_back(r);
            keys.push_back(c.getRotationKeyDevice(r));
        }
    // scale of the first term of the first giant step in use
    size_t first = 0;
    for (size_t t = 1; t < n; ++t)
        if ((size_t)idx[t] / bs < (size_t)idx[first] / bs) first = t;
    // terms of one giant step keep their order, so the first one listed wins
    for (size_t t = 0; t < n; ++t)
        if ((size_t)idx[t] / bs == (size_t)idx[first] / bs) {
            first = t;
            break;
        }
    const size_t slot_arg = indices.empty() ? giants * bs : slots;
    CKKSCiphertext out(c, 2, 0.0L, level, ct.getSecretOwner(), ct.getMessageEncodingState(), ct.getBatchSize());
    if (ct.getBatchSize() == 1) {
        mi_check(mi_linear_transform_q(c.handle(), (int)level, ct.polyData(1), ct.polyData(0), (int)n, idx.data(),
                                       diags.data(), diag_mont[0] ? 1 : 0, (int)bs, (int)slot_arg, (int)rots.size(),
                                       rots.data(), keys.data(), P.montgomery ? 1 : 0, out.polyData(1),
                                       out.polyData(0), c.stream()),
                 "ckks_linear_transform");
    } else {
        const size_t B = ct.getBatchSize();
        std::vector<const uint64_t*> ia(B), ib(B);
        std::vector<uint64_t*> oa(B), ob(B);
        for (size_t m = 0; m < B; ++m) {
            ia[m] = ct.polyData(1, m);
            ib[m] = ct.polyData(0, m);
            oa[m] = out.polyData(1, m);
            ob[m] = out.polyData(0, m);
        }
        mi_check(mi_linear_transform_q_batch(c.handle(), (int)level, (int)B, ia.data(), ib.data(), (int)n, idx.data(),
                                             diags.data(), diag_mont[0] ? 1 : 0, (int)bs, (int)slot_arg,
                                             (int)rots.size(), rots.data(), keys.data(), P.montgomery ? 1 : 0,
                                             oa.data(), ob.data(), c.stream()),
                 "ckks_linear_transform");
    }
    out.setScale((double)((long double)ct.getScale() * diag_scale[first]));
    if (rescale_after) ckks_rescale(c, out);
    out.setMessageEncodingState(out_state);
    return out;
}

}  // namespace

CKKSCiphertext ckks_apply_linear_transform_bsgs(const CKKSContext& c, const CKKSCiphertext& ct,
                                                const CKKSLinearTransformStage& stage) {
    const size_t level = ct.getLevel();
    std::vector<const uint64_t*> d;
    std::vector<long double> sc;
    std::vector<bool> mont;
    for (const auto& e : stage.diagonals) {
        if (e.getN() != c.getParams().getN() || e.getLevel() != level)
            throw std::invalid_argument("ckks_linear_transform: invalid diagonal plaintext metadata");
        d.push_back(e.data());
        sc.push_back(e.getScale());
        mont.push_back(e.isMontgomeryForm());
    }
    return apply_q_stage(c, ct, d, sc, mont, stage.diagonal_indices, stage.slot_count, stage.baby_step_count,
                         stage.rescale_after, stage.output_encoding_state);
}

CKKSCiphertext ckks_apply_linear_transform_owned_stage_q(const CKKSContext& c, const CKKSCiphertext& ct,
                                                         const CKKSLinearTransformOwnedStage& st) {
    const size_t level = ct.getLevel();
    std::vector<const uint64_t*> d;
    std::vector<long double> sc;
    std::vector<bool> mont;
    if (!st.diagonals.empty()) {
        for (const auto& e : st.diagonals) {
            if (e.getN() != c.getParams().getN() || e.getLevel() != level)
                throw std::invalid_argument("ckks_linear_transform: invalid diagonal plaintext metadata");
            d.push_back(e.data());
            sc.push_back(e.getScale());
            mont.push_back(e.isMontgomeryForm());
        }
    } else {
        // the first level+1 limbs of the QP-basis diagonals
        if (st.level != level || st.qp_limb_count != level + 1 + c.getParams().getPModuli().size())
            throw std::invalid_argument("ckks_linear_transform: invalid diagonal plaintext metadata");
        for (const auto& b : st.diagonal_qp_plaintexts) {
            d.push_back(b.data());
            sc.push_back((long double)st.plaintext_scale);
            mont.push_back(st.diagonal_qp_plaintexts_montgomery);
        }
    }
    return apply_q_stage(c, ct, d, sc, mont, std::span<const size_t>(st.diagonal_indices), st.slot_count,
                         st.baby_step_count, st.rescale_after, st.output_encoding_state);
}

CKKSCiphertext ckks_detail::apply_linear_transform_plan_from(const CKKSContext& c, const CKKSCiphertext& in,
                                                            const CKKSLinearTransformPlan& plan) {
    if (plan.stages.empty()) throw std::invalid_argument("ckks_linear_transform: empty transform plan");
    // every stage reads its input and writes a new ciphertext, so the first
    // one reads the caller's ciphertext in place (no copy of it)
    const auto stage = [&](const CKKSCiphertext& x, const CKKSLinearTransformOwnedStage& st) {
        // should_use_qp_linear_transform (ckks_linear_transform.cpp:67-74)
        if (!st.diagonal_qp_plaintexts.empty() && plan.use_qp_evaluator)
            return ckks_apply_linear_transform_owned_stage_qp(c, x, st);
        return ckks_apply_linear_transform_owned_stage_q(c, x, st);
    };
    CKKSCiphertext ct = stage(in, plan.stages.front());
    for (size_t i = 1; i < plan.stages.size(); ++i) ct = stage(ct, plan.stages[i]);
    return ct;
}

CKKSCiphertext ckks_apply_linear_transform_plan(const CKKSContext& c, CKKSCiphertext ct,
                                                const CKKSLinearTransformPlan& plan) {
    return ckks_detail::apply_linear_transform_plan_from(c, ct, plan);
}
