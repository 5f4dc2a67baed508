// Chebyshev-basis polynomial evaluation of an encrypted input
// (ckks_evalchebyshev_preconditioned, ckks_evalpoly.cpp:1209-1256).
//
// The schedule is the reference's: a power basis T_1..T_{2^k} built by
// T_{a+b} = 2 T_a T_b - T_{|a-b|} with exact scale tracking (seed^e over a
// product of moduli, ckks_evalpoly.cpp:24-141, 838-966), a recursive
// Paterson-Stockmeyer split planned on simulated (level, scale) pairs
// (:549-836), leaves evaluated as scale-aware scalar multiply-adds (:460-530,
// 998-1058) and the leaves combined by giant steps (:1060-1151).  Every
// constant is rounded in long double exactly as there, so the ciphertext
// limbs come out identical; the ring operations themselves run on the GPU.
#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <optional>
#include <sstream>

#include "ckks.hpp"

namespace {

using ld = long double;

size_t bit_length(size_t v) { return v ? (size_t)(std::numeric_limits<size_t>::digits - __builtin_clzll(v)) : 0; }

// seed^seed_pow / prod_i q_i^den[i]
struct ScaleTerm {
    uint64_t seed = 1;
    int seed_pow = 0;
    std::vector<int> den;

    static ScaleTerm of_integer_scale(double v, size_t nmod) {
        if (!std::isfinite(v) || v < 0.0) throw std::invalid_argument("ckks_evalchebyshev: invalid exact scale seed");
        const ld r = std::round((ld)v);
        const ld tol = std::max<ld>(1.0L, std::abs((ld)v)) * 0x1p-20L;
        if (std::abs((ld)v - r) > tol)
            throw std::invalid_argument("ckks_evalchebyshev: non-integral initial scale is unsupported");
        if (r > (ld)std::numeric_limits<uint64_t>::max())
            throw std::invalid_argument("ckks_evalchebyshev: initial scale exceeds uint64_t");
        ScaleTerm s;
        s.seed = (uint64_t)r;
        s.seed_pow = 1;
        s.den.assign(nmod, 0);
        return s;
    }
    ld value(const MyVector<uint64_t>& q) const {
        ld v = 1.0L;
        for (int i = 0; i < seed_pow; ++i) v *= (ld)seed;
        for (size_t i = 0; i < den.size(); ++i)
            for (int e = 0; e < den[i]; ++e) v /= (ld)q[i];
        return v;
    }
    double value_d(const MyVector<uint64_t>& q) const { return (double)value(q); }
};

ScaleTerm scale_product(const ScaleTerm& a, const ScaleTerm& b) {
    ScaleTerm o;
    o.seed = a.seed;
    o.seed_pow = a.seed_pow + b.seed_pow;
    o.den = a.den;
    if (o.den.size() < b.den.size()) o.den.resize(b.den.size(), 0);
    for (size_t i = 0; i < b.den.size(); ++i) o.den[i] += b.den[i];
    return o;
}

void scale_divide_by_modulus(ScaleTerm& s, size_t idx) {
    if (s.den.size() < idx + 1) s.den.resize(idx + 1, 0);
    s.den[idx] += 1;
}

uint64_t scale_ratio_floor(const ScaleTerm& num, const ScaleTerm& den, const MyVector<uint64_t>& q) {
    if (num.seed != den.seed) throw std::invalid_argument("ckks_evalchebyshev: mixed exact-scale seeds are unsupported");
    ld r = 1.0L;
    const int d = num.seed_pow - den.seed_pow;
    for (int i = 0; i < d; ++i) r *= (ld)num.seed;
    for (int i = 0; i < -std::min(d, 0); ++i) r /= (ld)num.seed;
    for (size_t i = 0; i < q.size(); ++i) {
        const int ne = i < num.den.size() ? num.den[i] : 0;
        const int de = i < den.den.size() ? den.den[i] : 0;
        const int diff = de - ne;
        for (int e = 0; e < diff; ++e) r *= (ld)q[i];
        for (int e = 0; e < -diff; ++e) r /= (ld)q[i];
    }
    const ld f = std::floor(r);
    if (!std::isfinite(f) || f < 0.0L || f > (ld)std::numeric_limits<uint64_t>::max())
        throw std::runtime_error("ckks_evalchebyshev: exact-scale ratio exceeds uint64_t");
    return (uint64_t)f;
}

size_t chebyshev_degree(const std::vector<double>& c) {
    if (c.empty()) throw std::invalid_argument("ckks_evalpoly: polynomial must have at least one coefficient");
    for (double v : c)
        if (!std::isfinite(v)) throw std::invalid_argument("ckks_evalpoly: polynomial coefficients must be finite");
    for (size_t i = c.size(); i > 0; --i)
        if (c[i - 1] != 0.0) return i - 1;
    return 0;
}

size_t ceil_log2(size_t d) { return d <= 1 ? 0 : bit_length(d - 1); }
size_t eval_depth(size_t d) { return d <= 1 ? 0 : bit_length(d) - 1; }

// power-of-two split of the PS recursion (the cheaper of the two neighbours)
size_t ps_split_log(size_t log_degree) {
    size_t ls = log_degree >> 1;
    const size_t cost_a = (size_t{1} << ls) + (size_t{1} << (log_degree - ls)) + log_degree - ls - 3;
    const size_t cost_b = (size_t{1} << (ls + 1)) + (size_t{1} << (log_degree - ls - 1)) + log_degree - ls - 4;
    if (cost_a > cost_b) ++ls;
    return ls;
}

// T_n = 2 T_a T_b - T_{|a-b|}
std::pair<size_t, size_t> chebyshev_split(size_t n) {
    if (n == 0) throw std::invalid_argument("ckks_evalchebyshev: cannot split degree zero");
    if ((n & (n - 1)) == 0) return {n >> 1, n >> 1};
    const size_t k = bit_length(n - 1) - 1;
    return {(size_t{1} << k) - 1, n + 1 - (size_t{1} << k)};
}

uint64_t scaled_constant_mod_q(double value, ld scale, uint64_t q) {
    const ld scaled = (ld)value * scale;
    if (!std::isfinite(scaled)) throw std::invalid_argument("ckks_evalpoly: non-finite scaled constant");
    ld res = std::fmod(std::round(scaled), (ld)q);
    if (res < 0.0L) res += (ld)q;
    uint64_t out = (uint64_t)res;
    if (out >= q) out %= q;
    return out;
}

bool near_integer(double v, double rel = 0x1p-20) {
    if (!std::isfinite(v)) return false;
    return std::abs(v - std::round(v)) <= rel * std::max(1.0, std::abs(v));
}

bool scales_match(double a, double b, double rel = 0x1p-45) {
    if (!std::isfinite(a) || !std::isfinite(b)) return false;
    const ld x = a, y = b;
    return std::abs(x - y) <= (ld)rel * std::max<ld>({1.0L, std::abs(x), std::abs(y)});
}

// ---- ciphertext helpers over the C-ABI
void zero_polys(const CKKSContext& c, CKKSCiphertext& ct, size_t from) {
    ckks_detail::zero_polys(c, ct, from, "ckks_evalpoly");
}

CKKSCiphertext zero_ct(const CKKSContext& c, size_t level, double scale, CKKSSecretOwner o,
                       CKKSMessageEncodingState st, size_t batch) {
    CKKSCiphertext ct(c, 2, scale, level, o, st, batch);
    zero_polys(c, ct, 0);
    return ct;
}

void add_scaled_constant(const CKKSContext& c, CKKSCiphertext& ct, double constant, ld scale) {
    if (constant == 0.0) return;
    const auto& q = c.getParams().getModuli();
    std::vector<uint64_t> r(ct.getLevel() + 1);
    for (size_t l = 0; l <= ct.getLevel(); ++l) r[l] = scaled_constant_mod_q(constant, scale, q[l]);
    ckks_detail::b_list_op(c, MI_LIST_ADD_SCALAR, ct, r.data(), "ckks_evalpoly");
}

size_t levels_per_rescale(const CKKSContext& c) { return std::log2(c.getParams().getScale()) <= 64.0 ? 1 : 2; }

ld rescale_divisor(const CKKSContext& c, size_t level) {
    const auto& q = c.getParams().getModuli();
    const size_t n = levels_per_rescale(c);
    if (level + 1 < n) throw std::invalid_argument("ckks_evalchebyshev: insufficient level for rescale modulus product");
    ld s = 1.0L;
    for (size_t i = 0; i < n; ++i) s *= (ld)q[level - i];
    return s;
}

CKKSCiphertext aligned(const CKKSContext& c, CKKSCiphertext ct, size_t level, double scale, size_t polys) {
    ckks_drop_to_level_inplace(ct, level);
    ct.setScale(scale);
    if (polys < ct.getNumPolys()) throw std::invalid_argument("ckks_evalchebyshev: cannot shrink ciphertext degree by padding");
    const size_t had = ct.getNumPolys();
    ct.resizePolys(polys);
    zero_polys(c, ct, had);
    return ct;
}

// A Chebyshev leaf as a linear form over the power basis, built on the host
// by the reference's operation sequence (evaluate_polynomial_from_power_basis,
// ckks_evalpoly.cpp:998-1049): its ciphertext metadata (level, scale, poly
// count) follows every step exactly, its residues per limb follow the data —
// the constant, each MulThenAdd term's rounded scalar, every integer
// scale-up of the accumulator applied to what it holds so far — and one
// mi_scaled_sum then writes the leaf.  All steps are exact mod q, so the
// leaf's residues equal the step-by-step evaluation's.
struct LeafForm {
    const CKKSContext& c;
    const MyVector<uint64_t>& q;
    size_t level = 0;
    ld scale = 0.0L;
    size_t polys = 2;
    std::vector<uint64_t> cst;                 // per limb
    bool has_cst = false;
    std::vector<const CKKSCiphertext*> terms;
    std::vector<std::vector<uint64_t>> r;      // per term, per limb

    static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
        return (uint64_t)((unsigned __int128)a * b % m);
    }
    // ckks_drop_to_level_inplace + setScale + pad_ciphertext_with_zero_polys
    void align(size_t lv, double sc, size_t target_polys) {
        if (lv > level) throw std::invalid_argument("ckks_drop_to_level_inplace: cannot raise ciphertext level");
        level = lv;
        scale = sc;
        if (target_polys < polys) throw std::invalid_argument("ckks_evalchebyshev: cannot shrink ciphertext degree by padding");
        polys = target_polys;
    }
    // mul_integer: the accumulator times round(scalar), checked as there
    void mul_integer(ld scalar, ld out_scale) {
        if (!std::isfinite(scalar) || scalar < 0.0L) throw std::invalid_argument("ckks_evalchebyshev: invalid scalar scale-up");
        const ld rr = std::round(scalar);
        if (rr < 1.0L || std::abs(scalar - rr) > std::max<ld>(1.0L, std::abs(scalar)) * 0x1p-20L)
            throw std::invalid_argument("ckks_evalchebyshev: non-integral scalar scale-up");
        if (rr > (ld)std::numeric_limits<uint64_t>::max())
            throw std::invalid_argument("ckks_evalchebyshev: scalar scale-up is too large");
        const uint64_t sv = (uint64_t)rr;
        if (sv != 1)
            for (size_t l = 0; l <= level; ++l) {
                const uint64_t m = sv % q[l];
                if (has_cst) cst[l] = mulmod(cst[l], m, q[l]);
                for (auto& t : r) t[l] = mulmod(t[l], m, q[l]);
            }
        scale = (double)out_scale;
    }
    // raise_scale_to
    void raise_scale_to(ld target) {
        const double cur = (double)scale;
        const double tgt = (double)target;
        const double m = std::max({1.0, std::abs(cur), std::abs(tgt)});
        if (std::isfinite(cur) && std::isfinite(tgt) && std::abs(cur - tgt) <= 0x1p-45 * m) {
            scale = (double)target;
            return;
        }
        if (scale > target)
            throw std::invalid_argument("ckks_evalchebyshev: cannot scale down ciphertext during scale alignment");
        mul_integer(target / scale, target);
    }
    // accumulator += scalar * op0 with the scales reconciled
    // (multhenadd_real_scalar_scale_aware_inplace, ckks_evalpoly.cpp:461-524)
    void mul_then_add(const CKKSCiphertext& op0, double scalar) {
        if (scalar == 0.0) return;
        const size_t lv = std::min(op0.getLevel(), level);
        const double out_scale = (double)scale;
        align(lv, out_scale, std::max(op0.getNumPolys(), polys));
        const double a_scale = (double)op0.getScale();
        ld s_rlwe = 0.0L;
        if (scales_match(a_scale, (double)scale)) {
            if (near_integer(scalar)) {
                s_rlwe = 1.0L;
            } else {
                s_rlwe = rescale_divisor(c, lv);
                raise_scale_to(scale * s_rlwe);
            }
        } else if (a_scale < scale) {
            s_rlwe = scale / a_scale;
        } else {
            std::ostringstream os;
            os << "ckks_evalchebyshev: MulThenAdd requires op0.Scale <= opOut.Scale (op0=" << a_scale
               << ", opOut=" << (double)scale << ", level=" << lv << ")";
            throw std::invalid_argument(os.str());
        }
        const ld term = a_scale * s_rlwe;
        if (!scales_match((double)term, (double)scale)) throw std::runtime_error("ckks_evalchebyshev: MulThenAdd scale mismatch");
        // fused_multiply_by_real_scalar_then_add_inplace (:352-398): + round(scalar s_rlwe) op0
        std::vector<uint64_t> rt(level + 1);
        for (size_t l = 0; l <= level; ++l) rt[l] = scaled_constant_mod_q(scalar, s_rlwe, q[l]);
        terms.push_back(&op0);
        r.push_back(std::move(rt));
    }
};

// ---- simulated (level, scale) planning
struct SimOp {
    size_t level = 0;
    ld scale = 0.0L;
};

struct Sim {
    const MyVector<uint64_t>& q;
    void rescale(SimOp& o) const {
        o.scale /= (ld)q[o.level];
        --o.level;
    }
    SimOp mul(const SimOp& a, const SimOp& b) const { return SimOp{std::min(a.level, b.level), a.scale * b.scale}; }
    std::pair<size_t, ld> baby(bool lead, size_t level, ld scale) const {
        ld s = scale;
        if (lead) s *= (ld)q[level];
        return {level, s};
    }
    std::pair<size_t, ld> giant(bool lead, size_t level, ld scale, ld xpow_scale) const {
        const uint64_t qi = lead ? q[level] : q[level + 1];
        return {level + 1, scale * (ld)qi / xpow_scale};
    }
};

void sim_power(size_t n, std::map<size_t, SimOp>& pb, const Sim& sim) {
    if (n < 2 || pb.count(n)) return;
    auto [a, b] = chebyshev_split(n);
    sim_power(a, pb, sim);
    sim_power(b, pb, sim);
    pb[n] = sim.mul(pb.at(a), pb.at(b));
    sim.rescale(pb.at(n));
}

struct Poly {
    std::vector<double> c;
    size_t max_degree = 0;
    bool lead = true;
    size_t level = 0;
    ld scale = 0.0L;
    bool even = false, odd = false;
    size_t degree() const { return c.empty() ? 0 : c.size() - 1; }
};

// p = q * T_split + r in the Chebyshev basis
std::pair<Poly, Poly> divide_by_chebyshev(const Poly& p, size_t split) {
    const size_t deg = p.degree();
    if (split < (deg >> 1)) throw std::invalid_argument("ckks_evalchebyshev: invalid recursive PS split");
    Poly quo, rem;
    quo.c.assign(deg - split + 1, 0.0);
    rem.c.assign(split, 0.0);
    for (size_t i = 0; i < split && i < p.c.size(); ++i) rem.c[i] = p.c[i];
    if (split < p.c.size()) quo.c[0] = p.c[split];
    for (size_t i = split + 1, j = 1; i <= deg; ++i, ++j) {
        const double v = p.c[i];
        if (v == 0.0) continue;
        quo.c[i - split] += 2.0 * v;
        rem.c[split - j] -= v;
    }
    quo.max_degree = p.max_degree;
    quo.lead = p.lead;
    quo.even = p.even;
    quo.odd = p.odd;
    rem.max_degree = p.max_degree == deg ? (split == 0 ? 0 : split - 1) : p.max_degree - (deg - split + 1);
    rem.lead = false;
    rem.even = p.even;
    rem.odd = p.odd;
    return {std::move(quo), std::move(rem)};
}

std::pair<std::vector<Poly>, SimOp> plan_leaves(size_t log_split, size_t target_level, const Poly& p,
                                                const std::map<size_t, SimOp>& pb, ld out_scale, const Sim& sim) {
    const size_t deg = p.degree();
    const size_t base = size_t{1} << log_split;
    if (deg < base) {
        if (p.lead && log_split > 1 &&
            p.max_degree > ((size_t{1} << bit_length(p.max_degree)) - (size_t{1} << (log_split - 1))))
            return plan_leaves(ps_split_log(bit_length(deg)), target_level, p, pb, out_scale, sim);
        Poly leaf = p;
        auto [lv, sc] = sim.baby(leaf.lead, target_level, out_scale);
        leaf.level = lv;
        leaf.scale = sc;
        return {std::vector<Poly>{std::move(leaf)}, SimOp{lv, sc}};
    }
    size_t next = base;
    while (next < (deg >> 1) + 1) next <<= 1;
    const SimOp& xpow = pb.at(next);
    auto [quo, rem] = divide_by_chebyshev(p, next);
    auto [lv_new, sc_new] = sim.giant(p.lead, target_level, out_scale, xpow.scale);
    auto [qv, res] = plan_leaves(log_split, lv_new, quo, pb, sc_new, sim);
    sim.rescale(res);
    res = sim.mul(res, xpow);
    auto [rv, tmp] = plan_leaves(log_split, target_level, rem, pb, res.scale, sim);
    if (!scales_match((double)res.scale, (double)tmp.scale))
        throw std::runtime_error("ckks_evalchebyshev: simulated PS scale mismatch");
    qv.insert(qv.end(), rv.begin(), rv.end());
    return {std::move(qv), res};
}

// ---- power basis on the device
class Powers {
public:
    Powers(const CKKSContext& c, const CKKSCiphertext& x) : c_(c), q_(c.getParams().getModuli()) {
        ct_.emplace(1, x.clone());
        sc_.emplace(1, ScaleTerm::of_integer_scale((double)x.getScale(), q_.size()));
    }
    void gen(size_t n) {
        if (n < 2 || ct_.count(n)) return;
        if (gen_unrescaled(n)) rescale_entry(n);
    }
    const CKKSCiphertext& at(size_t n) const { return ct_.at(n); }

private:
    void rescale_entry(size_t n) {
        CKKSCiphertext& v = ct_.at(n);
        const size_t lv = v.getLevel();
        ckks_rescale(c_, v);
        scale_divide_by_modulus(sc_.at(n), lv);
        v.setScale(sc_.at(n).value_d(q_));
    }
    bool gen_unrescaled(size_t n) {
        if (ct_.count(n)) return false;
        const auto [a, b] = chebyshev_split(n);
        const bool ra = gen_unrescaled(a);
        const bool rb = gen_unrescaled(b);
        if (ra) rescale_entry(a);
        if (rb) rescale_entry(b);
        const CKKSCiphertext& lhs_src = ct_.at(a);
        const size_t level = a == b ? lhs_src.getLevel() : std::min(lhs_src.getLevel(), ct_.at(b).getLevel());
        // the reference's clone of T_a (and T_b) at `level` followed by the
        // in-place product (ckks_evalpoly.cpp:908-921): here the product of
        // both operands read at `level` is written into a new ciphertext
        CKKSCiphertext out = ckks_detail::mul_into_at_level(c_, lhs_src, a == b ? nullptr : &ct_.at(b), level);
        const ScaleTerm prod = scale_product(sc_.at(a), sc_.at(b));
        ckks_relin_hybrid_inplace(c_, out);
        out.setScale(prod.value_d(q_));
        // T_n = 2 (T_a T_b) - T_d (d = |a - b| > 0) or 2 T_a^2 - 1 (d = 0): the
        // doubling (mul_integer by 2) and the subtraction of ratio * T_d (T_d
        // read at out's level) or the constant -1 at the product's scale, as
        // one scaled sum into a new ciphertext
        const size_t d = a > b ? a - b : b - a;
        if (d > 0) gen(d);
        ct_.insert_or_assign(n, doubled_minus(out, d, prod));
        sc_.insert_or_assign(n, prod);
        return true;
    }

    CKKSCiphertext doubled_minus(const CKKSCiphertext& out, size_t d, const ScaleTerm& prod) {
        const size_t L = out.getLevel() + 1, P = out.getNumPolys(), B = out.getBatchSize();
        CKKSCiphertext res(c_, P, prod.value_d(q_), out.getLevel(), out.getSecretOwner(), out.getMessageEncodingState(),
                           B);
        std::vector<uint64_t*> dst;
        for (size_t k = 0; k < P; ++k)
            for (size_t m = 0; m < B; ++m) dst.push_back(res.polyData(k, m));
        const size_t terms = d > 0 ? 2 : 1;
        std::vector<const uint64_t*> in(terms * P * B, nullptr);
        std::vector<uint64_t> r(terms * L), cst(L, 0);
        for (size_t k = 0; k < P; ++k)
            for (size_t m = 0; m < B; ++m) in[k * B + m] = out.polyData(k, m);
        for (size_t l = 0; l < L; ++l) r[l] = 2 % q_[l];
        if (d > 0) {
            const CKKSCiphertext& td = ct_.at(d);
            if (td.getLevel() < out.getLevel())
                throw std::invalid_argument("ckks_drop_to_level_inplace: cannot raise ciphertext level");
            if (td.getNumPolys() > P || td.getBatchSize() != B)
                throw std::invalid_argument("ckks_evalchebyshev: cannot shrink ciphertext degree by padding");
            const uint64_t ratio = scale_ratio_floor(prod, sc_.at(d), q_);
            for (size_t k = 0; k < td.getNumPolys(); ++k)
                for (size_t m = 0; m < B; ++m) in[P * B + k * B + m] = td.polyData(k, m);
            for (size_t l = 0; l < L; ++l) r[L + l] = (q_[l] - ratio % q_[l]) % q_[l];
        } else {
            for (size_t l = 0; l < L; ++l) cst[l] = scaled_constant_mod_q(-1.0, prod.value(q_), q_[l]);
        }
        mi_check(mi_scaled_sum(c_.handle(), (int)out.getLevel(), (int)(P * B), dst.data(), (int)terms, in.data(),
                               r.data(), d > 0 ? 0 : (int)B, cst.data(), c_.stream()),
                 "ckks_evalpoly");
        return res;
    }

    const CKKSContext& c_;
    const MyVector<uint64_t>& q_;
    std::map<size_t, CKKSCiphertext> ct_;
    std::map<size_t, ScaleTerm> sc_;
};

bool coeff_used(const Poly& p, size_t e) {
    return (!p.even && !p.odd) || ((e & 1) == 0 && p.even) || ((e & 1) == 1 && p.odd);
}

CKKSCiphertext eval_leaf(const CKKSContext& c, const Poly& p, const Powers& pw) {
    const size_t deg = p.degree();
    const CKKSCiphertext& x1 = pw.at(1);
    const size_t B = x1.getBatchSize();
    const auto& q = c.getParams().getModuli();
    LeafForm f{c, q};
    f.level = p.level;
    f.scale = (double)p.scale;
    f.cst.assign(p.level + 1, 0);
    if (p.even && p.c[0] != 0.0) {
        f.has_cst = true;
        for (size_t l = 0; l <= p.level; ++l) f.cst[l] = scaled_constant_mod_q(p.c[0], p.scale, q[l]);
    }
    if (deg > 0)
        for (size_t e = deg; e > 0; --e) {
            const double v = p.c[e];
            if (v == 0.0 || !coeff_used(p, e)) continue;
            f.align(p.level, (double)p.scale, f.polys);
            const CKKSCiphertext& op0 = pw.at(e);
            if (op0.getBatchSize() != B) throw std::invalid_argument("ckks_evalchebyshev: fused scalar multiply-add shape mismatch");
            f.mul_then_add(op0, v);
        }
    if (f.terms.empty()) {
        CKKSCiphertext res = zero_ct(c, p.level, (double)p.scale, x1.getSecretOwner(), x1.getMessageEncodingState(), B);
        if (f.has_cst) add_scaled_constant(c, res, p.c[0], p.scale);
        return res;
    }
    const size_t L = f.level, P = f.polys, T = f.terms.size();
    CKKSCiphertext res(c, P, f.scale, L, x1.getSecretOwner(), x1.getMessageEncodingState(), B);
    std::vector<uint64_t*> out;
    for (size_t k = 0; k < P; ++k)
        for (size_t m = 0; m < B; ++m) out.push_back(res.polyData(k, m));
    std::vector<const uint64_t*> in(T * P * B, nullptr);
    std::vector<uint64_t> resid(T * (L + 1));
    for (size_t t = 0; t < T; ++t) {
        const CKKSCiphertext& op0 = *f.terms[t];
        for (size_t k = 0; k < op0.getNumPolys(); ++k)
            for (size_t m = 0; m < B; ++m) in[t * P * B + k * B + m] = op0.polyData(k, m);
        for (size_t l = 0; l <= L; ++l) resid[t * (L + 1) + l] = f.r[t][l];
    }
    mi_check(mi_scaled_sum(c.handle(), (int)L, (int)(P * B), out.data(), (int)T, in.data(), resid.data(),
                           f.has_cst ? (int)B : 0, f.cst.data(), c.stream()),
             "ckks_evalpoly");
    return res;
}

// odd <- odd * x^(2^k) + even
void combine_giant(const CKKSContext& c, CKKSCiphertext& even, CKKSCiphertext& odd, const CKKSCiphertext& xpow) {
    if (odd.getNumPolys() == 3) ckks_relin_hybrid_inplace(c, odd);
    ckks_rescale(c, odd);
    if (xpow.getLevel() < odd.getLevel()) throw std::invalid_argument("ckks_drop_to_level_inplace: cannot raise ciphertext level");
    ckks_detail::mul_at_level(c, odd, xpow);  // x^(2^k) read at odd's level, as the reference's dropped copy
    const double es = even.getScale();
    CKKSCiphertext ea = aligned(c, std::move(even), odd.getLevel(), es, odd.getNumPolys());
    if (!scales_match(ea.getScale(), odd.getScale())) {
        std::ostringstream os;
        os << "ckks_evalchebyshev: EvaluateMonomial scale mismatch (even=" << (double)ea.getScale()
           << ", odd=" << (double)odd.getScale() << ", level=" << odd.getLevel() << ")";
        throw std::runtime_error(os.str());
    }
    const double common = std::max<double>(ea.getScale(), odd.getScale());
    ea.setScale(common);
    odd.setScale(common);
    ckks_add_inplace(c, odd, ea);
}

struct Leaf {
    size_t degree;
    CKKSCiphertext value;
};

CKKSCiphertext eval_plan(const CKKSContext& c, const std::vector<Poly>& leaves, const Powers& pw) {
    const size_t n = leaves.size();
    std::vector<std::optional<Leaf>> st(n);
    for (size_t i = 0; i < n; ++i) st[n - i - 1] = Leaf{leaves[i].degree(), eval_leaf(c, leaves[i], pw)};
    while (st.size() != 1) {
        std::vector<int> step(st.size(), 0);
        for (size_t i = 0; i < st.size(); ++i) {
            if (i == st.size() - 1) step[i] = 2;
            else if (st[i]->degree == st[i + 1]->degree) {
                step[i] = 1;
                ++i;
            }
        }
        for (size_t i = 0; i < st.size(); ++i) {
            if (step[i] == 2) {
                st[i]->degree = st[i - 1]->degree;
            } else if (step[i] == 1) {
                const size_t d = size_t{1} << bit_length(st[i]->degree);
                combine_giant(c, st[i]->value, st[i + 1]->value, pw.at(d));
                st[i + 1]->degree = 2 * d - 1;
                st[i].reset();
                ++i;
            }
        }
        size_t w = 0;
        for (size_t i = 0; i < st.size(); ++i)
            if (st[i].has_value()) {
                if (w != i) st[w] = std::move(st[i]);
                ++w;
            }
        st.resize(w);
    }
    CKKSCiphertext r = std::move(st.front()->value);
    if (r.getNumPolys() == 3) ckks_relin_hybrid_inplace(c, r);
    ckks_rescale(c, r);
    return r;
}

}  // namespace

CKKSCiphertext ckks_evalchebyshev_preconditioned(const CKKSContext& c, const CKKSCiphertext& x,
                                                 const std::vector<double>& coeffs, const CKKSEvalPolyOptions& opt) {
    const size_t deg = chebyshev_degree(coeffs);
    const double target = opt.target_scale > 0.0 ? opt.target_scale : (double)x.getScale();
    if (x.getN() != c.getParams().getN()) throw std::invalid_argument("ckks_evalpoly: ciphertext/context size mismatch");
    if (x.getNumPolys() != 2) throw std::invalid_argument("ckks_evalpoly: expected a 2-polynomial ciphertext");
    if (x.getSecretOwner() != CKKSSecretOwner::Dense)
        throw std::invalid_argument("ckks_evalpoly: polynomial evaluation requires a dense-secret ciphertext");
    if (x.getMessageEncodingState() != CKKSMessageEncodingState::Slots)
        throw std::invalid_argument("ckks_evalpoly: polynomial evaluation expects slot-encoded input");
    if (!std::isfinite(target) || target <= 0.0) throw std::invalid_argument("ckks_evalpoly: invalid scale");
    if (x.getLevel() < ceil_log2(deg))
        throw std::invalid_argument(
            "ckks_evalchebyshev_preconditioned: ciphertext level is too small for Chebyshev evaluation depth");
    if (deg == 0) {
        CKKSCiphertext out =
            zero_ct(c, x.getLevel(), target, x.getSecretOwner(), x.getMessageEncodingState(), x.getBatchSize());
        add_scaled_constant(c, out, coeffs[0], target);
        return out;
    }
    if (x.getLevel() < eval_depth(deg))
        throw std::invalid_argument("ckks_evalchebyshev: insufficient level for polynomial depth");
    Poly p;
    p.c.assign(coeffs.begin(), coeffs.begin() + deg + 1);
    p.max_degree = deg;
    p.lead = true;
    p.even = true;
    p.odd = true;
    for (size_t i = 1; i < coeffs.size(); i += 2) p.even = p.even && coeffs[i] == 0.0;
    for (size_t i = 0; i < coeffs.size(); i += 2) p.odd = p.odd && coeffs[i] == 0.0;
    const size_t log_deg = bit_length(deg);
    const size_t log_split = ps_split_log(log_deg);

    Powers pw(c, x);
    pw.gen(size_t{1} << (log_deg - 1));
    for (size_t i = (size_t{1} << log_split) - 1; i > 2; --i)
        if (coeff_used(p, i)) pw.gen(i);

    const auto& q = c.getParams().getModuli();
    Sim sim{q};
    std::map<size_t, SimOp> pb;
    pb[1] = SimOp{x.getLevel(), (ld)x.getScale()};
    sim_power(size_t{1} << log_deg, pb, sim);
    auto leaves = plan_leaves(log_split, x.getLevel() - eval_depth(deg), p, pb, (ld)target, sim).first;

    CKKSCiphertext out = eval_plan(c, leaves, pw);
    out.setScale(target);
    return out;
}
