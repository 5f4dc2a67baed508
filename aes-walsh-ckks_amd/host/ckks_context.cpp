// Parameters, device storage, ciphertext types, context construction and key
// generation of the host layer.
//
//   find_p_and_q_primes          random.cpp:66-131
//   CKKSParams                   ckks_params.cpp:13-110 (plain chain:
//                                q0_bit, then mult_depth x q_bit, ckks_params.cpp:213-269)
//   auxiliary-basis check        ckks_context.cpp:114-155
//   FFT twiddles, slot order     fft.cpp:37-80, ckks_context.cpp:56-68
//   switching keys               relinearization_key_generation.cpp:130-199,
//                                rotation_key_generation.cpp:90-199,
//                                conjugation_key_generation.cpp:59-167,
//                                hybrid_key_folding.hpp:50-77,
//                                hybrid_key_montgomery.hpp:12-36
// Key generation runs on the device through the C-ABI: the secret is
// NTT'd over Q u P, and b = P*s_src (digit limbs) + e - a*s_tgt, folded and
// put in Montgomery form with per-limb scalar products.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numbers>
#include <set>

#include <random>

#include "ckks.hpp"
#include "ckks_rng.hpp"

using ckks_host::GaussianSampler;
using ckks_host::Rng;

namespace {

typedef unsigned __int128 u128;

uint64_t mulm(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }
uint64_t powm(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = mulm(r, a, m);
        a = mulm(a, a, m);
        e >>= 1;
    }
    return r;
}
uint64_t invm(uint64_t a, uint64_t m) {
    __int128 t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        const __int128 q = r / nr;
        __int128 tmp = t - q * nt;
        t = nt;
        nt = tmp;
        tmp = r - q * nr;
        r = nr;
        nr = tmp;
    }
    if (r != 1) throw std::invalid_argument("ckks_context: modulus is not invertible");
    if (t < 0) t += m;
    return (uint64_t)t;
}
size_t bit_count(uint64_t m) { return m ? 64 - (size_t)__builtin_clzll(m) : 0; }

bool is_prime(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++s;
    }
    for (uint64_t a : bases) {
        if (n == a) return true;
        if (n % a == 0) return false;
        uint64_t x = powm(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool witness = true;
        for (int r = 1; r < s && witness; ++r) {
            x = mulm(x, x, n);
            if (x == n - 1) witness = false;
        }
        if (witness) return false;
    }
    return true;
}

size_t reverse_bits(size_t v, size_t width) {
    size_t r = 0;
    for (size_t i = 0; i < width; ++i) {
        r = (r << 1) | (v & 1);
        v >>= 1;
    }
    return r;
}

size_t log2_exact(size_t v) {
    size_t l = 0;
    while ((size_t{1} << l) < v) ++l;
    return l;
}

// s(X) -> s(X^g) on signed coefficients (rotation_key_generation.cpp:21-55)
MyVector<int8_t> automorphism_coeffs(const MyVector<int8_t>& s, uint64_t g) {
    const size_t N = s.size();
    MyVector<int8_t> out(N, 0);
    for (size_t i = 0; i < N; ++i) {
        const size_t mapped = (size_t)((u128)i * g % (2 * N));
        if (mapped >= N) out[mapped - N] = (int8_t)-s[i];
        else out[mapped] = s[i];
    }
    return out;
}

}  // namespace

void mi_check(int rc, const char* fn) {
    if (rc == MI_OK) return;
    const std::string msg = std::string(fn) + ": " + mi_last_error();
    if (rc == MI_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// ---------------------------------------------------------------- primes
std::pair<std::vector<uint64_t>, std::vector<uint64_t>> find_p_and_q_primes(
    const std::vector<uint64_t>& p_bits, const std::vector<uint64_t>& q_bits, size_t N) {
    std::map<uint64_t, uint64_t> cursor;  // bit size -> next k
    std::set<uint64_t> taken;
    std::vector<uint64_t> p, q;
    const size_t total = p_bits.size() + q_bits.size();
    for (size_t idx = 0; idx < total; ++idx) {
        const uint64_t bits = idx < p_bits.size() ? p_bits[idx] : q_bits[idx - p_bits.size()];
        if (bits < 2 || bits > 62)
            throw std::invalid_argument("find_p_and_q_primes: bit sizes must be in [2, 62]");
        auto it = cursor.find(bits);
        if (it == cursor.end()) it = cursor.emplace(bits, (uint64_t{1} << bits) / (2 * N)).first;
        const uint64_t floor_k = (uint64_t{1} << (bits - 1)) / (2 * N);
        uint64_t found = 0;
        while (it->second-- > floor_k) {
            const uint64_t cand = it->second * 2 * N + 1;
            if (taken.count(cand)) continue;
            if (is_prime(cand)) {
                found = cand;
                break;
            }
        }
        if (!found) throw std::runtime_error("find_p_and_q_primes: not enough NTT-friendly primes");
        taken.insert(found);
        (idx < p_bits.size() ? p : q).push_back(found);
    }
    return {p, q};
}

// ---------------------------------------------------------------- params
CKKSParams::CKKSParams(CKKSParamsConfig c) {
    if (c.N == 0 || (c.N & (c.N - 1)))
        throw std::invalid_argument("CKKSParams: N must be a nonzero power of two");
    if (c.slots != 0 && c.slots * 2 != c.N)
        throw std::invalid_argument("CKKSParams: slots must equal N / 2");
    if (c.q_bits.empty() && c.q_bit == 0)
        throw std::invalid_argument("CKKSParams: q_bit must be positive when q_bits is empty");
    if (c.dnum == 0) throw std::invalid_argument("CKKSParams: dnum must be positive");
    if (!c.q_bits.empty() && !c.q_mult_bits.empty())
        throw std::invalid_argument("CKKSParams: q_bits and q_mult_bits are mutually exclusive");
    if (!c.q_mult_bits.empty() && c.q_mult_bits.size() != c.mult_depth)
        throw std::invalid_argument("CKKSParams: q_mult_bits size must equal mult_depth");
    N = c.N;
    slots = N / 2;
    mult_depth = c.mult_depth;
    q0_bit = c.q0_bit;
    q_bit = c.q_bit;
    bts_keys = c.bts_keys;
    bootstrap = c.bootstrap_params;
    if (bts_keys && !bootstrap) throw std::runtime_error("CKKSParams: bootstrap parameters are required");
    dnum = c.dnum;
    p_bits = c.p_bits.empty() ? std::vector<uint64_t>(c.pmoduli_count, (uint64_t)c.p_bit) : c.p_bits;
    hamming_weight = c.secret_key_hamming_weight;
    sigma = c.sigma;
    scale = c.scale > 0.0 ? c.scale : std::exp2((double)c.q_bit);
    rotate_values = c.rotate_values;
    aes_walsh = c.aes_walsh_params;
    secure = c.secure;
    fold_hybrid_keyswitch = c.fold_hybrid_keyswitch;
    montgomery = c.montgomery;
    if (c.seed) {
        seed = *c.seed;
    } else {
        std::random_device rd;
        seed = ((uint64_t)rd() << 32) ^ (uint64_t)rd();
    }
    // chain: q0, S2C groups (bottom to top), mult limbs, EvalMod limbs, C2S
    // groups (ckks_params.cpp:213-296)
    size_t expect_depth = mult_depth;
    if (bts_keys) expect_depth += getBootstrapS2CDepth() + getBootstrapC2SDepth() + getBootstrapEvalModDepth();
    auto append_groups = [&](const CKKSBootstrapDftScaleSchedule& sched, size_t fallback_depth) {
        if (sched.empty()) {
            q_bits.insert(q_bits.end(), fallback_depth, (uint64_t)c.q_bit);
            return;
        }
        for (size_t r = 0; r < sched.size(); ++r) {
            const auto& g = sched[sched.size() - 1 - r];
            if (g.empty()) throw std::runtime_error("CKKSParams: DFT scale schedule groups must not be empty");
            uint64_t sum = 0;
            for (size_t v : g) sum += v;
            q_bits.push_back(sum);
        }
    };
    if (!c.q_bits.empty()) {
        q_bits = c.q_bits;
        depth = q_bits.size() - 1;
        if (c.depth != 0 && c.depth != depth)
            throw std::invalid_argument("CKKSParams: depth must match q_bits.size() - 1");
    } else {
        depth = c.depth == 0 ? expect_depth : c.depth;
        q_bits.push_back(c.q0_bit);
        if (bts_keys) append_groups(bootstrap->linear_transform_log_plaintext_scales[0], getBootstrapS2CDepth());
        if (c.q_mult_bits.empty()) q_bits.insert(q_bits.end(), mult_depth, (uint64_t)c.q_bit);
        else q_bits.insert(q_bits.end(), c.q_mult_bits.begin(), c.q_mult_bits.end());
        if (bts_keys) {
            const uint64_t eb = bootstrap->evalmod_bit_size == 0 ? c.q_bit : bootstrap->evalmod_bit_size;
            q_bits.insert(q_bits.end(), bootstrap->evalmod_depth, eb);
            append_groups(bootstrap->linear_transform_log_plaintext_scales[1], getBootstrapC2SDepth());
        }
        if (q_bits.size() != depth + 1) throw std::runtime_error("CKKSParams: Q bit schedule must match limb count");
    }
    if (depth != expect_depth)
        throw std::runtime_error("CKKSParams: depth must equal mult_depth plus requested bootstrap/S2C depths");
    if (p_bits.empty()) throw std::invalid_argument("CKKSParams: at least one P modulus is required");
    auto [pp, qq] = find_p_and_q_primes(p_bits, q_bits, N);
    moduli.assign(qq.begin(), qq.end());
    pmoduli.assign(pp.begin(), pp.end());
}

const CKKSBootstrapParams& CKKSParams::getBootstrapParams() const {
    if (!bts_keys) throw std::runtime_error("CKKSParams: bootstrap params are disabled because bts_keys is false");
    if (!bootstrap) throw std::runtime_error("CKKSParams: bootstrap parameters are required");
    return *bootstrap;
}

static size_t lt_depth(const CKKSBootstrapParams& b, size_t i) {
    const auto& s = b.linear_transform_log_plaintext_scales[i];
    return s.empty() ? b.linear_transform_depths[i] : s.size();
}

size_t CKKSParams::getAesWalshPackedSegmentCount() const {
    if (aes_walsh.packed_segment_count != 0) return aes_walsh.packed_segment_count;
    return aes_walsh.block_count == 0 ? 0 : slots / aes_walsh.block_count;
}

size_t CKKSParams::getBootstrapS2CDepth() const { return lt_depth(getBootstrapParams(), 0); }
size_t CKKSParams::getBootstrapC2SDepth() const { return lt_depth(getBootstrapParams(), 1); }

size_t CKKSParams::getBootstrapActiveSlotCount() const {
    const size_t a = getBootstrapParams().active_slot_count;
    if (a == 0) return slots;
    if (a & (a - 1)) throw std::runtime_error("CKKSParams: bootstrap active slot count must be a power of two");
    if (a > slots || slots % a != 0)
        throw std::runtime_error("CKKSParams: bootstrap active slot count must divide the slot count");
    return a;
}

size_t CKKSParams::getEvalModLogScaleBits() const {
    const auto& b = getBootstrapParams();
    return b.evalmod_bit_size == 0 ? q_bit : b.evalmod_bit_size;
}

// Smallest K with N * P(|sum of h uniform(-1/2,1/2)| > h/2 - K) below the
// target (the Irwin-Hall tail, binboot_k_bound.cpp:13-87); EvalMod uses K + 1.
size_t CKKSParams::getBootstrapEvalModK() const {
    const auto& b = getBootstrapParams();
    if (b.evalmod_k != 0) return b.evalmod_k;
    const long double fl = b.evalmod_k_failure_probability_log2;
    if (!std::isfinite(fl) || fl >= 0.0L)
        throw std::runtime_error("CKKSParams: automatic EvalMod K failure probability log2 must be finite and negative");
    const size_t h = b.sparse_secret_hamming_weight;
    const long double target = std::exp2(fl);
    if (h == 0) throw std::invalid_argument("binboot K bound: invalid coefficient count or Hamming weight");
    auto binom = [](size_t n, size_t k) {
        if (k > n) return 0.0L;
        k = std::min(k, n - k);
        long double o = 1.0L;
        for (size_t i = 1; i <= k; ++i) {
            o *= (long double)(n - k + i);
            o /= (long double)i;
        }
        return o;
    };
    auto cdf = [&](long double x) {
        if (x <= 0.0L) return 0.0L;
        if (x >= (long double)h) return 1.0L;
        const size_t up = (size_t)std::floor(x);
        long double sum = 0.0L;
        for (size_t k = 0; k <= up; ++k) {
            const long double t = binom(h, k) * std::pow(x - (long double)k, (long double)h);
            sum += (k & 1) == 0 ? t : -t;
        }
        return sum / std::tgamma((long double)h + 1.0L);
    };
    const size_t max_k = (h + 1) / 2;
    for (size_t K = 0; K <= max_k; ++K) {
        const long double half = (long double)h / 2.0L;
        const long double tail = (long double)K >= half ? 0.0L : 2.0L * cdf(half - (long double)K);
        const long double total = std::min(1.0L, (long double)N * tail);
        if (total <= target) return K + 1;
    }
    return max_k + 1;
}

// ---------------------------------------------------------------- storage
DeviceBuffer::DeviceBuffer(mi_context* ctx, size_t words) : ctx_(ctx), words_(words) {
    void* p = nullptr;
    mi_check(mi_alloc(ctx, words * sizeof(uint64_t), &p), "DeviceBuffer");
    ptr_ = (uint64_t*)p;
}

void DeviceBuffer::reset() {
    if (ptr_) (void)mi_free(ctx_, ptr_);
    ptr_ = nullptr;
    words_ = 0;
    ctx_ = nullptr;
}

static void upload(const CKKSContext& c, uint64_t* dst, const uint64_t* src, size_t words) {
    mi_check(mi_memcpy(c.handle(), dst, src, words * 8, MI_COPY_H2D, c.stream()), "upload");
}
static void download(const CKKSContext& c, uint64_t* dst, const uint64_t* src, size_t words) {
    mi_check(mi_memcpy(c.handle(), dst, src, words * 8, MI_COPY_D2H, c.stream()), "download");
}
static void copy_d2d(const CKKSContext& c, uint64_t* dst, const uint64_t* src, size_t words) {
    mi_check(mi_memcpy(c.handle(), dst, src, words * 8, MI_COPY_D2D, c.stream()), "copy");
}

CKKSCiphertext::CKKSCiphertext(const CKKSContext& context, size_t num_polys, long double s, size_t lv,
                               CKKSSecretOwner o, CKKSMessageEncodingState st, size_t b)
    : ctx(&context), N(context.getParams().getN()), scale(s), level(lv), secret_owner(o),
      message_encoding_state(st), batch(b) {
    if (lv > context.getParams().getMaxLevel())
        throw std::invalid_argument("CKKSCiphertext: invalid level");
    if (b == 0) throw std::invalid_argument("CKKSCiphertext: batch size must be positive");
    for (size_t i = 0; i < num_polys * b; ++i) polys.emplace_back(context.handle(), (lv + 1) * N);
}

CKKSCiphertext::CKKSCiphertext(const CKKSContext& context, MyVector<MyVector<uint64_t>> host,
                               long double s, size_t lv, CKKSSecretOwner o, CKKSMessageEncodingState st)
    : CKKSCiphertext(context, host.size(), s, lv, o, st) {
    for (size_t i = 0; i < host.size(); ++i) setPoly(i, host[i]);
}

CKKSCiphertext CKKSCiphertext::clone() const { return cloneAtLevel(level); }

CKKSCiphertext CKKSCiphertext::cloneAtLevel(size_t lv) const {
    if (lv > level) throw std::invalid_argument("ckks_drop_to_level_inplace: target level above current level");
    const size_t level = lv;
    CKKSCiphertext out(*ctx, getNumPolys(), scale, level, secret_owner, message_encoding_state, batch);
    if (polys.empty()) return out;
    std::vector<uint64_t*> dst(polys.size());
    std::vector<const uint64_t*> src(polys.size());
    for (size_t i = 0; i < polys.size(); ++i) {
        dst[i] = out.polys[i].data();
        src[i] = polys[i].data();
    }
    mi_check(mi_poly_list_op(ctx->handle(), MI_LIST_COPY, (int)level, (int)polys.size(), dst.data(), src.data(),
                             nullptr, ctx->stream()),
             "ckks_ciphertext");
    return out;
}

CKKSCiphertext CKKSCiphertext::stack(MyVector<CKKSCiphertext> members) {
    if (members.empty()) throw std::invalid_argument("CKKSCiphertext::stack: no members");
    const CKKSCiphertext& r = members.front();
    size_t total = 0;
    for (const auto& m : members) {
        if (m.ctx != r.ctx || m.N != r.N || m.level != r.level || m.getNumPolys() != r.getNumPolys() ||
            m.secret_owner != r.secret_owner || m.message_encoding_state != r.message_encoding_state ||
            m.scale != r.scale)
            throw std::invalid_argument("CKKSCiphertext::stack: members differ in context or metadata");
        total += m.batch;
    }
    const size_t np = r.getNumPolys();
    CKKSCiphertext out(*r.ctx, 0, r.scale, r.level, r.secret_owner, r.message_encoding_state, total);
    out.polys.resize(np * total);
    size_t base = 0;
    for (auto& m : members) {
        for (size_t k = 0; k < np; ++k)
            for (size_t j = 0; j < m.batch; ++j) out.polys[k * total + base + j] = std::move(m.polys[k * m.batch + j]);
        base += m.batch;
        m.polys.clear();
    }
    return out;
}

MyVector<CKKSCiphertext> CKKSCiphertext::unstack() {
    const size_t np = getNumPolys();
    MyVector<CKKSCiphertext> out;
    out.reserve(batch);
    for (size_t m = 0; m < batch; ++m) {
        out.emplace_back(*ctx, 0, scale, level, secret_owner, message_encoding_state, 1);
        out.back().polys.resize(np);
        for (size_t k = 0; k < np; ++k) out.back().polys[k] = std::move(polys[k * batch + m]);
    }
    polys.clear();
    batch = 1;
    return out;
}

void CKKSCiphertext::setLevel(size_t l) {
    for (auto& p : polys)
        if ((l + 1) * N > p.words()) throw std::invalid_argument("CKKSCiphertext: cannot raise level in place");
    level = l;
}

MyVector<uint64_t> CKKSCiphertext::getPoly(size_t idx) const {
    if (batch != 1) throw std::invalid_argument("CKKSCiphertext: getPoly of a batch needs a member index");
    return getPoly(idx, 0);
}

MyVector<uint64_t> CKKSCiphertext::getPoly(size_t idx, size_t member) const {
    MyVector<uint64_t> h((level + 1) * N);
    download(*ctx, h.data(), polyData(idx, member), h.size());
    return h;
}

void CKKSCiphertext::setPoly(size_t idx, const MyVector<uint64_t>& host) {
    if (host.size() != (level + 1) * N) throw std::invalid_argument("CKKSCiphertext: poly size mismatch");
    upload(*ctx, polyData(idx), host.data(), host.size());
}

uint64_t* CKKSCiphertext::polyData(size_t idx) {
    if (batch != 1) throw std::invalid_argument("CKKSCiphertext: this operation has no batched form");
    return polyData(idx, 0);
}
const uint64_t* CKKSCiphertext::polyData(size_t idx) const {
    if (batch != 1) throw std::invalid_argument("CKKSCiphertext: this operation has no batched form");
    return polyData(idx, 0);
}
uint64_t* CKKSCiphertext::polyData(size_t idx, size_t member) {
    if (idx >= getNumPolys() || member >= batch) throw std::out_of_range("CKKSCiphertext: poly index");
    return polys[idx * batch + member].data();
}
const uint64_t* CKKSCiphertext::polyData(size_t idx, size_t member) const {
    if (idx >= getNumPolys() || member >= batch) throw std::out_of_range("CKKSCiphertext: poly index");
    return polys[idx * batch + member].data();
}

std::vector<uint64_t*> CKKSCiphertext::allPolyData() {
    std::vector<uint64_t*> v(polys.size());
    for (size_t i = 0; i < polys.size(); ++i) v[i] = polys[i].data();
    return v;
}
std::vector<const uint64_t*> CKKSCiphertext::allPolyData() const {
    std::vector<const uint64_t*> v(polys.size());
    for (size_t i = 0; i < polys.size(); ++i) v[i] = polys[i].data();
    return v;
}

void CKKSCiphertext::addPoly() {
    for (size_t m = 0; m < batch; ++m) polys.emplace_back(ctx->handle(), (level + 1) * N);
}

void CKKSCiphertext::resizePolys(size_t count) {
    while (getNumPolys() > count)
        for (size_t m = 0; m < batch; ++m) polys.pop_back();
    while (getNumPolys() < count) addPoly();
}

CKKSEncoding::CKKSEncoding(const CKKSContext& context, MyVector<uint64_t> pt, long double s, size_t lv,
                           bool mont)
    : ctx(&context), N(context.getParams().getN()), scale(s), level(lv), montgomery_form(mont),
      buf(context.handle(), (lv + 1) * context.getParams().getN()) {
    if (pt.size() != (lv + 1) * N) throw std::invalid_argument("CKKSEncoding: invalid plaintext buffer size");
    upload(context, buf.data(), pt.data(), pt.size());
}

CKKSEncoding::CKKSEncoding(const CKKSContext& context, DeviceBuffer pt, long double s, size_t lv, bool mont)
    : ctx(&context), N(context.getParams().getN()), scale(s), level(lv), montgomery_form(mont),
      buf(std::move(pt)) {
    if (buf.words() < (lv + 1) * N) throw std::invalid_argument("CKKSEncoding: invalid plaintext buffer size");
}

MyVector<uint64_t> CKKSEncoding::getPlaintext() const {
    MyVector<uint64_t> h((level + 1) * N);
    download(*ctx, h.data(), buf.data(), h.size());
    return h;
}

// ---------------------------------------------------------------- context
int64_t normalize_rotation_value(size_t N, int64_t rotation) {
    const int64_t slots = (int64_t)(N >> 1);
    if (slots == 0) throw std::invalid_argument("ckks_context: invalid slot count");
    int64_t r = rotation % slots;
    if (r < 0) r += slots;
    return r;
}

uint64_t rotation_automorphism_index(size_t N, int64_t rotation) {
    return powm(5, (uint64_t)normalize_rotation_value(N, rotation), 2 * N);
}

uint64_t conjugation_automorphism_index(size_t N) { return 2 * N - 1; }

// Stage twiddles of the special FFT, bit-reversed per stage, with the
// embedding twist psi^k merged into stage k.
MyVector<Complex> twiddle_ifft_generation(size_t N) {
    MyVector<Complex> tw(N, Complex(0.0, 0.0));
    const double ang = 2.0 * std::numbers::pi / (double)N;
    const Complex w_inv(std::cos(ang), std::sin(ang));
    MyVector<Complex> stage(N >> 1);
    stage[0] = Complex(1.0, 0.0);
    for (size_t i = 1; i < (N >> 1); ++i) stage[i] = stage[i - 1] * w_inv;
    for (size_t m = 1; m < N; m <<= 1) {
        const size_t step = N / (m << 1);
        const size_t bits = log2_exact(m);
        for (size_t i = 0; i < m; ++i) tw[m + i] = stage[reverse_bits(i, bits) * step];
    }
    const double pang = std::numbers::pi / (double)N;
    const Complex psi(std::cos(pang), std::sin(pang));
    MyVector<Complex> psi_pows(N);
    psi_pows[0] = Complex(1.0, 0.0);
    for (size_t j = 1; j < N; ++j) psi_pows[j] = psi_pows[j - 1] * psi;
    for (size_t m = N >> 1, k = 1; m > 0; m >>= 1, k <<= 1)
        for (size_t i = 0; i < m; ++i) tw[m + i] = tw[m + i] * psi_pows[k];
    return tw;
}

// Slot i sits at evaluation index brev((5^i mod 2N - 1) / 2).
MyVector<size_t> perm_table_generation(size_t N) {
    const size_t logN = log2_exact(N);
    MyVector<size_t> perm(N >> 1);
    uint64_t cur = 1;
    for (size_t i = 0; i < (N >> 1); ++i) {
        perm[i] = reverse_bits((size_t)((cur - 1) >> 1), logN);
        cur = (cur * 5) % (2 * N);
    }
    return perm;
}

CKKSContext::CKKSContext(CKKSParams p) : params(std::move(p)) {
    const size_t N = params.getN();
    const auto& q = params.getModuli();
    const auto& pm = params.getPModuli();
    const size_t nq = q.size(), dnum = params.getDnum();
    if (dnum == 0 || dnum > nq)
        throw std::invalid_argument("ckks_context: invalid hybrid key-switch partition count");
    const size_t alpha = (nq + dnum - 1) / dnum;
    size_t p_bits = 0, widest = 0;
    for (uint64_t v : pm) p_bits += bit_count(v);
    for (size_t s0 = 0; s0 < nq; s0 += alpha) {
        size_t b = 0;
        for (size_t i = s0; i < std::min(s0 + alpha, nq); ++i) b += bit_count(q[i]);
        widest = std::max(widest, b);
    }
    if (p_bits < widest)
        throw std::invalid_argument("ckks_context: hybrid key-switch P basis is smaller than the widest Q digit; dnum = " +
                                    std::to_string(dnum) + ", P bits = " + std::to_string(p_bits) +
                                    ", widest Q digit bits = " + std::to_string(widest) +
                                    "; add P limbs or increase dnum");
    K().key_digits = std::min(dnum, (nq + alpha - 1) / alpha);

    twiddle_ifft = twiddle_ifft_generation(N);
    twiddle_fft.resize(N);
    for (size_t i = 0; i < N; ++i) twiddle_fft[i] = std::conj(twiddle_ifft[i]);
    perm_table = perm_table_generation(N);

    mi_check(mi_context_create(&mi, N, q.data(), (int)nq, pm.data(), (int)pm.size(), (int)dnum),
             "ckks_context");
    try {
        mi_check(mi_stream_create(mi, &stream_), "ckks_context");
        {
            static_assert(sizeof(Complex) == 2 * sizeof(double), "complex<double> layout");
            std::vector<uint64_t> perm(perm_table.begin(), perm_table.end());
            mi_check(mi_set_encoding_tables(mi, reinterpret_cast<const double*>(twiddle_fft.data()), perm.data()),
                     "ckks_context");
            mi_check(mi_set_decoding_tables(mi, reinterpret_cast<const double*>(twiddle_ifft.data()), perm.data()),
                     "ckks_context");
        }
        Rng rng(params.getSeed());
        K().secret_coeffs = ckks_host::hamming_weight_ternary(N, params.getSecretKeyHammingWeight(), rng);
        K().secret_qp = secret_qp_from_coeffs(K().secret_coeffs);
        const size_t KL = nq + pm.size();
        // relinearisation: s^2 -> s
        DeviceBuffer s2(mi, KL * N);
        copy_d2d(*this, s2.data(), K().secret_qp.data(), KL * N);
        mi_check(mi_poly_op(mi, MI_OP_MUL, s2.data(), K().secret_qp.data(), (int)KL, 0, 1, (int64_t)(KL * N), stream()),
                 "ckks_context");
        K().relin_key = generate_switch_key(s2.data(), K().secret_qp.data(), 1);
        // conjugation: conj(s) -> s
        DeviceBuffer sc = secret_qp_from_coeffs(automorphism_coeffs(K().secret_coeffs, conjugation_automorphism_index(N)));
        K().conj_key = generate_switch_key(sc.data(), K().secret_qp.data(), 2);
        for (int64_t r : params.getRotateValues()) {
            const int64_t n = normalize_rotation_value(N, r);
            if (n != 0 && !K().rotation_keys.count(n)) addRotationKey(n);
        }
        // append_aes_walsh_rotation_values (ckks_context.cpp:274-305):
        // +-segment * block_count for every provisioned segment
        if (params.usesAesWalsh()) {
            const size_t blocks = params.getAesWalshBlockCount(), segments = params.getAesWalshPackedSegmentCount();
            if (blocks == 0) throw std::invalid_argument("ckks_context: AES Walsh block count must be positive");
            if (segments < 30)
                throw std::invalid_argument("ckks_context: AES Walsh packed segment count must be at least 30");
            if (segments > params.getSlots() / blocks)
                throw std::invalid_argument("ckks_context: AES Walsh block count exceeds packed slot capacity");
            for (size_t s = 1; s < segments; ++s)
                for (int64_t r : {(int64_t)(s * blocks), -(int64_t)(s * blocks)}) {
                    const int64_t n = normalize_rotation_value(N, r);
                    if (n != 0 && !K().rotation_keys.count(n)) addRotationKey(n);
                }
        }
        if (params.bts_keys) {
            build_bootstrap_material(rng);
            K().evalmod_coeffs = default_evalmod_polynomial(params);
        }
    } catch (...) {
        release_all();
        if (stream_) mi_stream_destroy(mi, stream_);
        mi_context_destroy(mi);
        throw;
    }
}

CKKSContext::KeyMaterial& CKKSContext::K() {
    if (owner_) throw std::runtime_error("ckks_context: a lane shares its owner's keys; change them on the owner");
    return keys_;
}

// A lane: the owner's parameters, its own device context (tables, pool,
// workspaces) and stream; keys, plans and polynomial read from the owner.
CKKSContext::CKKSContext(const CKKSContext& owner, int)
    : params(owner.params), owner_(&owner.keyOwner()), twiddle_fft(owner.twiddle_fft),
      twiddle_ifft(owner.twiddle_ifft), perm_table(owner.perm_table) {
    const auto& q = params.getModuli();
    const auto& pm = params.getPModuli();
    // the owner's keys and plans are complete before another stream reads them
    mi_check(mi_stream_synchronize(owner.handle(), owner.stream()), "ckks_context");
    mi_check(mi_context_create(&mi, params.getN(), q.data(), (int)q.size(), pm.data(), (int)pm.size(),
                               (int)params.getDnum()),
             "ckks_context");
    try {
        mi_check(mi_stream_create(mi, &stream_), "ckks_context");
        std::vector<uint64_t> perm(perm_table.begin(), perm_table.end());
        mi_check(mi_set_encoding_tables(mi, reinterpret_cast<const double*>(twiddle_fft.data()), perm.data()),
                 "ckks_context");
        mi_check(mi_set_decoding_tables(mi, reinterpret_cast<const double*>(twiddle_ifft.data()), perm.data()),
                 "ckks_context");
    } catch (...) {
        if (stream_) mi_stream_destroy(mi, stream_);
        mi_context_destroy(mi);
        throw;
    }
    owner_->lanes_.fetch_add(1);
}

std::unique_ptr<CKKSContext> CKKSContext::makeLane(const CKKSContext& owner) {
    return std::unique_ptr<CKKSContext>(new CKKSContext(owner, 0));
}

void CKKSContext::release_all() {
    enc_cache.clear();
    scaled_s2c.clear();
    if (owner_) return;  // a lane holds no keys
    K().s2c_plan.reset();
    K().c2s_plan.reset();
    K().d2s_key.reset();
    K().s2d_key.reset();
    K().low_secret_qp.reset();
    K().relin_key.reset();
    K().conj_key.reset();
    K().rotation_keys.clear();
    K().secret_qp.reset();
}

// Sparse secret and its switching keys, the S2C / C2S DFT plans and their
// rotation keys (ckks_context.cpp:157-233, 389-414, 603-705).
void CKKSContext::build_bootstrap_material(Rng& rng) {
    const size_t N = params.getN(), nq = params.getModuli().size();
    const auto& b = params.getBootstrapParams();
    const auto& q = params.getModuli();
    K().low_secret_coeffs = ckks_host::hamming_weight_ternary(N, b.sparse_secret_hamming_weight, rng);
    K().low_secret_qp = secret_qp_from_coeffs(K().low_secret_coeffs);
    K().d2s_key = generate_q0p0_key(K().secret_qp.data(), K().low_secret_qp.data(), 3);
    K().s2d_key = generate_switch_key(K().low_secret_qp.data(), K().secret_qp.data(), 4);

    const size_t s2c_depth = params.getBootstrapS2CDepth(), c2s_depth = params.getBootstrapC2SDepth();
    const size_t active = params.getBootstrapActiveSlotCount();
    if (s2c_depth > 0) {
        CKKSBootstrapDftMatrixLiteral lit;
        lit.type = CKKSBootstrapDftType::SlotsToCoeffs;
        lit.level = s2c_depth;
        lit.baby_step_count = b.linear_transform_bsgs_baby_step_counts[0];
        lit.log_bsgs_ratio = b.linear_transform_log_bsgs_ratios[0];
        lit.stage_count = s2c_depth;
        lit.log_plaintext_scale_groups = b.linear_transform_log_plaintext_scales[0];
        lit.merge_depths = b.linear_transform_merge_depths[0];
        lit.active_slots = active;
        lit.transform_scale = ((double)q[0] * 0.5) / params.getScale();
        lit.plaintext_scale = (double)q[s2c_depth];
        lit.rescale_after_each_stage = true;
        K().s2c_plan = std::make_unique<CKKSLinearTransformPlan>(ckks_generate_bootstrap_dft_plan(*this, lit));
    }
    if (c2s_depth > 0) {
        const size_t level = params.getDepth();
        if (c2s_depth > level) throw std::runtime_error("ckks_context: EvalMod C2S depth exceeds bootstrap depth");
        const double evalmod_k = (double)params.getBootstrapEvalModK();
        const int shift = (int)params.getEvalModLogScaleBits() - (int)params.getQ0Bit();
        K().c2s_transform_scale = std::ldexp(params.getScale(), shift) / (double)q[0] / evalmod_k;
        CKKSBootstrapDftMatrixLiteral lit;
        lit.type = CKKSBootstrapDftType::CoeffsToSlots;
        lit.level = level;
        lit.baby_step_count = b.linear_transform_bsgs_baby_step_counts[1];
        lit.log_bsgs_ratio = b.linear_transform_log_bsgs_ratios[1];
        lit.stage_count = c2s_depth;
        lit.log_plaintext_scale_groups = b.linear_transform_log_plaintext_scales[1];
        lit.merge_depths = b.linear_transform_merge_depths[1];
        lit.active_slots = active;
        lit.transform_scale = K().c2s_transform_scale / (double)(params.getSlots() / active);
        lit.plaintext_scale = (double)q[level] / params.getScale();
        lit.rescale_after_each_stage = true;
        K().c2s_plan = std::make_unique<CKKSLinearTransformPlan>(ckks_generate_bootstrap_dft_plan(*this, lit));
    }
    auto add_plan_rotations = [&](const CKKSLinearTransformPlan* plan) {
        if (!plan) return;
        for (const auto& st : plan->stages)
            for (size_t d : st.diagonal_indices) {
                const size_t baby = d % st.baby_step_count;
                if (baby) addRotationKey((int64_t)baby);
                if (d - baby) addRotationKey((int64_t)(d - baby));
            }
    };
    add_plan_rotations(K().s2c_plan.get());
    add_plan_rotations(K().c2s_plan.get());
    if (c2s_depth > 0)
        for (size_t r = active; r < params.getSlots(); r <<= 1) addRotationKey((int64_t)r);
    (void)nq;
}

// Dense-to-sparse switch key over {q0, p0} (secret_key_generation.cpp:302-364):
// b = (p0 mod q0) * s_src on q0, 0 on p0, + e - a * s_tgt; folded and in
// Montgomery form.  Layout [b(q0), b(p0), a(q0), a(p0)].
DeviceBuffer CKKSContext::generate_q0p0_key(const uint64_t* src_qp, const uint64_t* tgt_qp, uint64_t stream_id) const {
    const size_t N = params.getN(), nq = params.getModuli().size();
    const uint64_t q0 = params.getModuli()[0], p0 = params.getPModuli()[0];
    const int64_t W = (int64_t)N;
    uint64_t st = params.getSeed() ^ (0xA5A5A5A5ull * (stream_id + 1));
    Rng rng(ckks_host::splitmix64(st));
    GaussianSampler gauss(params.getSigma());
    DeviceBuffer key(mi, 4 * N), tmp(mi, 2 * N);
    uint64_t* bq = key.data();
    uint64_t* bp = bq + N;
    uint64_t* aq = bp + N;  // a(p0) follows at aq + N
    const uint64_t* sq_src = src_qp;
    const uint64_t* sq_tgt = tgt_qp;
    const uint64_t* sp_tgt = tgt_qp + nq * N;
    copy_d2d(*this, bq, sq_src, N);
    const uint64_t pm = p0 % q0;
    mi_check(mi_poly_scalar(mi, MI_OP_MUL, bq, &pm, 1, 0, 1, W, stream()), "ckks_context");
    mi_check(mi_memset_zero(mi, bp, N * 8, stream()), "ckks_context");
    std::vector<uint64_t> h(2 * N);
    for (size_t i = 0; i < N; ++i) h[i] = rng.below(q0);
    for (size_t i = 0; i < N; ++i) h[N + i] = rng.below(p0);
    upload(*this, aq, h.data(), 2 * N);
    for (size_t i = 0; i < N; ++i) {
        const int64_t e = gauss.sample(rng);
        h[i] = e < 0 ? q0 - (uint64_t)(-e) : (uint64_t)e;
        h[N + i] = e < 0 ? p0 - (uint64_t)(-e) : (uint64_t)e;
    }
    upload(*this, tmp.data(), h.data(), 2 * N);
    mi_check(mi_ntt_forward(mi, tmp.data(), 1, 1, W, 0, stream()), "ckks_context");
    mi_check(mi_ntt_forward(mi, tmp.data() + N, 1, 1, W, (int)nq, stream()), "ckks_context");
    mi_check(mi_poly_op(mi, MI_OP_ADD, bq, tmp.data(), 1, 0, 1, W, stream()), "ckks_context");
    mi_check(mi_poly_op(mi, MI_OP_ADD, bp, tmp.data() + N, 1, (int)nq, 1, W, stream()), "ckks_context");
    copy_d2d(*this, tmp.data(), aq, 2 * N);
    mi_check(mi_poly_op(mi, MI_OP_MUL, tmp.data(), sq_tgt, 1, 0, 1, W, stream()), "ckks_context");
    mi_check(mi_poly_op(mi, MI_OP_MUL, tmp.data() + N, sp_tgt, 1, (int)nq, 1, W, stream()), "ckks_context");
    mi_check(mi_poly_op(mi, MI_OP_SUB, bq, tmp.data(), 1, 0, 1, W, stream()), "ckks_context");
    mi_check(mi_poly_op(mi, MI_OP_SUB, bp, tmp.data() + N, 1, (int)nq, 1, W, stream()), "ckks_context");
    uint64_t fq = 1, fp = 1;
    if (params.fold_hybrid_keyswitch) fq = invm(pm, q0);
    if (params.montgomery) {
        fq = mulm(fq, (uint64_t)(((u128)1 << 64) % q0), q0);
        fp = mulm(fp, (uint64_t)(((u128)1 << 64) % p0), p0);
    }
    if (params.fold_hybrid_keyswitch || params.montgomery) {
        // b and a of each limb: q0 limbs at 0 and 2N, p0 limbs at N and 3N
        mi_check(mi_poly_scalar(mi, MI_OP_MUL, bq, &fq, 1, 0, 2, 2 * W, stream()), "ckks_context");
        mi_check(mi_poly_scalar(mi, MI_OP_MUL, bp, &fp, 1, (int)nq, 2, 2 * W, stream()), "ckks_context");
    }
    mi_check(mi_stream_synchronize(mi, stream()), "ckks_context");
    return key;
}

const uint64_t* CKKSContext::getDenseToSparseSwitchKeyDevice() const {
    if (!K().d2s_key.data()) throw std::runtime_error("ckks_context: dense-to-sparse switch key is unavailable");
    return K().d2s_key.data();
}
const uint64_t* CKKSContext::getSparseToDenseSwitchKeyDevice() const {
    if (!K().s2d_key.data()) throw std::runtime_error("ckks_context: sparse-to-dense switch key is unavailable");
    return K().s2d_key.data();
}
void CKKSContext::setDenseToSparseSwitchKey(const MyVector<uint64_t>& key) {
    const size_t N = params.getN();
    if (key.size() != 4 * N) throw std::invalid_argument("ckks_context: dense-to-sparse key size mismatch");
    DeviceBuffer d(mi, key.size());
    upload(*this, d.data(), key.data(), key.size());
    K().d2s_key = std::move(d);
}
void CKKSContext::setSparseToDenseSwitchKey(const MyVector<uint64_t>& key) { K().s2d_key = upload_key(key); }
MyVector<uint64_t> CKKSContext::downloadDenseToSparseSwitchKey() const {
    MyVector<uint64_t> h(4 * params.getN());
    download(*this, h.data(), getDenseToSparseSwitchKeyDevice(), h.size());
    return h;
}

const CKKSLinearTransformPlan& CKKSContext::getEvalModS2CPlan() const {
    if (!K().s2c_plan) throw std::runtime_error("ckks_context: EvalMod S2C plan is unavailable");
    return *K().s2c_plan;
}
const CKKSLinearTransformPlan& CKKSContext::getEvalModScaledC2SPlan() const {
    if (!K().c2s_plan) throw std::runtime_error("ckks_context: EvalMod C2S plan is unavailable");
    return *K().c2s_plan;
}

const CKKSLinearTransformPlan& CKKSContext::getScaledS2CPlan(uint64_t ratio) const {
    if (ratio == 0) throw std::invalid_argument("ckks EvalMod: invalid S2C scale ratio");
    if (ratio == 1) return getEvalModS2CPlan();
    std::lock_guard<std::mutex> g(scaled_mu);
    auto it = scaled_s2c.find(ratio);
    if (it != scaled_s2c.end()) return *it->second;
    const auto& b = params.getBootstrapParams();
    const size_t s2c_depth = params.getBootstrapS2CDepth();
    if (s2c_depth == 0) throw std::runtime_error("ckks EvalMod: scaled S2C depth must be positive");
    CKKSBootstrapDftMatrixLiteral lit;
    lit.type = CKKSBootstrapDftType::SlotsToCoeffs;
    lit.level = s2c_depth;
    lit.baby_step_count = b.linear_transform_bsgs_baby_step_counts[0];
    lit.log_bsgs_ratio = b.linear_transform_log_bsgs_ratios[0];
    lit.stage_count = s2c_depth;
    lit.log_plaintext_scale_groups = b.linear_transform_log_plaintext_scales[0];
    lit.merge_depths = b.linear_transform_merge_depths[0];
    lit.active_slots = params.getBootstrapActiveSlotCount();
    lit.transform_scale = ((double)params.getModuli()[0] * 0.5) / (params.getScale() * (double)ratio);
    lit.plaintext_scale = (double)params.getModuli()[s2c_depth];
    lit.rescale_after_each_stage = true;
    auto plan = std::make_unique<CKKSLinearTransformPlan>(ckks_generate_bootstrap_dft_plan(*this, lit));
    const CKKSLinearTransformPlan& ref = *plan;
    scaled_s2c.emplace(ratio, std::move(plan));
    return ref;
}

CKKSBootstrapParams aes_walsh15_bootstrap_params() {
    CKKSBootstrapParams b;
    b.sparse_secret_hamming_weight = 32;
    b.linear_transform_log_plaintext_scales = {CKKSBootstrapDftScaleSchedule{{18, 20, 23}},
                                               CKKSBootstrapDftScaleSchedule{{25, 25}, {25, 25}}};
    b.linear_transform_depths = {1, 2};
    b.evalmod_depth = 7;
    b.evalmod_bit_size = 38;
    b.evalmod_k_failure_probability_log2 = -16.0L;
    b.evalmod_polynomial_degree = 100;
    b.evalmod_double_angle = 0;
    b.evalmod_coefficient_prune_threshold = 1e-10;
    b.evalmod_discrete_interval_radius = std::nextafter(0.25, 0.0);
    b.evalmod_discrete_validation_points_per_interval = 128;
    b.linear_transform_log_bsgs_ratios = {1, 1};
    b.linear_transform_use_qp = {true, true};
    return b;
}

CKKSParamsConfig aes_walsh15_config() {
    CKKSParamsConfig c;
    c.secure = true;
    c.bts_keys = true;
    c.N = size_t{1} << 15;
    c.mult_depth = 2;
    c.q0_bit = 36;
    c.q_mult_bits = {42, 42};
    c.scale = std::exp2(35.0);
    c.dnum = 3;
    c.pmoduli_count = 4;
    c.p_bits = {61, 61, 61, 61};
    c.secret_key_hamming_weight = 256;
    c.aes_walsh_params.enabled = true;
    c.aes_walsh_params.block_count = 512;
    c.aes_walsh_params.packed_segment_count = 32;
    c.bootstrap_params = aes_walsh15_bootstrap_params();
    return c;
}

CKKSParams CKKSParams::AESWalsh15() { return CKKSParams(aes_walsh15_config()); }

namespace {

// make_evalmod_bootstrap_params (ckks_params.hpp:140-156) with the XBOOT
// Mod1 settings (mod1.CosDiscrete, no double angle, LogMessageRatio 1,
// discrete interval radius just below 1/4, 64 validation points, no pruning).
CKKSBootstrapParams xboot_bootstrap_params(size_t h, CKKSBootstrapDftScaleSchedule s2c,
                                           CKKSBootstrapDftScaleSchedule c2s, size_t evalmod_depth,
                                           size_t evalmod_bits, size_t K, size_t degree,
                                           std::array<int, 2> log_bsgs_ratios) {
    CKKSBootstrapParams b;
    b.sparse_secret_hamming_weight = h;
    b.linear_transform_depths = {s2c.size(), c2s.size()};
    b.linear_transform_log_plaintext_scales = {std::move(s2c), std::move(c2s)};
    b.evalmod_depth = evalmod_depth;
    b.evalmod_bit_size = evalmod_bits;
    b.evalmod_k = K;
    b.evalmod_polynomial_degree = degree;
    b.evalmod_double_angle = 0;
    b.evalmod_log_message_ratio = 1;
    b.evalmod_coefficient_prune_threshold = 0.0;
    b.evalmod_discrete_interval_radius = std::nextafter(0.25, 0.0);
    b.evalmod_discrete_validation_points_per_interval = 64;
    b.linear_transform_log_bsgs_ratios = log_bsgs_ratios;
    b.linear_transform_use_qp = {true, true};
    return b;
}

}  // namespace

// XBootParamAes13ReferenceConfig (ckks_aes_params.hpp:12-33, 83-109): XBOOT
// Table 2 Param-AES-13 at N = 2^14 (the reference marks it not secure under
// its fixed-h bound), dnum 13 over one 32-bit special prime.
CKKSParamsConfig xboot_aes13_reference_config() {
    CKKSParamsConfig c;
    c.secure = false;
    c.bts_keys = true;
    c.N = size_t{1} << 14;
    c.mult_depth = 3;
    c.q0_bit = 32;
    c.q_bit = 32;
    c.q_bits = {32, 28, 31, 31, 31, 32, 32, 32, 32, 32, 32, 28, 28};
    c.scale = std::exp2(31.0);
    c.dnum = 13;
    c.pmoduli_count = 1;
    c.p_bits = {32};
    c.secret_key_hamming_weight = 256;
    c.bootstrap_params = xboot_bootstrap_params(16, CKKSBootstrapDftScaleSchedule{{28}},
                                                CKKSBootstrapDftScaleSchedule{{28}, {28}}, 6, 32, 7, 60, {1, 2});
    return c;
}

// AESXBoot14SecureConfig (ckks_aes_params.hpp:111-148): every Q limb one bit
// smaller than Param-AES-13, secure.
CKKSParamsConfig aes_xboot14_secure_config() {
    CKKSParamsConfig c = xboot_aes13_reference_config();
    c.secure = true;
    c.q0_bit = 31;
    c.q_bit = 31;
    c.q_bits = {31, 27, 30, 30, 30, 31, 31, 31, 31, 31, 31, 27, 27};
    c.scale = std::exp2(30.0);
    c.bootstrap_params = xboot_bootstrap_params(16, CKKSBootstrapDftScaleSchedule{{27}},
                                                CKKSBootstrapDftScaleSchedule{{27}, {27}}, 6, 31, 7, 60, {1, 2});
    return c;
}

// AESXBoot15Config (ckks_aes_params.hpp:156-203): N = 2^15, dnum 5 over five
// 38-bit special primes, EvalMod degree 120 with K = 12.
CKKSParamsConfig aes_xboot15_config() {
    CKKSParamsConfig c;
    c.secure = true;
    c.bts_keys = true;
    c.N = size_t{1} << 15;
    c.mult_depth = 3;
    c.q0_bit = 38;
    c.q_bit = 38;
    c.q_bits = {38, 36, 36, 37, 37, 37, 38, 38, 38, 38, 38, 38, 38, 38, 38, 38, 38};
    c.scale = std::exp2(37.0);
    c.dnum = 5;
    c.pmoduli_count = 5;
    c.p_bits = {38, 38, 38, 38, 38};
    c.secret_key_hamming_weight = 256;
    c.bootstrap_params =
        xboot_bootstrap_params(32, CKKSBootstrapDftScaleSchedule{{36}, {36}},
                               CKKSBootstrapDftScaleSchedule{{38}, {38}, {38}, {38}}, 7, 38, 12, 120, {2, 1});
    return c;
}

// The Sparse1024 variants (ckks_aes_params.hpp:150-154, 205-209): 1024
// active bootstrap slots.
static CKKSParamsConfig sparse1024(CKKSParamsConfig c) {
    c.bootstrap_params->active_slot_count = 1024;
    return c;
}

CKKSParams CKKSParams::XBootParamAes13Reference() { return CKKSParams(xboot_aes13_reference_config()); }
CKKSParams CKKSParams::AESXBoot14Secure() { return CKKSParams(aes_xboot14_secure_config()); }
CKKSParams CKKSParams::AESXBoot14SecureSparse1024() { return CKKSParams(sparse1024(aes_xboot14_secure_config())); }
CKKSParams CKKSParams::AESXBoot15() { return CKKSParams(aes_xboot15_config()); }
CKKSParams CKKSParams::AESXBoot15Sparse1024() { return CKKSParams(sparse1024(aes_xboot15_config())); }

namespace {

// The reference's EvalMod fit for the parameters of its presets
// (ckks_evalmod_tables.inc, made by evalmod_fit.py at 256-bit precision).
struct EvalModFit {
    double K;
    size_t degree, double_angle;
    double interval_radius;
    size_t validation_points;
    double prune_threshold, interval_min, interval_max, max_error;
    std::vector<double> coeffs;
};

const std::vector<EvalModFit>& evalmod_fits() {
    static const std::vector<EvalModFit> fits = {
#include "ckks_evalmod_tables.inc"
    };
    return fits;
}

// automatic_binboot_discrete_interval_radius (ckks_params.hpp:34-45) behind
// getBootstrapEvalModDiscreteIntervalRadius (ckks_params.cpp:403-414)
double evalmod_interval_radius(const CKKSParams& P) {
    const auto& b = P.getBootstrapParams();
    const double r = b.evalmod_discrete_interval_radius;
    if (!std::isfinite(r) || r < 0.0 || r >= 0.25)
        throw std::runtime_error("CKKSParams: EvalMod discrete interval radius must be in [0, 0.25)");
    if (r > 0.0) return r;
    const size_t lm = b.evalmod_log_message_ratio;
    if (lm == 0) return 1.0 / 32.0;
    if (lm >= (size_t)std::numeric_limits<double>::max_exponent) return std::numeric_limits<double>::min();
    return std::min(std::ldexp(1.0, -(int)lm), std::nextafter(0.25, 0.0));
}

// getBootstrapEvalModDiscreteValidationPointsPerInterval (ckks_params.cpp:416-427)
size_t evalmod_validation_points(const CKKSParams& P) {
    const size_t n = P.getBootstrapParams().evalmod_discrete_validation_points_per_interval;
    if (n == 0) return 64;
    if (n < 2) throw std::runtime_error("CKKSParams: EvalMod discrete validation points per interval must be at least 2");
    return n;
}

const EvalModFit* find_reference_fit(const CKKSParams& P) {
    const auto& b = P.getBootstrapParams();
    const double K = (double)P.getBootstrapEvalModK();
    const double radius = evalmod_interval_radius(P);
    const size_t points = evalmod_validation_points(P);
    for (const auto& f : evalmod_fits())
        if (f.K == K && f.degree == b.evalmod_polynomial_degree && f.double_angle == b.evalmod_double_angle &&
            f.interval_radius == radius && f.validation_points == points &&
            f.prune_threshold == b.evalmod_coefficient_prune_threshold)
            return &f;
    return nullptr;
}

}  // namespace

bool evalmod_polynomial_is_reference_fit(const CKKSParams& params) { return find_reference_fit(params) != nullptr; }

// The reference's fit (binboot_han_ki_discrete_multi_interval_approximate,
// poly_approx.cpp:565-681) where its parameters are those of a preset; for
// other parameters without double-angle steps, the degree-d Chebyshev
// interpolant (points cos(pi (j + 1/2) / (d + 1))) of 0.5 - 0.5 cos(2 pi K u)
// on [-1, 1], odd terms zero; with double-angle steps there is no built-in
// polynomial (empty: evalmod_fit.py computes the reference's fit for any
// parameters, installed with CKKSContext::setEvalModPolynomial).
std::vector<double> default_evalmod_polynomial(const CKKSParams& params) {
    if (const EvalModFit* f = find_reference_fit(params)) return f->coeffs;
    if (params.getBootstrapEvalModDoubleAngle() != 0) return {};
    const size_t d = params.getBootstrapEvalModPolynomialDegree(), n = d + 1;
    const long double K = (long double)params.getBootstrapEvalModK();
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> f(n);
    for (size_t j = 0; j < n; ++j) {
        const long double x = std::cos(pi * ((long double)j + 0.5L) / (long double)n);
        f[j] = 0.5L - 0.5L * std::cos(2.0L * pi * K * x);
    }
    std::vector<double> c(n, 0.0);
    for (size_t k = 0; k < n; k += 2) {
        long double acc = 0.0L;
        for (size_t j = 0; j < n; ++j) acc += f[j] * std::cos(pi * (long double)k * ((long double)j + 0.5L) / (long double)n);
        c[k] = (double)(acc * (k == 0 ? 1.0L : 2.0L) / (long double)n);
    }
    return c;
}

void CKKSContext::setEvalModPolynomial(const std::vector<double>& c) {
    if (c.empty()) throw std::invalid_argument("ckks_context: empty EvalMod polynomial");
    K().evalmod_coeffs = c;
}

const std::vector<double>& CKKSContext::getEvalModPolynomial() const {
    if (K().evalmod_coeffs.empty())
        throw std::runtime_error(
            "ckks_context: EvalMod polynomial is unavailable: no built-in fit for these EvalMod parameters "
            "(double-angle steps); set the reference's fit with setEvalModPolynomial");
    return K().evalmod_coeffs;
}

CKKSContext::~CKKSContext() {
    release_all();
    if (owner_) owner_->lanes_.fetch_sub(1);
    if (mi && stream_) mi_stream_destroy(mi, stream_);
    if (mi) mi_context_destroy(mi);
}

DeviceBuffer CKKSContext::secret_qp_from_coeffs(const MyVector<int8_t>& coeffs) const {
    const size_t N = params.getN();
    std::vector<uint64_t> basis(params.getModuli().begin(), params.getModuli().end());
    basis.insert(basis.end(), params.getPModuli().begin(), params.getPModuli().end());
    const size_t KL = basis.size();
    std::vector<uint64_t> h(KL * N);
    for (size_t l = 0; l < KL; ++l)
        for (size_t i = 0; i < N; ++i) h[l * N + i] = coeffs[i] < 0 ? basis[l] - (uint64_t)(-coeffs[i]) : (uint64_t)coeffs[i];
    DeviceBuffer d(mi, KL * N);
    upload(*this, d.data(), h.data(), h.size());
    mi_check(mi_ntt_forward(mi, d.data(), 1, (int)KL, (int64_t)(KL * N), 0, stream()), "ckks_context");
    return d;
}

size_t CKKSContext::switchKeyWords() const {
    return K().key_digits * 2 * (params.getModuli().size() + params.getPModuli().size()) * params.getN();
}

DeviceBuffer CKKSContext::generate_switch_key(const uint64_t* src, const uint64_t* tgt, uint64_t stream_id) const {
    const size_t N = params.getN();
    const auto& q = params.getModuli();
    const auto& pm = params.getPModuli();
    const size_t nq = q.size(), np = pm.size(), KL = nq + np;
    const size_t alpha = params.getAlpha();
    std::vector<uint64_t> basis(q.begin(), q.end());
    basis.insert(basis.end(), pm.begin(), pm.end());
    std::vector<uint64_t> p_mod(KL), scale(KL);
    for (size_t l = 0; l < KL; ++l) {
        uint64_t acc = 1;
        for (uint64_t pv : pm) acc = mulm(acc, pv % basis[l], basis[l]);
        p_mod[l] = acc;
        uint64_t f = 1;
        if (params.fold_hybrid_keyswitch) {
            if (l < nq) {
                f = invm(acc, basis[l]);
            } else {
                uint64_t ph = 1;
                for (size_t j = 0; j < np; ++j)
                    if (j != l - nq) ph = mulm(ph, pm[j] % basis[l], basis[l]);
                f = invm(ph, basis[l]);
            }
        }
        if (params.montgomery) f = mulm(f, (uint64_t)(((u128)1 << 64) % basis[l]), basis[l]);
        scale[l] = f;
    }
    uint64_t st = params.getSeed() ^ (0xA5A5A5A5ull * (stream_id + 1));
    Rng rng(ckks_host::splitmix64(st));
    GaussianSampler gauss(params.getSigma());
    DeviceBuffer key(mi, switchKeyWords());
    DeviceBuffer tmp(mi, KL * N);
    std::vector<uint64_t> h(KL * N);
    const int64_t W = (int64_t)(KL * N);
    for (size_t d = 0; d < K().key_digits; ++d) {
        uint64_t* b = key.data() + (2 * d) * KL * N;
        uint64_t* a = b + KL * N;
        copy_d2d(*this, b, src, KL * N);
        std::vector<uint64_t> r(KL, 0);
        for (size_t l = d * alpha; l < std::min(d * alpha + alpha, nq); ++l) r[l] = p_mod[l];
        mi_check(mi_poly_scalar(mi, MI_OP_MUL, b, r.data(), (int)KL, 0, 1, W, stream()), "ckks_context");
        for (size_t l = 0; l < KL; ++l)
            for (size_t i = 0; i < N; ++i) h[l * N + i] = rng.below(basis[l]);
        upload(*this, a, h.data(), h.size());
        for (size_t i = 0; i < N; ++i) {
            const int64_t e = gauss.sample(rng);
            for (size_t l = 0; l < KL; ++l) h[l * N + i] = e < 0 ? basis[l] - (uint64_t)(-e) : (uint64_t)e;
        }
        upload(*this, tmp.data(), h.data(), h.size());
        mi_check(mi_ntt_forward(mi, tmp.data(), 1, (int)KL, W, 0, stream()), "ckks_context");
        mi_check(mi_poly_op(mi, MI_OP_ADD, b, tmp.data(), (int)KL, 0, 1, W, stream()), "ckks_context");
        copy_d2d(*this, tmp.data(), a, KL * N);
        mi_check(mi_poly_op(mi, MI_OP_MUL, tmp.data(), tgt, (int)KL, 0, 1, W, stream()), "ckks_context");
        mi_check(mi_poly_op(mi, MI_OP_SUB, b, tmp.data(), (int)KL, 0, 1, W, stream()), "ckks_context");
        if (params.fold_hybrid_keyswitch || params.montgomery)
            mi_check(mi_poly_scalar(mi, MI_OP_MUL, b, scale.data(), (int)KL, 0, 2, W, stream()), "ckks_context");
    }
    mi_check(mi_stream_synchronize(mi, stream()), "ckks_context");
    return key;
}

void CKKSContext::addRotationKey(int64_t rotation) {
    const size_t N = params.getN();
    const int64_t r = normalize_rotation_value(N, rotation);
    if (r == 0 || K().rotation_keys.count(r)) return;
    // keys switch s to pi^-1(s); the rotation applies pi after the switch
    DeviceBuffer tgt = secret_qp_from_coeffs(automorphism_coeffs(K().secret_coeffs, rotation_automorphism_index(N, -r)));
    K().rotation_keys.emplace(r, generate_switch_key(K().secret_qp.data(), tgt.data(), 1000 + (uint64_t)r));
}

DeviceBuffer CKKSContext::upload_key(const MyVector<uint64_t>& key) const {
    if (key.size() != switchKeyWords())
        throw std::invalid_argument("ckks_context: switching key size mismatch (expected " +
                                    std::to_string(switchKeyWords()) + " words)");
    DeviceBuffer d(mi, key.size());
    upload(*this, d.data(), key.data(), key.size());
    return d;
}

void CKKSContext::setRelinKey(const MyVector<uint64_t>& key) { K().relin_key = upload_key(key); }
void CKKSContext::setConjugationKey(const MyVector<uint64_t>& key) { K().conj_key = upload_key(key); }
void CKKSContext::setRotationKey(int64_t rotation, const MyVector<uint64_t>& key) {
    const int64_t r = normalize_rotation_value(params.getN(), rotation);
    if (r == 0) throw std::invalid_argument("ckks_context: rotation by 0 needs no key");
    K().rotation_keys[r] = upload_key(key);
}
void CKKSContext::setSecretKeyEval(const MyVector<uint64_t>& q_limbs) {
    const size_t N = params.getN(), nq = params.getModuli().size();
    if (q_limbs.size() != nq * N) throw std::invalid_argument("ckks_context: secret key size mismatch");
    upload(*this, K().secret_qp.data(), q_limbs.data(), q_limbs.size());
}

bool CKKSContext::hasRotationKey(int64_t rotation) const {
    return K().rotation_keys.count(normalize_rotation_value(params.getN(), rotation)) != 0;
}

const uint64_t* CKKSContext::getRotationKeyDevice(int64_t rotation) const {
    auto it = K().rotation_keys.find(normalize_rotation_value(params.getN(), rotation));
    if (it == K().rotation_keys.end())
        throw std::runtime_error("ckks_context: rotation key is unavailable for rotation " + std::to_string(rotation));
    return it->second.data();
}

std::vector<int64_t> CKKSContext::getRotationKeyValues() const {
    std::vector<int64_t> v;
    for (auto& kv : K().rotation_keys) v.push_back(kv.first);
    return v;
}

const uint64_t* CKKSContext::getConjugationKeyDevice() const {
    if (!K().conj_key.data()) throw std::runtime_error("ckks_context: conjugation key is unavailable");
    return K().conj_key.data();
}

MyVector<MyVector<uint64_t>> CKKSContext::getSecretKey() const {
    const size_t N = params.getN(), nq = params.getModuli().size();
    MyVector<uint64_t> all(nq * N);
    download(*this, all.data(), K().secret_qp.data(), all.size());
    MyVector<MyVector<uint64_t>> out(nq);
    for (size_t l = 0; l < nq; ++l) out[l].assign(all.begin() + l * N, all.begin() + (l + 1) * N);
    return out;
}

MyVector<uint64_t> CKKSContext::downloadKey(const uint64_t* key) const {
    MyVector<uint64_t> h(switchKeyWords());
    download(*this, h.data(), key, h.size());
    return h;
}

uint64_t CKKSContext::nextRandomnessSeed() const {
    uint64_t st = params.getSeed() ^ 0x3C6EF372FE94F82Bull;
    st += 0x9E3779B97F4A7C15ull * (K().randomness_counter.fetch_add(1) + 1);
    return ckks_host::splitmix64(st);
}
