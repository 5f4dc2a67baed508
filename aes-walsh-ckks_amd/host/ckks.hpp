// C++ host API of the MI355X CKKS library.
//
// The reference's public types and free functions (aes-walsh-ckks
// src/ckks/{parameters,context,ciphertext,encoding,operations}) with the same
// names, argument meaning and exceptions (std::invalid_argument for bad shape,
// level or metadata; std::runtime_error for missing keys or device failures;
// messages prefixed "<function>: ").  Ciphertext, plaintext and key storage is
// resident in device memory and every ring operation runs in the HIP kernels
// behind include/mickks.h; this layer only holds metadata, host precompute
// (moduli, FFT tables, key sampling) and the floating-point encode/decode.
//
// Differences from the reference API, all forced by device residency:
//   * CKKSCiphertext / CKKSEncoding carry the CKKSContext they live on;
//     getPoly()/getPlaintext() return host copies, setPoly() uploads, and
//     polyData() exposes the device pointer.
//   * randomness comes from one 64-bit seed per CKKSParams (drawn from
//     std::random_device unless CKKSParamsConfig::seed sets it): it drives key
//     generation and each context's encryption stream (the reference draws
//     from getrandom directly).  Contexts built from one CKKSParams share keys.
#pragma once

#include <atomic>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <array>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <span>
#include <vector>

#include "../../include/mickks.h"

using Complex = std::complex<double>;
template <class T>
using MyVector = std::vector<T>;

enum class CKKSSecretOwner { Dense, BootstrapSparse };
enum class CKKSMessageEncodingState { Slots, Coefficients };
enum class CKKSDecryptDomain { Slots, Coefficients };
enum class CKKSHybridKeySwitchCombine { Add, Assign };

// ---------------------------------------------------------------- parameters
using CKKSBootstrapDftScaleSchedule = std::vector<std::vector<size_t>>;

// CKKSBootstrapParams (ckks_params.hpp:47-70): index 0 is SlotsToCoeffs,
// index 1 CoeffsToSlots.
struct CKKSBootstrapParams {
    size_t sparse_secret_hamming_weight = 0;
    size_t active_slot_count = 0;  // 0: N/2
    std::array<size_t, 2> linear_transform_depths{0, 0};
    std::array<CKKSBootstrapDftScaleSchedule, 2> linear_transform_log_plaintext_scales;
    size_t evalmod_depth = 0;
    size_t evalmod_bit_size = 0;
    size_t evalmod_k = 0;  // 0: from the sparse-secret overflow bound
    long double evalmod_k_failure_probability_log2 = -128.0L;
    size_t evalmod_polynomial_degree = 120;
    size_t evalmod_double_angle = 0;
    size_t evalmod_log_message_ratio = 0;
    // controls of the reference's EvalMod fit: a context with bootstrap keys
    // installs the fit for its parameters (default_evalmod_polynomial);
    // CKKSContext::setEvalModPolynomial replaces it
    double evalmod_coefficient_prune_threshold = 1e-10;
    double evalmod_discrete_interval_radius = 0.0;
    size_t evalmod_discrete_validation_points_per_interval = 0;
    std::array<size_t, 2> linear_transform_bsgs_baby_step_counts{0, 0};
    std::array<int, 2> linear_transform_log_bsgs_ratios{1, 1};
    std::array<std::vector<size_t>, 2> linear_transform_merge_depths;
    // QP evaluator for S2C / C2S stages (the presets turn both on)
    std::array<bool, 2> linear_transform_use_qp{false, false};
};

// CKKSParamsConfig (ckks_params.hpp:85-117).
// CKKSAesWalshParams (ckks_params.hpp:72-81): the AES Walsh path's active
// blocks per bit ciphertext (also the packed segment stride) and the number
// of segments to provision rotation keys for (0: slots / block_count).
struct CKKSAesWalshParams {
    bool enabled = false;
    size_t block_count = 0;
    size_t packed_segment_count = 0;
};

struct CKKSParamsConfig {
    CKKSAesWalshParams aes_walsh_params;
    double scale = 0.0;
    double sigma = 3.19;
    size_t N = size_t{1} << 15;
    size_t slots = 0;
    size_t depth = 0;
    size_t mult_depth = 0;
    size_t q0_bit = 33;
    size_t q_bit = 32;
    std::vector<uint64_t> q_mult_bits;
    size_t dnum = 3;
    size_t p_bit = 61;
    size_t pmoduli_count = 1;
    std::vector<uint64_t> p_bits;
    size_t secret_key_hamming_weight = 256;
    std::vector<uint64_t> q_bits;
    MyVector<int64_t> rotate_values;
    bool bts_keys = false;
    std::optional<CKKSBootstrapParams> bootstrap_params;
    bool secure = true;
    bool fold_hybrid_keyswitch = true;
    bool montgomery = true;
    // Drives key generation and each context's encryption randomness.  Unset
    // (the default), CKKSParams draws it from std::random_device (getrandom),
    // as the reference draws its keys; set it only to replay a key set (tests,
    // contexts that must share keys).
    std::optional<uint64_t> seed;
};

// AESWalsh15BootstrapParams / AESWalsh15Config (ckks_aes_params.hpp:35-82):
// N = 2^15, q0 36 bits, two 42-bit mult limbs, scale 2^35, dnum 3, P = 4 x
// 61 bits, Hamming weight 256, 512 AES blocks per lane with 32 packed
// segments; bootstrap: sparse weight 32, S2C one stage of 18+20+23, C2S two
// stages of 25+25, EvalMod 7 x 38 bits with K at failure probability 2^-16,
// degree 100, QP evaluators for both transforms.
CKKSBootstrapParams aes_walsh15_bootstrap_params();
CKKSParamsConfig aes_walsh15_config();
// The XBOOT presets of aes_bench's xboot rows (ckks_aes_params.hpp:12-211).
CKKSParamsConfig xboot_aes13_reference_config();
CKKSParamsConfig aes_xboot14_secure_config();
CKKSParamsConfig aes_xboot15_config();

// find_p_and_q_primes (random.cpp:66-131): P bit sizes first, then Q; one
// descending cursor per bit size from 2^bits/(2N); q = 2kN+1; Miller-Rabin.
std::pair<std::vector<uint64_t>, std::vector<uint64_t>> find_p_and_q_primes(
    const std::vector<uint64_t>& p_bits, const std::vector<uint64_t>& q_bits, size_t N);

class CKKSParams {
public:
    explicit CKKSParams(CKKSParamsConfig config);
    // CKKSParams::AESWalsh15 (ckks_aes_params.hpp:213-215)
    static CKKSParams AESWalsh15();
    // ckks_params.hpp:232-237 / ckks_aes_params.hpp:217-235
    static CKKSParams XBootParamAes13Reference();
    static CKKSParams AESXBoot14Secure();
    static CKKSParams AESXBoot14SecureSparse1024();
    static CKKSParams AESXBoot15();
    static CKKSParams AESXBoot15Sparse1024();
    size_t getN() const { return N; }
    size_t getSlots() const { return slots; }
    size_t getDepth() const { return depth; }
    size_t getMaxLevel() const { return moduli.size() - 1; }
    size_t getDnum() const { return dnum; }
    size_t getAlpha() const { return (moduli.size() + dnum - 1) / dnum; }
    double getScale() const { return scale; }
    double getSigma() const { return sigma; }
    size_t getSecretKeyHammingWeight() const { return hamming_weight; }
    const MyVector<uint64_t>& getModuli() const { return moduli; }
    const MyVector<uint64_t>& getPModuli() const { return pmoduli; }
    const std::vector<uint64_t>& getQBits() const { return q_bits; }
    const std::vector<uint64_t>& getPBits() const { return p_bits; }
    const MyVector<int64_t>& getRotateValues() const { return rotate_values; }
    uint64_t getSeed() const { return seed; }
    size_t getQ0Bit() const { return q0_bit; }
    size_t getDefaultQBit() const { return q_bit; }
    // bootstrapping (ckks_params.cpp:300-455)
    bool hasBootstrapParams() const { return bootstrap.has_value(); }
    const CKKSBootstrapParams& getBootstrapParams() const;
    size_t getBootstrapS2CDepth() const;
    size_t getBootstrapC2SDepth() const;
    size_t getBootstrapEvalModDepth() const { return getBootstrapParams().evalmod_depth; }
    size_t getBootstrapEvalModPolynomialDegree() const { return getBootstrapParams().evalmod_polynomial_degree; }
    size_t getBootstrapEvalModDoubleAngle() const { return getBootstrapParams().evalmod_double_angle; }
    bool getBootstrapS2CUseQpLinearTransform() const { return getBootstrapParams().linear_transform_use_qp[0]; }
    bool getBootstrapC2SUseQpLinearTransform() const { return getBootstrapParams().linear_transform_use_qp[1]; }
    size_t getBootstrapActiveSlotCount() const;
    bool usesSparseBootstrapSlots() const { return getBootstrapActiveSlotCount() != slots; }
    size_t getBootstrapEvalModK() const;
    size_t getEvalModLogScaleBits() const;
    // ckks_params.cpp:184-205
    bool usesAesWalsh() const { return aes_walsh.enabled; }
    size_t getAesWalshBlockCount() const { return aes_walsh.block_count; }
    size_t getAesWalshPackedSegmentCount() const;

    bool bts_keys = false;
    bool secure = true;
    bool fold_hybrid_keyswitch = true;
    bool montgomery = true;

private:
    size_t N = 0, slots = 0, depth = 0, mult_depth = 0, dnum = 0, hamming_weight = 0;
    size_t q0_bit = 0, q_bit = 0;
    std::optional<CKKSBootstrapParams> bootstrap;
    double scale = 0.0, sigma = 0.0;
    std::vector<uint64_t> q_bits, p_bits;
    MyVector<uint64_t> moduli, pmoduli;
    MyVector<int64_t> rotate_values;
    CKKSAesWalshParams aes_walsh;
    uint64_t seed = 0;
};

class CKKSContext;
namespace ckks_host {
class Rng;
}

// ------------------------------------------------------------ device storage
// One allocation from the context's device pool (mi_alloc / mi_free).
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(mi_context* ctx, size_t words);
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept { swap(o); }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            reset();
            swap(o);
        }
        return *this;
    }
    ~DeviceBuffer() { reset(); }
    uint64_t* data() const { return ptr_; }
    size_t words() const { return words_; }
    void reset();
    void swap(DeviceBuffer& o) noexcept {
        std::swap(ctx_, o.ctx_);
        std::swap(ptr_, o.ptr_);
        std::swap(words_, o.words_);
    }

private:
    mi_context* ctx_ = nullptr;
    uint64_t* ptr_ = nullptr;
    size_t words_ = 0;
};

// A switching key as held in HBM: [parts][2 (b, a)][nq+np][N], folded (the
// form the context generates).  CKKSContext::getRotationKey, getRelinHybrid,
// getConjugationKey and the dense/sparse switch-key getters return it where
// the reference returns a span of per-digit key ciphertexts.
struct CKKSSwitchingKey {
    const uint64_t* device = nullptr;
};

// ckks_keyswitch.hpp:70-86: the hoisted QP digits of a ciphertext's a, kept
// in HBM ([parts][L+K][N], evaluation domain).  The workspace is the
// context's own per-stream scratch, so CKKSHoistWorkspace carries nothing.
struct CKKSEvalHoistedCiphertextQP {
    size_t N = 0;
    size_t level = 0;
    size_t q_limb_count = 0;
    size_t p_limb_count = 0;
    long double scale = 0.0L;
    MyVector<uint64_t> moduli;
    DeviceBuffer digits;
};
struct CKKSHoistWorkspace {};

// ---------------------------------------------------------- ciphertext types
// CKKSCiphertext (ckks_ciphertext.hpp:18-80): polys [b, a, (c2)], each
// (level+1)*N words in the evaluation domain.  Move-only; clone() copies.
//
// A CKKSCiphertext may also hold a batch of `getBatchSize()` ciphertexts
// that share every piece of metadata (level, scale, degree, secret owner,
// encoding state): the reference runs one evaluation schedule per
// ciphertext on OpenMP threads (16 S-boxes per AES round, the final-refresh
// packs); a batch runs the schedule once, every ring step one launch for all
// members, and each member's limbs equal what the one-ciphertext evaluation
// gives.  stack() / unstack() move members in and out without copies.
// polyData(k) names a poly of a single ciphertext; polyData(k, m) the poly k
// of member m.  Functions that have no batched form reject batches.
class CKKSCiphertext {
public:
    // Uninitialised device polys.
    CKKSCiphertext(const CKKSContext& context, size_t num_polys, long double scale, size_t level,
                   CKKSSecretOwner secret_owner = CKKSSecretOwner::Dense,
                   CKKSMessageEncodingState message_encoding_state = CKKSMessageEncodingState::Slots,
                   size_t batch = 1);
    // Uploads host polys (ckks_ciphertext.hpp:20-33 with the context added).
    CKKSCiphertext(const CKKSContext& context, MyVector<MyVector<uint64_t>> polys, long double scale,
                   size_t level, CKKSSecretOwner secret_owner = CKKSSecretOwner::Dense,
                   CKKSMessageEncodingState message_encoding_state = CKKSMessageEncodingState::Slots);
    CKKSCiphertext(const CKKSCiphertext&) = delete;
    CKKSCiphertext& operator=(const CKKSCiphertext&) = delete;
    CKKSCiphertext(CKKSCiphertext&&) noexcept = default;
    CKKSCiphertext& operator=(CKKSCiphertext&&) noexcept = default;

    CKKSCiphertext clone() const;
    // clone() followed by ckks_drop_to_level_inplace(out, level), copying
    // only limbs 0..level.
    CKKSCiphertext cloneAtLevel(size_t level) const;

    long double getScale() const { return scale; }
    size_t getLevel() const { return level; }
    size_t getN() const { return N; }
    size_t getNumPolys() const { return polys.size() / batch; }
    CKKSSecretOwner getSecretOwner() const { return secret_owner; }
    CKKSMessageEncodingState getMessageEncodingState() const { return message_encoding_state; }
    size_t getBatchSize() const { return batch; }
    // A batch of the members (same context, N, level, degree, owner, encoding
    // state and scale; std::invalid_argument otherwise), buffers moved in.
    static CKKSCiphertext stack(MyVector<CKKSCiphertext> members);
    // The members, buffers moved out; *this is left empty.
    MyVector<CKKSCiphertext> unstack();
    void setScale(long double s) { scale = s; }
    void setLevel(size_t l);
    void setSecretOwner(CKKSSecretOwner o) { secret_owner = o; }
    void setMessageEncodingState(CKKSMessageEncodingState s) { message_encoding_state = s; }

    // Host copies of the active limbs ((level+1)*N words).
    MyVector<uint64_t> getPoly(size_t idx) const;
    MyVector<uint64_t> getA() const { return getPoly(1); }
    MyVector<uint64_t> getB() const { return getPoly(0); }
    void setPoly(size_t idx, const MyVector<uint64_t>& host);
    // Device pointers (limb-major, limb l at l*N).
    uint64_t* polyData(size_t idx);
    const uint64_t* polyData(size_t idx) const;
    uint64_t* polyData(size_t idx, size_t member);
    const uint64_t* polyData(size_t idx, size_t member) const;
    // Every device poly, poly-major (poly k of member m at k * batch + m).
    std::vector<uint64_t*> allPolyData();
    std::vector<const uint64_t*> allPolyData() const;
    // Host copy of poly idx of one member.
    MyVector<uint64_t> getPoly(size_t idx, size_t member) const;
    // Appends an uninitialised poly (the c2 slot of a tensor product).
    void addPoly();
    void resizePolys(size_t count);
    const CKKSContext& getContext() const { return *ctx; }

private:
    const CKKSContext* ctx;
    size_t N;
    long double scale;
    size_t level;
    CKKSSecretOwner secret_owner;
    CKKSMessageEncodingState message_encoding_state;
    size_t batch = 1;
    std::vector<DeviceBuffer> polys;  // poly k of member m at k * batch + m
};

// CKKSEncoding (ckks_ciphertext.hpp:83-108): one (level+1)-limb plaintext.
class CKKSEncoding {
public:
    CKKSEncoding(const CKKSContext& context, MyVector<uint64_t> plaintext, long double scale,
                 size_t level, bool montgomery_form = false);
    CKKSEncoding(const CKKSContext& context, DeviceBuffer plaintext, long double scale, size_t level,
                 bool montgomery_form = false);
    CKKSEncoding(const CKKSEncoding&) = delete;
    CKKSEncoding& operator=(const CKKSEncoding&) = delete;
    CKKSEncoding(CKKSEncoding&&) noexcept = default;
    CKKSEncoding& operator=(CKKSEncoding&&) noexcept = default;

    MyVector<uint64_t> getPlaintext() const;
    const uint64_t* data() const { return buf.data(); }
    bool isMontgomeryForm() const { return montgomery_form; }
    const CKKSContext& getContext() const { return *ctx; }
    long double getScale() const { return scale; }
    size_t getLevel() const { return level; }
    size_t getN() const { return N; }

private:
    const CKKSContext* ctx;
    size_t N;
    long double scale;
    size_t level;
    bool montgomery_form;
    DeviceBuffer buf;
};

// ---------------------------------------------------- bootstrapping DFT plans
// ckks_bootstrap_dft.hpp:10-40, ckks_linear_transform_plan.hpp:9-47.  The
// generated DFT stages keep their diagonals once, in the Q u P basis at the
// stage level on the device; the Q evaluator reads the first level+1 limbs
// of those (the Q-basis encoding is the same per-modulus encode).  Stages a
// caller builds for the Q evaluator hold CKKSEncoding diagonals.
enum class CKKSBootstrapDftType { SlotsToCoeffs, CoeffsToSlots };

struct CKKSBootstrapDftMatrixLiteral {
    CKKSBootstrapDftType type = CKKSBootstrapDftType::SlotsToCoeffs;
    size_t level = 0;
    size_t baby_step_count = 0;  // 0: chosen from the diagonal indices
    size_t stage_count = 1;
    CKKSBootstrapDftScaleSchedule log_plaintext_scale_groups;
    std::vector<size_t> merge_depths;
    int log_bsgs_ratio = 1;
    size_t active_slots = 0;     // 0: N/2
    double transform_scale = 1.0;
    double plaintext_scale = 1.0;
    size_t input_selector_start = 0;
    size_t input_selector_size = 0;
    bool rescale_after_each_stage = true;
};

// A stage over caller-owned diagonals (ckks_linear_transform_plan.hpp:9-17):
// diagonal_indices empty means term t is diagonal t; slot_count 0 means N/2.
struct CKKSLinearTransformStage {
    std::span<const CKKSEncoding> diagonals;
    std::span<const size_t> diagonal_indices;
    size_t slot_count = 0;
    size_t baby_step_count = 0;
    bool rescale_after = true;
    CKKSMessageEncodingState output_encoding_state = CKKSMessageEncodingState::Slots;
};

struct CKKSLinearTransformOwnedStage {
    CKKSLinearTransformOwnedStage() = default;
    CKKSLinearTransformOwnedStage(const CKKSLinearTransformOwnedStage&) = delete;
    CKKSLinearTransformOwnedStage& operator=(const CKKSLinearTransformOwnedStage&) = delete;
    CKKSLinearTransformOwnedStage(CKKSLinearTransformOwnedStage&&) noexcept = default;
    CKKSLinearTransformOwnedStage& operator=(CKKSLinearTransformOwnedStage&&) noexcept = default;

    MyVector<CKKSEncoding> diagonals;               // Q-evaluator diagonals (may be empty, see above)
    MyVector<DeviceBuffer> diagonal_qp_plaintexts;  // (level+1+K)*N each
    MyVector<size_t> diagonal_indices;
    size_t slot_count = 0;
    size_t baby_step_count = 0;
    size_t qp_limb_count = 0;
    size_t qp_p_limb_count = 0;
    size_t level = 0;
    double plaintext_scale = 1.0;
    bool diagonal_qp_plaintexts_montgomery = false;
    bool rescale_after = true;
    CKKSMessageEncodingState output_encoding_state = CKKSMessageEncodingState::Slots;
    MyVector<uint64_t> downloadDiagonalQp(const CKKSContext& context, size_t term) const;
};

struct CKKSLinearTransformPlan {
    CKKSLinearTransformPlan() = default;
    CKKSLinearTransformPlan(const CKKSLinearTransformPlan&) = delete;
    CKKSLinearTransformPlan& operator=(const CKKSLinearTransformPlan&) = delete;
    CKKSLinearTransformPlan(CKKSLinearTransformPlan&&) noexcept = default;
    CKKSLinearTransformPlan& operator=(CKKSLinearTransformPlan&&) noexcept = default;

    MyVector<CKKSLinearTransformOwnedStage> stages;
    bool use_qp_evaluator = false;
};

// ------------------------------------------------------------------- context
// CKKSContext (ckks_context.hpp:59-90; constructor ckks_context.cpp:603-705):
// FFT tables and the slot permutation on the host; NTT, ModUp/ModDown and
// rescale tables on the device (mi_context_create); a seeded secret key and
// the relinearisation, conjugation and rotation keys generated on the device.
class CKKSContext {
public:
    explicit CKKSContext(CKKSParams params);
    ~CKKSContext();
    // A lane of `owner`: its own device tables, memory pool, workspaces and
    // stream, and the owner's secret, switching keys, bootstrapping plans and
    // EvalMod polynomial read in place (nothing regenerated or copied), so a
    // second evaluation stream costs its working memory only.  The owner
    // must outlive its lanes; key-changing calls on a lane throw.  Lanes on
    // separate host threads overlap on the device.
    static std::unique_ptr<CKKSContext> makeLane(const CKKSContext& owner);
    bool isLane() const { return owner_ != nullptr; }
    const CKKSContext& keyOwner() const { return owner_ ? *owner_ : *this; }
    CKKSContext(const CKKSContext&) = delete;
    CKKSContext& operator=(const CKKSContext&) = delete;

    const CKKSParams& getParams() const { return params; }
    mi_context* handle() const { return mi; }
    // The context's own stream: every evaluation of this context is ordered
    // on it; it synchronises with the null stream, not with other contexts'
    // streams, so contexts on separate host threads overlap on the device.
    void* stream() const { return stream_; }
    const MyVector<Complex>& getTwiddleFft() const { return twiddle_fft; }
    const MyVector<Complex>& getTwiddleIfft() const { return twiddle_ifft; }
    const MyVector<size_t>& getPermTable() const { return perm_table; }

    const MyVector<int8_t>& getSecretKeyCoeffs() const { return K().secret_coeffs; }
    // Evaluation-domain secret over q_0..q_Lmax, p_0..p_{K-1} (device).
    const uint64_t* secretKeyDevice() const { return K().secret_qp.data(); }
    // Host copy of the evaluation-domain secret, one vector per Q limb.
    MyVector<MyVector<uint64_t>> getSecretKey() const;

    // Switching keys: [digits][2 (b, a)][nq+np][N] on the device.
    size_t keyDigitCount() const { return K().key_digits; }
    size_t switchKeyWords() const;
    const uint64_t* getRelinKeyDevice() const { return K().relin_key.data(); }
    bool hasConjugationKey() const { return K().conj_key.data() != nullptr; }
    const uint64_t* getConjugationKeyDevice() const;
    size_t getRotationKeyCount() const { return K().rotation_keys.size(); }
    bool hasRotationKey(int64_t rotation) const;
    const uint64_t* getRotationKeyDevice(int64_t rotation) const;
    // the reference's key getters (ckks_context.hpp:78-110) over the device keys
    CKKSSwitchingKey getRotationKey(int64_t rotation) const { return {getRotationKeyDevice(rotation)}; }
    CKKSSwitchingKey getRelinHybrid(size_t /*level*/) const { return {getRelinKeyDevice()}; }
    CKKSSwitchingKey getConjugationKey() const { return {getConjugationKeyDevice()}; }
    CKKSSwitchingKey getDenseToSparseSwitchKey() const { return {getDenseToSparseSwitchKeyDevice()}; }
    CKKSSwitchingKey getSparseToDenseSwitchKey() const { return {getSparseToDenseSwitchKeyDevice()}; }
    // build_ntt_automorphism_map (rotation_key_generation.cpp:80-90) of a
    // rotation's and of the conjugation's Galois element
    MyVector<size_t> getRotationNttMap(int64_t rotation) const;
    MyVector<size_t> getConjugationNttMap() const;
    std::vector<int64_t> getRotationKeyValues() const;
    // Generates the key for one more rotation (normalised mod N/2).
    void addRotationKey(int64_t rotation);
    // Installs keys made elsewhere (a saved key set, or another
    // implementation's keys for a cross-check): switching keys in the layout
    // above; the secret as evaluation-domain Q limbs (used by decryption and
    // encryption; key generation keeps the context's own secret).
    void setRelinKey(const MyVector<uint64_t>& key);
    void setConjugationKey(const MyVector<uint64_t>& key);
    void setRotationKey(int64_t rotation, const MyVector<uint64_t>& key);
    void setSecretKeyEval(const MyVector<uint64_t>& q_limbs);
    // Host copy of a switching key (tests and serialisation).
    MyVector<uint64_t> downloadKey(const uint64_t* key) const;

    // Encryption randomness: a fresh stream per call, derived from the seed.
    uint64_t nextRandomnessSeed() const;

    // ---- bootstrapping material (ckks_context.cpp:603-705 with bts_keys)
    bool hasSparseSecretEncapsulationKeys() const { return K().d2s_key.data() != nullptr; }
    const MyVector<int8_t>& getLowSecretKeyCoeffs() const { return K().low_secret_coeffs; }
    const uint64_t* lowSecretKeyDevice() const { return K().low_secret_qp.data(); }
    // [b(q0), b(p0), a(q0), a(p0)], N words each
    const uint64_t* getDenseToSparseSwitchKeyDevice() const;
    const uint64_t* getSparseToDenseSwitchKeyDevice() const;
    void setDenseToSparseSwitchKey(const MyVector<uint64_t>& key);
    void setSparseToDenseSwitchKey(const MyVector<uint64_t>& key);
    MyVector<uint64_t> downloadDenseToSparseSwitchKey() const;
    bool hasEvalModBootstrapDftPlans() const { return K().s2c_plan != nullptr || K().c2s_plan != nullptr; }
    const CKKSLinearTransformPlan& getEvalModS2CPlan() const;
    const CKKSLinearTransformPlan& getEvalModScaledC2SPlan() const;
    double getEvalModScaledC2STransformScale() const { return K().c2s_transform_scale; }
    // S2C plan for an input whose scale is `ratio` times the default scale
    // (cached; ckks_bootstrap_evalmod_inplace.cpp:114-154).
    const CKKSLinearTransformPlan& getScaledS2CPlan(uint64_t ratio) const;
    // The EvalMod Chebyshev coefficients.  The reference derives them with
    // MPFR at context construction; here they are supplied by the caller.
    void setEvalModPolynomial(const std::vector<double>& chebyshev_coeffs);
    bool hasEvalModPolynomial() const { return !K().evalmod_coeffs.empty(); }
    const std::vector<double>& getEvalModPolynomial() const;

private:
    void release_all();
    void build_bootstrap_material(ckks_host::Rng& rng);
    DeviceBuffer generate_q0p0_key(const uint64_t* src_qp, const uint64_t* tgt_qp, uint64_t stream_id) const;
    DeviceBuffer upload_key(const MyVector<uint64_t>& key) const;
    DeviceBuffer generate_switch_key(const uint64_t* src_secret_qp, const uint64_t* tgt_secret_qp,
                                     uint64_t stream_id) const;
    DeviceBuffer secret_qp_from_coeffs(const MyVector<int8_t>& coeffs) const;

    // Secret, switching keys, bootstrapping plans and the EvalMod polynomial:
    // owned by a key-owning context, read through K() by it and its lanes.
    struct KeyMaterial {
        MyVector<int8_t> secret_coeffs;
        DeviceBuffer secret_qp;
        size_t key_digits = 0;
        DeviceBuffer relin_key, conj_key;
        std::map<int64_t, DeviceBuffer> rotation_keys;
        mutable std::atomic<uint64_t> randomness_counter{0};
        MyVector<int8_t> low_secret_coeffs;
        DeviceBuffer low_secret_qp, d2s_key, s2d_key;
        std::unique_ptr<CKKSLinearTransformPlan> s2c_plan, c2s_plan;
        double c2s_transform_scale = 1.0;
        std::vector<double> evalmod_coeffs;
    };
    const KeyMaterial& K() const { return owner_ ? owner_->keys_ : keys_; }
    KeyMaterial& K();  // the owner's own material; a lane throws
    CKKSContext(const CKKSContext& owner, int lane_tag);

    CKKSParams params;
    mi_context* mi = nullptr;
    void* stream_ = nullptr;
    const CKKSContext* owner_ = nullptr;  // set for a lane: the context whose keys it reads
    mutable std::atomic<int> lanes_{0};   // live lanes of this key-owning context
    MyVector<Complex> twiddle_fft, twiddle_ifft;
    MyVector<size_t> perm_table;
    KeyMaterial keys_;
    // scaled S2C plans, per context (their diagonals live in its own pool)
    mutable std::mutex scaled_mu;
    mutable std::map<uint64_t, std::unique_ptr<CKKSLinearTransformPlan>> scaled_s2c;

public:
    // An encoding that depends only on `key` (what is encoded, its scale and
    // level), made by `make` on first use and kept for the context's
    // lifetime (at most 256 entries; past that each call makes its own).
    // Used for the constant-slot and slot-mask plaintexts the evaluator
    // multiplies by, which repeat in every S-box and bootstrap.
    template <class Make>
    const CKKSEncoding* cachedEncoding(const std::array<uint64_t, 4>& key, Make make,
                                       std::optional<CKKSEncoding>& fallback) const {
        std::lock_guard<std::mutex> g(enc_mu);
        auto it = enc_cache.find(key);
        if (it != enc_cache.end()) return it->second.get();
        if (enc_cache.size() >= 256) {
            fallback.emplace(make());
            return &*fallback;
        }
        auto e = std::make_unique<CKKSEncoding>(make());
        const CKKSEncoding* ptr = e.get();
        enc_cache.emplace(key, std::move(e));
        return ptr;
    }

private:
    mutable std::mutex enc_mu;
    mutable std::map<std::array<uint64_t, 4>, std::unique_ptr<CKKSEncoding>> enc_cache;
};

// ------------------------------------------------------------ encode / crypt
// ckks_encoding.hpp:15-77.
CKKSEncoding ckks_encode(const CKKSContext& context, const MyVector<Complex>& values, size_t level,
                         bool eval = false);
CKKSEncoding ckks_encode(const CKKSContext& context, const MyVector<Complex>& values, double scale,
                         size_t level, bool eval = false);
CKKSEncoding ckks_encode_coefficients(const CKKSContext& context, const MyVector<Complex>& values,
                                      double scale, size_t level, bool eval = false);
// Integer plaintext of the slot values (before any NTT): limb-major residues.
MyVector<uint64_t> encode_slots_custom_basis(size_t N, const MyVector<Complex>& values, double scale,
                                             const MyVector<uint64_t>& moduli,
                                             const MyVector<Complex>& twiddle_fft,
                                             const MyVector<size_t>& perm_table);
CKKSCiphertext ckks_encrypt(const CKKSContext& context, const MyVector<Complex>& values, size_t level,
                            bool eval = false,
                            CKKSMessageEncodingState state = CKKSMessageEncodingState::Slots);
CKKSCiphertext ckks_encrypt(const CKKSContext& context, const CKKSEncoding& encd, bool eval = false,
                            CKKSMessageEncodingState state = CKKSMessageEncodingState::Slots);
MyVector<Complex> ckks_decrypt(const CKKSContext& context, const CKKSCiphertext& ct,
                               CKKSDecryptDomain domain = CKKSDecryptDomain::Slots,
                               bool use_low_secret = false);
// ckks_decrypt of many ciphertexts: the device parts in order on the
// context's stream, the host decodes (CRT and FFT) on up to 16 threads; each
// result equals ckks_decrypt's.  No counterpart in the reference (its bench
// decrypts one ciphertext at a time, aes_bench.cpp:156-196).
std::vector<MyVector<Complex>> ckks_decrypt_batch(const CKKSContext& context,
                                                  const std::vector<const CKKSCiphertext*>& cts,
                                                  CKKSDecryptDomain domain = CKKSDecryptDomain::Slots);
// The decrypted plaintext's coefficients (b + a s, inverse NTT), L x N
// residues, downloaded: the input of ckks_decode_plaintext.
MyVector<uint64_t> ckks_decrypt_coefficients(const CKKSContext& context, const CKKSCiphertext& ct);
// Decode of a coefficient-domain integer plaintext (first limbs whose product
// fits 128 bits, centred) at `scale` (ckks_encoding.cpp:75-170, 240-329).
MyVector<Complex> ckks_decode_plaintext(const CKKSContext& context, const MyVector<uint64_t>& coeffs,
                                        size_t limb_count, long double scale,
                                        CKKSDecryptDomain domain = CKKSDecryptDomain::Slots);
MyVector<Complex> decode_plaintext_custom_basis(size_t N, const MyVector<uint64_t>& coeffs,
                                                const MyVector<uint64_t>& moduli, long double scale,
                                                const MyVector<Complex>& twiddle_ifft,
                                                const MyVector<size_t>& perm_table,
                                                CKKSDecryptDomain domain = CKKSDecryptDomain::Slots);
// Host tables (fft.cpp:37-80, ckks_context.cpp:56-68).
MyVector<Complex> twiddle_ifft_generation(size_t N);
MyVector<size_t> perm_table_generation(size_t N);

// ----------------------------------------------------------------- evaluator
// ckks_mul.hpp, ckks_rescale.hpp, ckks_addition.hpp, ckks_relin.hpp,
// ckks_rotation.hpp, ckks_conjugate.hpp.
void ckks_mul_inplace(const CKKSContext& context, CKKSCiphertext& ct1, const CKKSEncoding& pt);
void ckks_mul_inplace(const CKKSContext& context, CKKSCiphertext& ct1, const CKKSCiphertext& ct2);
void ckks_mul_inplace(const CKKSContext& context, CKKSCiphertext& ct, int64_t scalar);
void ckks_mul_inplace(const CKKSContext& context, CKKSCiphertext& ct, Complex value,
                      double plaintext_scale);
CKKSCiphertext ckks_mul(const CKKSContext& context, const CKKSCiphertext& ct, Complex value,
                        double plaintext_scale);
void ckks_square_inplace(const CKKSContext& context, CKKSCiphertext& ct);
void ckks_rescale(const CKKSContext& context, CKKSCiphertext& ct);
void ckks_drop_to_level_inplace(CKKSCiphertext& ct, size_t target_level);
void ckks_add_inplace(const CKKSContext& context, CKKSCiphertext& ct1, const CKKSEncoding& pt);
void ckks_add_inplace(const CKKSContext& context, CKKSCiphertext& ct1, const CKKSCiphertext& ct2);
void ckks_sub_inplace(const CKKSContext& context, CKKSCiphertext& ct1, const CKKSCiphertext& ct2);
void ckks_add_integer_constant_inplace(const CKKSContext& context, CKKSCiphertext& ct, int64_t constant);
void ckks_relin_hybrid_inplace(const CKKSContext& context, CKKSCiphertext& ct);
CKKSCiphertext ckks_rotate(const CKKSContext& context, const CKKSCiphertext& ct, int64_t rotation);
void ckks_rotate_inplace(const CKKSContext& context, CKKSCiphertext& ct, int64_t rotation);
CKKSCiphertext ckks_conjugate(const CKKSContext& context, const CKKSCiphertext& ct);
// ckks_conjugate of each ciphertext (one level, 2 polys), the key switches as
// one batch (mi_conjugate_batch); each result equals ckks_conjugate's.
MyVector<CKKSCiphertext> ckks_conjugate_batch(const CKKSContext& context, const MyVector<CKKSCiphertext>& cts);
void ckks_conjugate_inplace(const CKKSContext& context, CKKSCiphertext& ct);

// ------------------------------------------------------------ Walsh S-box
// ckks_walsh_sbox.cpp.  Walsh coefficients / 4 of the AES S-box output bits,
// [output bit][low mask][high mask] (build_walsh_table, :130-162).
using CKKSWalshTable = std::array<std::array<std::array<int16_t, 16>, 16>, 8>;
const CKKSWalshTable& ckks_aes_walsh_table();
// walsh_sum_for_output_pair (:984-1043): the Walsh sum of output bits
// (real_out_bit, imag_out_bit) over the 16 + 16 nibble sign ciphertexts (mask
// 0 is the constant 1); the linear combinations, i-combination and raw 4+4
// products run as one device pass, then one relinearisation.  All signs at
// one level; the high signs' products must share one scale.
CKKSCiphertext ckks_walsh_sum_for_output_pair(const CKKSContext& context, const MyVector<CKKSCiphertext>& low_signs,
                                              const MyVector<CKKSCiphertext>& high_signs, size_t real_out_bit,
                                              size_t imag_out_bit, const CKKSEncoding& encoded_i);
// walsh_sum_to_bit_pair (:1045-1078): z = (1/2 + i/2) - D / 128.
CKKSCiphertext ckks_walsh_sum_to_bit_pair(const CKKSContext& context, CKKSCiphertext walsh_sum);
// recombine_mask_packed_signs_to_paired_outputs (:1096-1150): the four output
// pairs (bit p real, bit p + 4 imaginary) of one S-box.
MyVector<CKKSCiphertext> ckks_walsh_recombine_signs_to_paired_outputs(const CKKSContext& context,
                                                                      const MyVector<CKKSCiphertext>& low_signs,
                                                                      const MyVector<CKKSCiphertext>& high_signs);

// build_mask_packed_parity_input (:521-546): the 30 nibble parity sums of
// the 8 input bits, each rotated by its segment * block_count (one batched
// rotation) and summed.  The caller bootstraps the result next, so the
// reference's requirement that the bits sit at the S2C entry level is left
// to it.
CKKSCiphertext ckks_walsh_build_mask_packed_parity_input(const CKKSContext& context,
                                                         const MyVector<CKKSCiphertext>& input, size_t block_count);
// unpack_mask_packed_signs (:551-596): the 16 low and 16 high sign
// ciphertexts (mask 0: the constant 1) from the packed signs, with one hoist
// and the 29 segment rotations as one batch.
std::pair<MyVector<CKKSCiphertext>, MyVector<CKKSCiphertext>> ckks_walsh_unpack_mask_packed_signs(
    const CKKSContext& context, const CKKSCiphertext& packed, size_t block_count);

// ckks_walsh_sbox_complex_blocks_mask_packed (ckks_walsh_sbox.cpp:1280-1323):
// the AES S-box on 8 bit ciphertexts (slot j: real = bit of block lane 0,
// imaginary = bit of lane 1, at the S2C entry level): parity packing, one
// complex-packed bootstrap to parity bits, signs, unpacking, the Walsh
// output pairs of both lanes and their split into the 8 output bits.
MyVector<CKKSCiphertext> ckks_walsh_sbox_complex_blocks_mask_packed(CKKSContext& context,
                                                                    MyVector<CKKSCiphertext> input,
                                                                    size_t block_count);

// The AES-CTR masking loops (ckks_aes_ctr_walsh_sbox.cpp:217-329): every
// state ciphertext times the active-slot mask (encoded at q_level) and
// rescaled back to its scale; the final refresh mask (scale q_level and a
// rescale while the level is above minimum_level_to_keep, scale 1
// otherwise).  All ciphertexts go through one batched product and rescale.
void ckks_aes_mask_inactive_slots_inplace(const CKKSContext& context, MyVector<CKKSCiphertext>& state,
                                          size_t block_count);
void ckks_aes_mask_final_refresh_slots_inplace(const CKKSContext& context, MyVector<CKKSCiphertext>& bits,
                                               size_t active_slot_count, size_t minimum_level_to_keep);

// ------------------------------------------------------------ AES-128 CTR
// aes.cpp (math/aes): FIPS-197 key expansion and plain AES-128 of one block,
// the host-side reference for the homomorphic path.
std::array<uint8_t, 176> aes128_expand_key(const std::array<uint8_t, 16>& key);
std::array<uint8_t, 16> aes128_encrypt_block_expanded(const std::array<uint8_t, 176>& expanded_round_keys,
                                                      const std::array<uint8_t, 16>& block);
std::array<uint8_t, 16> aes128_encrypt_block(const std::array<uint8_t, 16>& key, const std::array<uint8_t, 16>& block);
std::array<uint8_t, 16> aes128_ctr_crypt_block(const std::array<uint8_t, 16>& key,
                                               const std::array<uint8_t, 16>& counter_block,
                                               const std::array<uint8_t, 16>& input_block);
// The default EvalMod polynomial of a bootstrapping context (Chebyshev
// coefficients on [-1, 1] / 2^double_angle): the reference's discrete
// multi-interval fit (poly_approx.cpp:565-681) for the EvalMod parameters of
// its presets; otherwise, without double-angle steps, the degree-d
// interpolant of 0.5 - 0.5 cos(2 pi K u) on [-1, 1]; empty with double-angle
// steps (the context then needs setEvalModPolynomial).
std::vector<double> default_evalmod_polynomial(const CKKSParams& params);
// Whether default_evalmod_polynomial returns the reference's fit.
bool evalmod_polynomial_is_reference_fit(const CKKSParams& params);
const std::array<uint8_t, 256>& aes_sbox_table();

// A HIP graph of device work recorded from a context's stream: launch()
// replays the same launches on the same buffers (no host work, no
// allocation).  ckks_capture_graph records everything `body` issues on the
// context's stream (and the side streams it forks); nothing executes while
// recording.  Run the same body eagerly once first, so that workspaces,
// pool blocks and cached encodings exist; buffers the graph reads or writes
// must stay alive while it does (ciphertexts created inside `body` are
// kept by whoever holds them), and pool blocks released during the capture
// stay reserved while any graph of the context lives.
class CKKSGraph {
public:
    CKKSGraph() = default;
    CKKSGraph(const CKKSGraph&) = delete;
    CKKSGraph& operator=(const CKKSGraph&) = delete;
    CKKSGraph(CKKSGraph&& o) noexcept { std::swap(ctx_, o.ctx_), std::swap(g_, o.g_); }
    CKKSGraph& operator=(CKKSGraph&& o) noexcept {
        if (this != &o) {
            reset();
            std::swap(ctx_, o.ctx_);
            std::swap(g_, o.g_);
        }
        return *this;
    }
    ~CKKSGraph() { reset(); }
    void launch() const;  // on the context's stream
    void reset();

private:
    friend CKKSGraph ckks_capture_graph(const CKKSContext& context, const std::function<void()>& body);
    const CKKSContext* ctx_ = nullptr;
    mi_graph* g_ = nullptr;
};
CKKSGraph ckks_capture_graph(const CKKSContext& context, const std::function<void()>& body);
// dst's polys (all members) <- src's, same shape (a copy kernel; capturable)
void ckks_copy_into(const CKKSContext& context, CKKSCiphertext& dst, const CKKSCiphertext& src);

namespace ckks_detail {
// The integer factor by which the AES circuit's add_term scales a ciphertext
// of scale `from` up to `to` (ckks_aes_ctr_walsh_sbox.cpp:80-124): 1 when the
// scales agree to 2^-45, else the rounded ratio when it is within the
// reference's 1e-3 integrality tolerance; returns the reference's error text
// where it would throw, nullptr otherwise.
const char* aes_scale_up_factor(long double from, long double to, int64_t* factor);
// Every poly of `ct` through mi_poly_list_op `op` (y = the same polys of
// `other` where the op reads one, residues per limb where it scales).
void poly_list_op(const CKKSContext& c, int op, CKKSCiphertext& ct, const CKKSCiphertext* other,
                  const uint64_t* residues, const char* fn);
// The b poly of every member of `ct` through a y-less mi_poly_list_op `op`
// (MI_LIST_ADD_SCALAR: the constant terms of a batch).
void b_list_op(const CKKSContext& c, int op, CKKSCiphertext& ct, const uint64_t* residues, const char* fn);
// Polys from..degree-1 of every member set to zero (one launch per 64 polys).
void zero_polys(const CKKSContext& c, CKKSCiphertext& ct, size_t from, const char* fn);
// scale_up_ciphertext_to_match_inplace (ckks_walsh_sbox.cpp:228-270): the
// scale raised to `target` by an integer factor (an all-zero ciphertext only
// takes the new scale).
void scale_up_to(const CKKSContext& context, CKKSCiphertext& ct, long double target);
// ckks_mul_inplace(ct1, ct2) with ct2 read at ct1's level (ct2 at that level
// or above): the result equals the product with a copy of ct2 dropped to
// ct1's level, without the copy.
void mul_at_level(const CKKSContext& c, CKKSCiphertext& ct1, const CKKSCiphertext& ct2);
// the tensor product of lhs and rhs (rhs null: the square of lhs), both read
// at `level`, into a new 3-poly ciphertext at `level`; the inputs unchanged
CKKSCiphertext mul_into_at_level(const CKKSContext& c, const CKKSCiphertext& lhs, const CKKSCiphertext* rhs,
                                 size_t level);
// ckks_apply_linear_transform_plan reading `in` in place (its stages never
// write their input, so callers that keep `in` need no copy of it)
CKKSCiphertext apply_linear_transform_plan_from(const CKKSContext& c, const CKKSCiphertext& in,
                                                const CKKSLinearTransformPlan& plan);
// ckks_batch_bootstrap_evalmod_from_complex_packed with its two output lanes
// stacked into one ciphertext of twice the members (out0's members, then
// out1's), the EvalMods of both lanes run as that one batch.
CKKSCiphertext bootstrap_lanes_stacked(const CKKSContext& c, CKKSCiphertext packed);
}  // namespace ckks_detail

// ckks_aes_ctr_walsh_sbox.hpp:16-21 / .cpp:979-1035: the 1408 expanded
// round-key bits, each encrypted in the first `active_complex_block_count`
// slots (real and imaginary lanes) at the AES state entry level (S2C depth
// + 1); rounds 1..9 at scale / 4, rounds 0 and 10 at the scale.
MyVector<CKKSCiphertext> ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(
    const CKKSContext& context, const std::array<uint8_t, 176>& expanded_round_keys,
    size_t active_complex_block_count);

// ckks_aes_ctr_walsh_sbox.hpp:12-14.  parallelize_sbox_pairs: the 16
// S-boxes of a round run concurrently, one host thread per context, over the
// main context and `sbox_contexts` (contexts built from the same parameters
// and seed, hence the same keys, each with its own stream and memory pool).
struct CKKSAesCtrWalshSboxOptions {
    bool parallelize_sbox_pairs = true;
    // The 16 S-boxes of a round (and the final refresh's packs) evaluated as
    // one ciphertext batch on the main context: one launch sequence for all
    // of them.  When false they run one after another on the main stream,
    // or over sbox_contexts (one host thread each) when those are given.
    bool batch_sboxes = true;
    std::vector<CKKSContext*> sbox_contexts;
    // MixColumns + AddRoundKey of a round as one mi_linear_combine, and the
    // 128 public counter / ciphertext bit encodings as one batched encoding
    // (equal limb for limb to the per-operation sequence, which runs when false)
    bool fuse_linear_layer = true;
};

// ckks_aes_ctr_walsh_sbox.hpp:23-35 / .cpp:936-977, 1037-1051: AES-128 CTR
// decryption with the mask-packed Walsh S-box.  Block s of the first half
// of each span rides in the real lane of slot s, block n/2 + s in the
// imaginary lane; the 128 returned bit ciphertexts have been through the
// final binary refresh.
MyVector<CKKSCiphertext> ckks_aes_ctr_walsh_sbox_complex_blocks(
    CKKSContext& context, const MyVector<CKKSCiphertext>& encrypted_expanded_round_key_bits,
    std::span<const std::array<uint8_t, 16>> counter_blocks,
    std::span<const std::array<uint8_t, 16>> ciphertext_blocks, const CKKSAesCtrWalshSboxOptions& options = {});


// Served walsh15 AES-CTR with continuous batching (no counterpart in the
// reference, whose bench evaluates one fixed batch).  Lanes of the key owner
// evaluate whole requests packed first come first served into 1024-block
// evaluations; see ckks_aes_serve.cpp for admission.
struct CKKSAesServerOptions {
    size_t lanes = 2;                   // evaluation streams, each a lane of the key owner
    int64_t memory_budget_bytes = 0;    // device bytes the lanes may take (0: 90 % of the free memory)
    size_t max_request_blocks = 1024;   // at most one evaluation
    size_t max_queued_evaluations = 4;  // waiting work admitted per lane, in evaluations
};
struct CKKSAesRequestResult {
    uint64_t id = 0;
    bool refused = false;               // refused at admission (error says why)
    std::string error;                  // empty on success
    std::vector<std::array<uint8_t, 16>> plaintexts;
    std::vector<std::array<double, 128>> values;  // decrypted bit values per block
    double arrival = 0, start = 0, finish = 0;    // steady-clock seconds
    size_t lane = 0;
};
// The packing rule of a lane: the offsets of the queued requests (block
// counts, first come first served) taken whole into one evaluation of
// `capacity` blocks; the first request that does not fit ends the batch.
std::vector<size_t> ckks_aes_serve_take(const std::vector<size_t>& queued_blocks, size_t capacity);
class CKKSAesCtrServer {
public:
    // Encrypts the round keys under `owner` and opens the lanes (a warm-up
    // evaluation each).  The owner must outlive the server.
    CKKSAesCtrServer(CKKSContext& owner, const std::array<uint8_t, 16>& aes_key,
                     const CKKSAesServerOptions& options = {});
    ~CKKSAesCtrServer();  // finishes the admitted requests, then stops
    // Queues a request (or records its refusal); returns its id at once.
    uint64_t submit(std::vector<std::array<uint8_t, 16>> counter_blocks,
                    std::vector<std::array<uint8_t, 16>> ciphertext_blocks);
    // Blocks until the request has completed (or was refused).
    CKKSAesRequestResult wait(uint64_t id);
    size_t laneCount() const;
    int64_t evaluationBytes() const;
    size_t evaluations() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// ckks_aes_ctr.hpp:14-28: AES-128 CTR with the XBOOT monomial S-box and a
// bootstrap of the whole state every round (aes_bench's xboot rows).
// parallelize_independent_work is the reference's OpenMP switch (kept for
// the API); batch_independent_work evaluates the 16 S-boxes of a round and
// the 64 pair bootstraps each as one ciphertext batch (limb for limb the
// one-at-a-time result, which runs when it is false).
struct CKKSAes128CtrOptions {
    bool parallelize_independent_work = true;
    bool batch_independent_work = true;
};
// ckks_aes_ctr.cpp:687-725: round-key bits in every slot at the S2C level
// (round 0 `first_round_key_level` levels higher).
MyVector<CKKSCiphertext> ckks_aes128_encrypt_expanded_round_key_bits(const CKKSContext& context,
                                                                     const std::array<uint8_t, 176>& expanded_round_keys,
                                                                     size_t first_round_key_level);
// ckks_aes_ctr.cpp:757-779: one AES block per slot (block s % active slots),
// 128 refreshed plaintext bit ciphertexts.
MyVector<CKKSCiphertext> ckks_aes_ctr(CKKSContext& context,
                                      const MyVector<CKKSCiphertext>& encrypted_expanded_round_key_bits,
                                      std::span<const std::array<uint8_t, 16>> counter_blocks,
                                      std::span<const std::array<uint8_t, 16>> ciphertext_blocks,
                                      const CKKSAes128CtrOptions& options = {});

// Chebyshev-basis polynomial evaluation (ckks_evalpoly.hpp:9-20): the
// coefficients c_0..c_d of sum c_i T_i(x), x already mapped into [-1, 1].
struct CKKSEvalPolyOptions {
    // Scale assigned to the result; zero keeps the input's scale.
    double target_scale = 0.0;
};
CKKSCiphertext ckks_evalchebyshev_preconditioned(const CKKSContext& context, const CKKSCiphertext& x,
                                                 const std::vector<double>& chebyshev_coefficients,
                                                 const CKKSEvalPolyOptions& options = {});

CKKSLinearTransformPlan ckks_generate_bootstrap_dft_plan(const CKKSContext& context,
                                                         const CKKSBootstrapDftMatrixLiteral& literal);
// A one-stage plan from caller diagonals (index -> N/2 slot values), encoded
// as the DFT stages are (encode_sparse_dft_stage, ckks_bootstrap_dft.cpp:
// 789-917: pre-rotated by the giant offset, Q u P basis, Montgomery form when
// the context's keys are): the QP evaluator for arbitrary matrices.
CKKSLinearTransformPlan ckks_encode_linear_transform_plan(const CKKSContext& context,
                                                          const std::map<size_t, MyVector<Complex>>& diagonals,
                                                          size_t level, size_t baby_step_count, double plaintext_scale,
                                                          bool use_qp_evaluator, bool rescale_after = true);
CKKSCiphertext ckks_apply_linear_transform_owned_stage_qp(const CKKSContext& context, const CKKSCiphertext& ct,
                                                          const CKKSLinearTransformOwnedStage& stage);
CKKSCiphertext ckks_apply_linear_transform_owned_stage_q(const CKKSContext& context, const CKKSCiphertext& ct,
                                                         const CKKSLinearTransformOwnedStage& stage);
CKKSCiphertext ckks_apply_linear_transform_bsgs(const CKKSContext& context, const CKKSCiphertext& ct,
                                                const CKKSLinearTransformStage& stage);
CKKSCiphertext ckks_apply_linear_transform_plan(const CKKSContext& context, CKKSCiphertext ct,
                                                const CKKSLinearTransformPlan& plan);

// ---------------------------------------------------------- bootstrapping
// ckks_keyswitch.cpp:1844-1870, ckks_modraise.cpp:76-110,
// ckks_s2c_generalized.cpp, ckks_c2s_generalized.cpp,
// ckks_bootstrap_evalmod_inplace.cpp:156-434.
void ckks_key_switch_dense_to_sparse_inplace(const CKKSContext& context, CKKSCiphertext& ct);
void ckks_key_switch_sparse_to_dense_inplace(const CKKSContext& context, CKKSCiphertext& ct);

// ---- the keyswitch-level API (ckks_keyswitch.hpp:59-149) ----
// Polynomials are host vectors as in the reference (uploaded, switched on the
// device, downloaded); switching keys are the device-resident CKKSSwitchingKey.
void ckks_apply_ntt_automorphism_to_flat_rns_poly_inplace(MyVector<uint64_t>& poly, size_t N, size_t limb_count,
                                                          std::span<const size_t> dst_to_src_map);
void ckks_hoist_ciphertext_a_qp_eval(const CKKSContext& context, const CKKSCiphertext& ct,
                                     CKKSHoistWorkspace& workspace, CKKSEvalHoistedCiphertextQP& out);
void ckks_hybrid_key_switch_inplace(const CKKSContext& context, size_t level, MyVector<uint64_t>& key_term,
                                    CKKSSwitchingKey switching_key, MyVector<uint64_t>& b_out,
                                    MyVector<uint64_t>& a_out, CKKSHybridKeySwitchCombine b_mode,
                                    CKKSHybridKeySwitchCombine a_mode, bool switching_key_montgomery = false);
void ckks_hybrid_key_switch_hoisted_inplace(const CKKSContext& context,
                                            const CKKSEvalHoistedCiphertextQP& hoisted_key_term,
                                            CKKSSwitchingKey switching_key, MyVector<uint64_t>& b_out,
                                            MyVector<uint64_t>& a_out, CKKSHybridKeySwitchCombine b_mode,
                                            CKKSHybridKeySwitchCombine a_mode, bool switching_key_montgomery,
                                            std::span<const size_t> ntt_map = {});
void ckks_hybrid_key_switch_hoisted_qp_rotated(const CKKSContext& context,
                                               const CKKSEvalHoistedCiphertextQP& hoisted_key_term,
                                               CKKSSwitchingKey switching_key, const MyVector<uint64_t>& b_q_eval,
                                               MyVector<uint64_t>& b_qp_out, MyVector<uint64_t>& a_qp_out,
                                               bool switching_key_montgomery, std::span<const size_t> ntt_map);
void ckks_hybrid_key_switch_qp_from_key_term(const CKKSContext& context, size_t level,
                                             const MyVector<uint64_t>& key_term_q_eval,
                                             CKKSSwitchingKey switching_key, MyVector<uint64_t>& b_qp_out,
                                             MyVector<uint64_t>& a_qp_out, bool switching_key_montgomery);
void ckks_key_switch_inplace(const CKKSContext& context, CKKSCiphertext& ct, CKKSSwitchingKey switching_key,
                             bool switching_key_montgomery = false);
void ckks_modraise_inplace(const CKKSContext& context, CKKSCiphertext& ct, size_t target_level);
CKKSCiphertext ckks_s2c_generalized(const CKKSContext& context, const CKKSCiphertext& ct, size_t stage_count);
CKKSCiphertext ckks_c2s_generalized(const CKKSContext& context, const CKKSCiphertext& ct, double transform_scale);
// EvalMod after the packed S2C: (real, imaginary) halves, each refreshed to
// the top of the chain.
std::pair<CKKSCiphertext, CKKSCiphertext> ckks_batch_bootstrap_evalmod_from_complex_packed(
    const CKKSContext& context, CKKSCiphertext packed);
void ckks_batch_bootstrap_evalmod_inplace(const CKKSContext& context, CKKSCiphertext& ct0, CKKSCiphertext& ct1);
// The steps after S2C, exposed for the parity tests: dense->sparse,
// ModRaise, sparse->dense, trace, C2S and the conjugate split; EvalMod is not
// applied.
std::pair<CKKSCiphertext, CKKSCiphertext> ckks_bootstrap_pre_evalmod(const CKKSContext& context,
                                                                     CKKSCiphertext packed_coeffs);
// One EvalMod (Chebyshev polynomial, then the double-angle steps).
void ckks_evalmod_binary_inplace(const CKKSContext& context, CKKSCiphertext& ct);

// Galois element of a slot rotation (rotation_key_generation.cpp:58-78).
int64_t normalize_rotation_value(size_t N, int64_t rotation);
uint64_t rotation_automorphism_index(size_t N, int64_t rotation);
uint64_t conjugation_automorphism_index(size_t N);

// Throws std::runtime_error (MI_EHIP / MI_ERUNTIME) or std::invalid_argument
// (MI_EINVAL) with mi_last_error() for a failed C-ABI call.
// ---- byte formats (ckks_serialize.cpp).  The reference has no wire or disk
// format; a blob records the context fingerprint (N and the moduli) and a
// checksum, and loads only into a context with the same ring and chain.
std::vector<uint8_t> ckks_serialize_ciphertext(const CKKSCiphertext& ct);
CKKSCiphertext ckks_deserialize_ciphertext(const CKKSContext& context, std::span<const uint8_t> bytes);
std::vector<uint8_t> ckks_serialize_encoding(const CKKSEncoding& pt);
CKKSEncoding ckks_deserialize_encoding(const CKKSContext& context, std::span<const uint8_t> bytes);
// Relinearisation, conjugation, rotation and sparse-secret switching keys.
std::vector<uint8_t> ckks_serialize_evaluation_keys(const CKKSContext& context);
void ckks_load_evaluation_keys(CKKSContext& context, std::span<const uint8_t> bytes);

void mi_check(int rc, const char* fn);
