// Internal context layout shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <set>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace mi {

typedef uint64_t u64;

// Per-modulus constants, indexed by basis index (q_0..q_{nq-1}, p_0..p_{np-1}).
struct BasisConst {
    u64 q;
    u64 mu_lo, mu_hi;   // floor((2^128-1)/q)
    u64 qninv;          // -q^-1 mod 2^64
    u64 n_inv, n_inv_shoup;
    u64 r_mod_q;        // 2^64 mod q
    double inv_d;       // 1.0 / (double)q (HPS estimates)
    u64 negq;           // 2^64 - q
    u64 w1_ninv, w1_ninv_shoup;  // inverse twiddle 1 times N^-1: the last inverse
                                 //   stage fused with the scaling
};

// HPS ModUp tables of one (level, digit): source limbs are the digit's chain
// limbs [s0, s1); destinations are the remaining chain limbs of the level
// followed by the P limbs, in the reference's order.
struct ModUpTables {
    int ns, nd, s0, s1;
    const u64* qhat_inv;        // [ns]
    const u64* qhat_inv_shoup;  // [ns]
    const double* src_inv;      // [ns]
    const u64* qhat_mod_dst;    // [nd][ns], Montgomery form (x 2^64 mod m_dst)
    const u64* alpha_corr;      // [nd][ns+1]
    const int* dst_basis;       // [nd]
};

// Folded ModDown (ckks_keyswitch.cpp:521-607) tables at one level.
struct ModDownTables {
    const u64* pinv;            // [(level+1)][np] p_j^-1 mod q_i, Montgomery form
};

struct RescaleTables {
    const u64* inv_p;           // [level] (q_level)^-1 mod q_i
    const u64* inv_p_shoup;
    const u64* half_ntt;        // [level][N] NTT of the all-floor(p/2) polynomial
    const u64* half_mod;        // [level] floor(p/2) mod q_i
};

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Tuned launch configuration of a context (mi_tune.hip): the NTT kernel
// variant per launch-size bucket (1 = one limb per workgroup, 2 = two limbs
// per workgroup, 3 = one limb with 32-bit LDS exchanges), and key switches
// per launch group.
constexpr int TUNE_BUCKETS = 6;
constexpr int NTT_VARIANTS = 3;
struct TunePlan {
    int ntt_variant[TUNE_BUCKETS] = {1, 1, 1, 1, 1, 1};
    int ks_group = 32;
    bool tuned = false;       // measured (now or in an earlier process)
    bool from_cache = false;  // read from the tuning cache file
};

struct Context {
    u64 N = 0;
    int logN = 0, n1 = 0, n2 = 0;
    int nq = 0, np = 0, dnum = 0, alpha = 0;
    int device = 0;
    std::vector<u64> q, p;
    // device tables
    u64* d_fwd = nullptr;        // [nb][N]
    u64* d_fwd_shoup = nullptr;  // [nb][N]
    u64* d_inv = nullptr;        // [nb][N]
    u64* d_inv_shoup = nullptr;  // [nb][N]
    BasisConst* d_bc = nullptr;  // [nb]
    std::vector<BasisConst> h_bc;
    void* d_blob = nullptr;      // ModUp / ModDown / rescale tables
    size_t blob_bytes = 0;
    std::vector<std::vector<ModUpTables>> modup;  // [level][part]
    std::vector<ModDownTables> moddown;           // [level]
    std::vector<RescaleTables> rescale;           // [level] (level >= 1)
    u64* d_half = nullptr;                        // rescale half_ntt storage
    u64 q0p0_inv = 0, q0p0_inv_shoup = 0;         // p_0^-1 mod q_0 (level-0 switch)
    double* d_enc_tw = nullptr;                   // encoding FFT twiddles [N] complex
    int* d_enc_inv = nullptr;                     // encoding slot placement [N]
    double* d_dec_tw = nullptr;                   // decoding (inverse special FFT) twiddles [N] complex
    int* d_dec_perm = nullptr;                    // decoding slot permutation [N/2]
    std::mutex gauss_mu;                          // discrete Gaussian table of the last sigma
    double gauss_sigma = 0.0;
    double* d_gauss_cdf = nullptr;
    int gauss_bound = 0, gauss_n = 0;
    size_t table_bytes = 0;
    std::mutex ws_mu;
    std::map<hipStream_t, Workspace> ws;
    // Workspaces replaced while a graph is alive (a graph may replay into
    // them); freed when the last graph is destroyed, or with the context.
    std::vector<void*> retired;
    // Side streams (default priority) that independent groups of a batch
    // run on concurrently; created on first use, joined back with events.
    std::mutex side_mu;
    std::vector<hipStream_t> side;
    // Device memory pool for ciphertexts, plaintexts and keys: freed blocks
    // are cached by size and reused; byte counts feed admission decisions.
    std::mutex pool_mu;
    std::multimap<size_t, void*> pool_free;
    std::map<void*, size_t> pool_live;
    size_t pool_live_bytes = 0, pool_peak_bytes = 0, pool_cached_bytes = 0;
    size_t pool_limit = 0;  // 0 = no limit
    // Graph capture (pool_mu): blocks allocated or released while a capture
    // is open can be addressed by the graph, so once released they are held
    // back from reuse until no capture is open and no graph is alive.
    int capturing = 0, live_graphs = 0;
    std::set<void*> captured_blocks;
    std::vector<std::pair<size_t, void*>> graph_held;

    // NTT probe (mi_probe_begin / mi_probe_end): while on, every forward or
    // inverse NTT launch pair is bracketed by events on its stream and its
    // algorithmic bytes (16 N per limb) are counted.  Work that would spread
    // over the side streams runs on the caller's stream meanwhile, so each
    // timed pair has the device to itself; mi_probe_end restores the setting.
    struct NttProbe {
        bool on = false;
        bool saved_multi_stream = true;
        std::vector<hipEvent_t> events;  // start/stop pairs, reused across probes
        size_t used = 0;
        double bytes = 0.0;
        long long limbs = 0;
    };
    mutable std::mutex probe_mu;
    mutable NttProbe probe;
    // probe_open reserves an event pair for a launch pair of `limbs` limbs and
    // records its start event on s (returns -1 when the probe is off);
    // probe_close records the pair's stop event on s.
    long probe_open(hipStream_t s, long long limbs) const;
    void probe_close(long slot, hipStream_t s) const;

    // NTT plan: the kernel variant (1..NTT_VARIANTS) of a pass over `count`
    // limbs.  ntt_forced (environment MICKKS_NTT_VARIANT) overrides the plan.
    int ntt_forced = 0;
    // Key switches fold the key inner product into ModDown's NTT passes
    // (MICKKS_FUSE_MAC=0: the separate k_mac pass, for comparison).
    bool fuse_mac = true;
    // Independent chunks of a batch run on the side streams too
    // (MICKKS_MULTI_STREAM=0: all on the caller's stream).
    bool multi_stream = true;
    u64* d_zero = nullptr;  // N zero words (a zero row for branch-free jobs)
    int ntt_variant(int count) const;
    TunePlan tune;
    std::string tune_key;  // "<gfx arch> <compute units> <log2 N>"

    int nb() const { return nq + np; }
    int parts(int level) const {
        int L = level + 1;
        int pp = (L + alpha - 1) / alpha;
        return pp < dnum ? pp : dnum;
    }
    void* workspace(hipStream_t s, size_t bytes);
    const std::vector<hipStream_t>& side_streams();
    void* pool_alloc(size_t bytes);
    void pool_release(void* p);
    void pool_trim();
    void capture_begin();
    // made_graph: the capture produced an instantiated graph
    void capture_end(bool made_graph);
    void graph_destroyed();
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MI_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::mi::Error(-3, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Division of a uniform index k < 2^20 by d < 2^20 as a multiply and shift
// with m = ceil(2^40 / d): k*m / 2^40 exceeds k/d by less than 2^-20, which is
// below the gap 1/d between frac(k/d) and 1, so the floor is exact.  Block
// index arithmetic in the NTT jobs uses it so the quotient stays on the
// scalar unit (a runtime '/' is expanded on the vector unit).
struct FastDiv {
    u64 m;
    __host__ __device__ __forceinline__ int div(int k) const { return (int)(((u64)(unsigned)k * m) >> 40); }
};
inline FastDiv fast_div(int d) {
    if (d < 1 || d >= (1 << 20)) throw Error(-1, "batch dimension out of range");
    return FastDiv{((1ull << 40) + (u64)d - 1) / (u64)d};
}

// ---- Walsh S-box product sum (mi_walsh.hip) ----
// 15 sign ciphertexts per nibble (masks 1..15), coefficient tables [16][16]
// (int16, [high mask][low mask]), rounded constants |V| (lo, hi) and signs
// for (h, real) at 2h and (h, imag) at 2h + 1.
void walsh_product_sum(const Context& c, int level, const u64* const* low_b, const u64* const* low_a,
                       const u64* const* high_b, const u64* const* high_a, const short* cr, const short* ci,
                       const u64* v_abs, const int* v_neg, const u64* pt, bool pt_mont, u64* acc_b, u64* acc_a,
                       u64* raw_b, u64* raw_a, u64* raw_c2, hipStream_t s);

// ---- NTT launchers (mi_ntt.hip) ----
// A batch of limbs: limb k of the batch lives at base + (k / lpp) * stride +
// (k % lpp) * N and uses basis index basis0 + (k % lpp).
struct LimbBatch {
    u64* base;
    long long stride;
    int lpp;
    int basis0;
    int count;
};
void ntt_forward(const Context& c, const LimbBatch& b, hipStream_t s);
void ntt_inverse(const Context& c, const LimbBatch& b, hipStream_t s);

Context* context_create(u64 N, const u64* q, int nq, const u64* p, int np, int dnum);
// Decoding on the device (mi_decode.hip): tables from the host context, then
// CRT + scale division + inverse special FFT of `count` coefficient-domain
// plaintexts (limbs x N each); out: count x N/2 complex (2 doubles each).
void set_decoding_tables(Context& c, const double* twiddles, const u64* perm);
void decode_batch(Context& c, int count, const u64* const* coeffs, int limbs, const u64* scale_mant,
                  const int* scale_exp, bool coeff_domain, double* out, hipStream_t s);
// Installs the tuned plan for the context's device and ring dimension,
// measuring it first when neither this process nor the cache file has it.
void tune_context(Context& c);
int tune_bucket(int count);
void context_destroy(Context* c);

// ---- evaluator ops (mi_ops.hip) ----
enum EwOp { EW_ADD = 0, EW_SUB = 1, EW_MUL = 2, EW_MUL_MONT = 3 };
// x[k] (op)= y[k] (y stride may be 0 to broadcast one poly)
void ew_binary(const Context& c, int op, u64* x, const u64* y, int limbs, int npolys,
               long long xstride, long long ystride, hipStream_t s, int basis0 = 0);
// x[k][l] = x[k][l] (op) scalar[l]   (op: add or mul), scalars on the host
constexpr int MAX_LIMBS = 64;
void ew_scalar(const Context& c, int op, u64* x, const u64* scalars_host, int limbs, int npolys,
               long long stride, hipStream_t s, int basis0 = 0);
void ew_scalar_mac(const Context& c, u64* dst, const u64* src, const u64* scalars_host, int limbs,
                   int npolys, long long dstride, long long sstride, hipStream_t s);
void tensor(const Context& c, int limbs, u64* b1, u64* a1, const u64* b2, const u64* a2, u64* c2,
            hipStream_t s);
void square(const Context& c, int limbs, u64* b, u64* a, u64* c2, hipStream_t s);
// listed tensor products (b2 == nullptr: squares), level+1 limbs each; the
// first operand read from (s1b, s1a) when given, else from (b1, a1)
void tensor_list(const Context& c, int level, int count, u64* const* b1, u64* const* a1, const u64* const* b2,
                 const u64* const* a2, u64* const* c2, hipStream_t s, const u64* const* s1b = nullptr,
                 const u64* const* s1a = nullptr);
void automorphism(const Context& c, int limbs, u64 galois, const u64* src, u64* dst, int npolys,
                  long long sstride, long long dstride, hipStream_t s);
// dst (+)= automorphism(src); src and dst must not overlap
void automorphism_acc(const Context& c, int limbs, u64 galois, const u64* src, u64* dst, int npolys,
                      long long sstride, long long dstride, bool accumulate, hipStream_t s);
void rescale(Context& c, int level, u64* polys, int npolys, long long stride, hipStream_t s);
// the same for listed polys (any number; launches of up to 64)
void rescale_list(Context& c, int level, u64* const* polys, int npolys, hipStream_t s);
// polys[k] *= pt over limbs 0..level (pt Montgomery form if mont)
void linear_combine(const Context& c, int level, int nout, u64* const* out, int nterms, const u64* const* in,
                    const unsigned* coef, hipStream_t s);
void scaled_sum(const Context& c, int level, int npolys, u64* const* out, int nterms, const u64* const* in,
                const u64* residues, int nconst, const u64* constant, hipStream_t s);
void pair_list_op(const Context& c, int op, int level, u64* const* x, const u64* const* y, int npolys,
                  const u64* residues, hipStream_t s);
void mul_plain_list(const Context& c, int level, u64* const* polys, int npolys, const u64* pt, bool mont,
                    hipStream_t s);
// bytes of zeros at p by a kernel (no memset node, so safe inside a captured graph)
void fill_zero(void* p, size_t bytes, hipStream_t s);
void add_plain_mont(const Context& c, int limbs, u64* b, const u64* pt, hipStream_t s);

// ---- key switching (mi_keyswitch.hip) ----
struct KsItem {
    const u64* key_term;
    const u64* key;
    u64* b_out;
    u64* a_out;
    int level, b_mode, a_mode;
    const u64* init_b = nullptr;  // optional L-limb term added to the switched b before it is combined
};
void keyswitch(Context& c, const KsItem& it, bool montgomery, hipStream_t s);

// Items per batched key-switch launch (pointer lists travel as kernel
// arguments, so a launch needs no device-side argument buffer and can be
// captured in a graph).
constexpr int KS_MAX_ITEMS = 32;

// Key inner product over the QP slots of one level, for `items` items that
// share the level.  Item i reads digits + i * digits_stride and key[i].
struct MacArgs {
    const u64* digits;                 // [parts][L+K][N] evaluation domain, per item
    long long digits_stride;
    const u64* key[KS_MAX_ITEMS];      // [parts][2 (b, a)][nq+np][N]
    u64* b_out[KS_MAX_ITEMS];          // split: L limbs; qp_out: L+K limbs
    u64* a_out[KS_MAX_ITEMS];
    u64* accp;                         // split only: item i's P slots of a at
    long long accp_stride;             //   accp + i * accp_stride, of b K*N later
    const u64* init_b[KS_MAX_ITEMS];   // Q-limb b term added to the b result before it is combined
                                       //   (qp_out: read through the automorphism)
    const u64* add_b[KS_MAX_ITEMS];    // qp_out: L+K-limb b term added (read through the automorphism)
    u64 galois;                        // qp_out: 0, or the Galois element applied to the result
    int items;
    int L, K, nq, parts, logN;
    int b_mode, a_mode;                // split: 0 = add into outputs, 1 = assign
    int montgomery;
    int qp_out;                        // 1: whole result in the QP basis
    int accumulate;                    // qp_out: add into the outputs instead of assigning
    const BasisConst* bc;
};
MacArgs mac_args(const Context& c, int level, const u64* digits, const u64* key, bool montgomery);
// Device steps (scratch supplied by the caller):
//   ks_modup_n: n <= KS_MAX_ITEMS key terms (L limbs each, evaluation
//     domain); C = n*L*N words; item i's digits at D + i*dstride
//     (parts*(L+K)*N words, canonical).
//   ks_moddown_n: npoly <= 2*KS_MAX_ITEMS polys; poly k's P limbs at p[k]
//     (K*N words, destroyed); out[k] = base[k] - NTT(contribution_k), base
//     may alias out; T = npoly*L*N words.
void ks_modup_n(Context& c, int level, int n, const u64* const* key_terms, u64* D, long long dstride,
                u64* C, hipStream_t s, bool cols_only = false);
// ModDown of n key switches (one level and modes) with the key inner
// product of their digits folded in (split form); at most 3 digits.
void ks_moddown_mac(Context& c, int level, int n, const KsItem* items, const u64* D, long long dstride, u64* AP,
                    u64* T, bool montgomery, hipStream_t s);
void ks_modup(Context& c, int level, const u64* key_term, u64* D, u64* C, hipStream_t s);
void ks_mac(const Context& c, const MacArgs& a, int level, hipStream_t s);
void ks_moddown_n(Context& c, int level, int npoly, u64* const* p, const u64* const* base, u64* const* out,
                  u64* T, hipStream_t s);
void ks_moddown(Context& c, int level, u64* p0, u64* p1, const u64* base0, const u64* base1, u64* out0,
                u64* out1, u64* T, hipStream_t s);
// A batch of independent key switches at one level with one pair of combine
// modes (callers group by these): items[i] outputs must not overlap another
// item's inputs or outputs.
void keyswitch_items(Context& c, const KsItem* items, int n, bool montgomery, hipStream_t s);
// The same with caller scratch of n * ks_item_words(c, level) words.
size_t ks_item_words(const Context& c, int level);
void keyswitch_items_with(Context& c, const KsItem* items, int n, bool montgomery, u64* ws, hipStream_t s);
// Whole batch: items are grouped by (level, modes) into launches of up to
// KS_MAX_ITEMS, keeping the order of items whose buffers overlap.
void keyswitch_batch(Context& c, const KsItem* items, int n, bool montgomery, hipStream_t s);
// In-place conjugations of independent ciphertexts at one level: one
// automorphism launch per KS_MAX_ITEMS ciphertexts into scratch, then their
// key switches as one batch (the conjugated b enters as the switch's init_b).
void conjugate_batch(Context& c, int level, int n, u64* const* b, u64* const* a, const u64* key, bool montgomery,
                     hipStream_t s);
// One key switch with its scratch in the library's fixed device array
// (mi_noalloc.hip): no allocation.  noalloc_setup takes that array's address
// on `device` (current) once; keyswitch_noalloc refuses a switch whose
// scratch exceeds noalloc_scratch_words().
void noalloc_setup(int device);
size_t noalloc_scratch_words();
void keyswitch_noalloc(Context& c, const KsItem& it, bool montgomery, hipStream_t s);

// Hoisting and QP-basis switches (ckks_keyswitch.cpp:344-373, 675-1328).
void hoist(Context& c, int level, const u64* a_eval, u64* digits, hipStream_t s);
void hoist_batch(Context& c, int level, int n, const u64* const* a_eval, u64* digits, long long dstride,
                 hipStream_t s);
void keyswitch_hoisted(Context& c, int level, const u64* digits, const u64* key, u64* b_out,
                       u64* a_out, int b_mode, int a_mode, u64 galois, bool montgomery, hipStream_t s);
// Rotations i < n of one hoisted ciphertext (digits of a, b): (b_out[i],
// a_out[i]) = g_i(b + KS_i(a)) with galois[i] and keys[i]; outputs must not
// overlap digits or b.
void rotate_batch(Context& c, int level, int n, u64* const* b, u64* const* a, const u64* galois,
                  const u64* const* keys, bool montgomery, hipStream_t s);
// Hoisted rotations, item i from the digits at digits + (i mod period) *
// dstride (dstride 0: one shared hoist) with b[i] added to its b.
void rotate_hoisted_items(Context& c, int level, const u64* digits, long long dstride, int period,
                          const u64* const* b, int n, const u64* galois, const u64* const* keys, u64* const* b_out,
                          u64* const* a_out, bool montgomery, hipStream_t s);
void rotate_hoisted_batch(Context& c, int level, const u64* digits, const u64* b, int n, const u64* galois,
                          const u64* const* keys, u64* const* b_out, u64* const* a_out, bool montgomery,
                          hipStream_t s);
void keyswitch_hoisted_qp(Context& c, int level, const u64* digits, const u64* key, const u64* b_q,
                          u64* b_qp, u64* a_qp, u64 galois, bool montgomery, hipStream_t s);
void keyswitch_qp(Context& c, int level, const u64* key_term, const u64* key, u64* b_qp, u64* a_qp,
                  bool montgomery, hipStream_t s);
void moddown_qp(Context& c, int level, u64* qp, hipStream_t s);

// BSGS linear-transform stage in the QP basis (mi_lt.hip).
u64 rotation_galois(u64 N, long long rotation);
struct LtStage {
    int level, nterms;
    const long long* diag_idx;     // host [nterms]
    const u64* const* diags;       // host array of device pointers, each (L+K)*N
    int pt_montgomery, baby_steps, slots;
    int nkeys;
    const long long* key_rot;      // host [nkeys]
    const u64* const* keys;        // host array of device pointers
    int montgomery;
};
void linear_transform_qp(Context& c, const LtStage& st, const u64* a, const u64* b, u64* a_out,
                         u64* b_out, hipStream_t s);
// The same stage applied to n ciphertexts at the stage's level (any n; they
// are processed KS_MAX_ITEMS at a time).
void linear_transform_qp_batch(Context& c, const LtStage& st, int n, const u64* const* a,
                               const u64* const* b, u64* const* a_out, u64* const* b_out, hipStream_t s);
// BSGS stage on the Q evaluator (ckks_apply_linear_transform_bsgs,
// ckks_linear_transform.cpp:818-996): baby steps are full hoisted rotations
// in Q, diagonals are level+1 limb plaintexts, each giant step's inner sum is
// rotated by a full key switch.
void linear_transform_q_batch(Context& c, const LtStage& st, int n, const u64* const* a,
                              const u64* const* b, u64* const* a_out, u64* const* b_out, hipStream_t s);

// Slot encoding on the device (mi_encode.hip): twiddles = N complex doubles
// (re, im), perm = the N/2-entry slot permutation; values = count x N/2
// complex doubles on the host; out poly b at out + b*stride.
void set_encoding_tables(Context& c, const double* twiddles, const u64* perm);
void encode_slots(Context& c, const double* values, int count, double scale, int limbs, bool eval, u64* out,
                  long long stride, hipStream_t s);

// Samplers (mi_encode.hip), counter-based (Philox4x32-10): a limb range of
// uniform residues (basis indices basis0..), and one discrete Gaussian
// integer per coefficient as residues of limbs 0..limbs-1.
void sample_uniform(Context& c, u64* out, int limbs, int basis0, u64 seed, u64 counter, hipStream_t s);
void sample_gaussian(Context& c, u64* out, int limbs, double sigma, u64 seed, u64 counter, hipStream_t s);

// Bootstrapping pieces (mi_boot.hip).
void modraise(Context& c, int target_level, const u64* poly_q0, u64* out, hipStream_t s);
void modraise_batch(Context& c, int target_level, int n, const u64* const* poly_q0, u64* const* out, hipStream_t s);
void q0p0_keyswitch_batch(Context& c, int n, u64* const* b, u64* const* a, const u64* key, bool montgomery,
                          hipStream_t s);
void q0p0_keyswitch(Context& c, u64* b, u64* a, const u64* key, bool montgomery, hipStream_t s);

// Independent chunks of a batch spread over the caller's stream and the
// context's side streams (each with its own workspace of `ws_bytes`): the
// side streams start after the work already on s, and s continues after all
// of them.  A chunk's launches overlap the fill and drain of another's.
// fn(stream, workspace, chunk index).
template <class F>
inline void over_streams(Context& c, hipStream_t s, int nchunks, size_t ws_bytes, F&& fn) {
    if (nchunks <= 1 || !c.multi_stream) {
        u64* ws = (u64*)c.workspace(s, ws_bytes);
        for (int i = 0; i < nchunks; ++i) fn(s, ws, i);
        return;
    }
    const auto& side = c.side_streams();
    const int ns = std::min(nchunks, 1 + (int)side.size());
    std::vector<hipStream_t> st(ns);
    std::vector<u64*> ws(ns);
    st[0] = s;
    for (int k = 1; k < ns; ++k) st[k] = side[k - 1];
    for (int k = 0; k < ns; ++k) ws[k] = (u64*)c.workspace(st[k], ws_bytes);
    hipEvent_t fork;
    MI_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    MI_HIP(hipEventRecord(fork, s));
    for (int k = 1; k < ns; ++k) MI_HIP(hipStreamWaitEvent(st[k], fork, 0));
    for (int i = 0; i < nchunks; ++i) fn(st[i % ns], ws[i % ns], i);
    for (int k = 1; k < ns; ++k) {
        hipEvent_t j;
        MI_HIP(hipEventCreateWithFlags(&j, hipEventDisableTiming));
        MI_HIP(hipEventRecord(j, st[k]));
        MI_HIP(hipStreamWaitEvent(s, j, 0));
        MI_HIP(hipEventDestroy(j));
    }
    MI_HIP(hipEventDestroy(fork));
}

}  // namespace mi
