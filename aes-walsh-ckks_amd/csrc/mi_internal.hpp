// Internal context layout shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
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
};

// HPS ModUp tables of one (level, digit): source limbs are the digit's chain
// limbs [s0, s1); destinations are the remaining chain limbs of the level
// followed by the P limbs, in the reference's order.
struct ModUpTables {
    int ns, nd, s0, s1;
    const u64* qhat_inv;        // [ns]
    const u64* qhat_inv_shoup;  // [ns]
    const double* src_inv;      // [ns]
    const u64* qhat_mod_dst;    // [nd][ns]
    const u64* alpha_corr;      // [nd][ns+1]
    const int* dst_basis;       // [nd]
};

// Folded ModDown (ckks_keyswitch.cpp:521-607) tables at one level.
struct ModDownTables {
    const u64* pinv;            // [(level+1)][np] p_j^-1 mod q_i
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
    size_t table_bytes = 0;
    std::mutex ws_mu;
    std::map<hipStream_t, Workspace> ws;

    int nb() const { return nq + np; }
    int parts(int level) const {
        int L = level + 1;
        int pp = (L + alpha - 1) / alpha;
        return pp < dnum ? pp : dnum;
    }
    void* workspace(hipStream_t s, size_t bytes);
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
void context_destroy(Context* c);

// ---- evaluator ops (mi_ops.hip) ----
enum EwOp { EW_ADD = 0, EW_SUB = 1, EW_MUL = 2, EW_MUL_MONT = 3 };
// x[k] (op)= y[k] (y stride may be 0 to broadcast one poly)
void ew_binary(const Context& c, int op, u64* x, const u64* y, int limbs, int npolys,
               long long xstride, long long ystride, hipStream_t s);
// x[k][l] = x[k][l] (op) scalar[l]   (op: add or mul), scalars on the host
constexpr int MAX_LIMBS = 64;
void ew_scalar(const Context& c, int op, u64* x, const u64* scalars_host, int limbs, int npolys,
               long long stride, hipStream_t s);
void tensor(const Context& c, int limbs, u64* b1, u64* a1, const u64* b2, const u64* a2, u64* c2,
            hipStream_t s);
void square(const Context& c, int limbs, u64* b, u64* a, u64* c2, hipStream_t s);
void automorphism(const Context& c, int limbs, u64 galois, const u64* src, u64* dst, int npolys,
                  long long sstride, long long dstride, hipStream_t s);
void rescale(Context& c, int level, u64* polys, int npolys, long long stride, hipStream_t s);
void add_plain_mont(const Context& c, int limbs, u64* b, const u64* pt, hipStream_t s);

// ---- key switching (mi_keyswitch.hip) ----
struct KsItem {
    const u64* key_term;
    const u64* key;
    u64* b_out;
    u64* a_out;
    int level, b_mode, a_mode;
};
void keyswitch(Context& c, const KsItem& it, bool montgomery, hipStream_t s);

}  // namespace mi
