/*
 * mickks.h -- C-ABI of the MI355X (gfx950) CKKS evaluation library.
 *
 * Plain extern "C" entry points over device pointers and sizes.  Every call
 * returns 0 on success or a negative MI_E* code; mi_last_error() gives the
 * thread's last message, prefixed "<function>: " like the reference's
 * exceptions.  Streams are passed as void* (a hipStream_t; NULL = default).
 *
 * Data conventions are the reference's (aes-walsh-ckks src/): an RNS
 * polynomial is limb-major uint64[limbs][N]; evaluation-domain index i holds
 * A(psi^(2*brev(i)+1)); stored residues are canonical [0, q).  A ciphertext is
 * the polys [b, a, (c2)].  A switching key is uint64[parts][2][nq+np][N]
 * (poly 0 = b, poly 1 = a, Q limbs 0..nq-1 then P limbs), i.e. the
 * reference's key digits (CKKSCiphertext at the maximum level over Q u P,
 * ckks_keyswitch.cpp:20-35) stored back to back, already folded and (when
 * montgomery=1) in Montgomery form as CKKSContext generates them.
 *
 * Basis indices: 0..nq-1 name the chain moduli q_i, nq..nq+np-1 the special
 * moduli p_j.
 */
#ifndef MICKKS_H
#define MICKKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MI_OK = 0,
    MI_EINVAL = -1,   /* bad argument (reference: std::invalid_argument) */
    MI_ERUNTIME = -2, /* missing precompute / internal (std::runtime_error) */
    MI_EHIP = -3      /* HIP runtime failure */
};

typedef struct mi_context mi_context;

/* Thread-local text of the last error. */
const char* mi_last_error(void);

/* ---- context ------------------------------------------------------------
 * Replaces the device-relevant part of CKKSContext's constructor
 * (ckks_context.cpp:603-705): NTT twiddle tables (ntt.cpp:54-148), HPS ModUp
 * tables per (level, digit) (ckks_context.cpp:468-535,
 * relinearization_key_generation.cpp:45-129), folded P^-1 tables
 * (ckks_context.cpp:537-560) and rescale constants (ckks_context.cpp:565-599),
 * all uploaded once to HBM.  q: chain moduli (limb 0..Lmax), p: special moduli.
 */
int mi_context_create(mi_context** out, uint64_t N, const uint64_t* q, int nq,
                      const uint64_t* p, int np, int dnum);
int mi_context_destroy(mi_context* ctx);
/* info[0..5] = N, nq, np, dnum, alpha, device id */
int mi_context_info(const mi_context* ctx, int64_t* info);
/* Device memory the context holds (tables + workspaces), in bytes. */
int64_t mi_context_device_bytes(const mi_context* ctx);

/* ---- device memory ----------------------------------------------------------
 * Replaces the host MyVector storage of CKKSCiphertext / CKKSEncoding and the
 * context's key vectors (ckks_ciphertext.hpp:18-108, ckks_context.hpp:94-133):
 * a per-context pool with byte accounting, used by the C++ host layer and by
 * admission in the serving loop.  A freed block may be handed out again at
 * once, so free a block only after the work using it is stream-ordered before
 * any later use.  mi_memory_stats: live, peak, cached, limit (bytes).  With a
 * limit set, an allocation that would exceed it fails with MI_ERUNTIME. */
int mi_alloc(mi_context* ctx, size_t bytes, void** out);
int mi_free(mi_context* ctx, void* ptr);
int mi_memory_trim(mi_context* ctx);
int mi_memory_stats(mi_context* ctx, int64_t* stats);
int mi_set_memory_limit(mi_context* ctx, int64_t bytes);
/* Autotuned launch plan (no counterpart in the reference): at context
 * creation the library times the candidate configurations of its tunable
 * launches on the device (the NTT kernel variant per launch-size bucket of
 * 0-3, 4-15, 16-63, 64-255, 256-1023 and 1024+ limbs: 1 = one limb per
 * workgroup, 2 = two limbs per workgroup, 3 = one limb with 32-bit LDS
 * exchanges; key switches per launch group, 8/16/32) and keeps the fastest,
 * cached per device and ring dimension in memory and in $MICKKS_TUNE_CACHE
 * (else $HOME/.cache/mickks/tune.txt); MICKKS_TUNE=0 keeps the defaults and
 * MICKKS_NTT_VARIANT=v forces one NTT variant.  Every plan computes the same
 * residues.  mi_tune_report writes the context's plan as a JSON object into
 * buf; mi_tune_set replaces it (six variants, each 1, 2 or 3, and a group
 * size). */
int mi_tune_report(mi_context* ctx, char* buf, size_t len);
int mi_tune_set(mi_context* ctx, const int* ntt_variant, int keyswitch_group);
/* NTT probe (measurement only; no counterpart in the reference): between
 * mi_probe_begin and mi_probe_end every forward or inverse NTT launch pair
 * of the context (standalone, or inside key switches, rescales, ModRaise and
 * linear transforms) is bracketed by HIP events on its stream, and work that
 * would spread over the side streams stays on the caller's stream, so each
 * timed pair has the device to itself (mi_probe_end restores the spreading;
 * outputs are the same either way).  mi_probe_end
 * waits for them and returns stats[0] = summed event time of the pairs (ms),
 * stats[1] = launch pairs, stats[2] = algorithmic bytes (16 N per limb: each
 * limb read and written once), stats[3] = limbs transformed. */
int mi_probe_begin(mi_context* ctx);
int mi_probe_end(mi_context* ctx, double* stats);
/* Free and total bytes of the context's device (hipMemGetInfo): what the
 * serving loop sizes its lanes against.  No counterpart in the reference,
 * whose data lives in host memory. */
int mi_device_memory_info(mi_context* ctx, int64_t* free_bytes, int64_t* total_bytes);
enum { MI_COPY_H2D = 0, MI_COPY_D2H = 1, MI_COPY_D2D = 2 };
/* Host copies complete before the call returns; device-to-device copies are
 * stream-ordered. */
int mi_memcpy(mi_context* ctx, void* dst, const void* src, size_t bytes, int kind, void* stream);
int mi_stream_synchronize(mi_context* ctx, void* stream);
/* A stream of default priority that synchronises with the null stream (no
 * counterpart in the reference, whose evaluator runs on host threads): each
 * C++ context keeps one, so that contexts driven from separate host threads
 * run concurrently.  Destroying it waits for its work first. */
int mi_stream_create(mi_context* ctx, void** out);
int mi_stream_destroy(mi_context* ctx, void* stream);

/* ---- element-wise over a basis range (limb l uses basis index basis0 + l) ---
 * pointwise_multiply*_dispatch and poly_add/sub_modq (ntt.cpp:450-579,
 * polynomial.cpp:6-32) as key generation, key folding and plaintext
 * Montgomery conversion use them over Q u P (hybrid_key_folding.hpp:50-77,
 * hybrid_key_montgomery.hpp:12-36, ckks_encoding.cpp:172-190).
 * op: x (op)= y, or x (op)= residues[l]. */
enum { MI_OP_ADD = 0, MI_OP_SUB = 1, MI_OP_MUL = 2, MI_OP_MUL_MONT = 3 };
int mi_poly_op(mi_context* ctx, int op, uint64_t* x, const uint64_t* y, int limbs, int basis0,
               int npolys, int64_t poly_stride, void* stream);
int mi_poly_scalar(mi_context* ctx, int op, uint64_t* x, const uint64_t* residues, int limbs,
                   int basis0, int npolys, int64_t poly_stride, void* stream);

/* ---- negacyclic NTT (ntt.cpp:235-447) -----------------------------------
 * Batched over npolys polys of `limbs` limbs each; limb l of every poly uses
 * basis index basis0+l; polys are poly_stride words apart (>= limbs*N).
 * Forward: natural -> bit-reversed evaluation order.  Inverse includes N^-1.
 * Inputs and outputs canonical. */
int mi_ntt_forward(mi_context* ctx, uint64_t* data, int npolys, int limbs, int64_t poly_stride,
                   int basis0, void* stream);
int mi_ntt_inverse(mi_context* ctx, uint64_t* data, int npolys, int limbs, int64_t poly_stride,
                   int basis0, void* stream);

/* ---- element-wise ciphertext ops (level = limbs-1 over q_0..q_level) ----- */
/* ckks_mul_inplace(ct, ct) tensor, ckks_mul.cpp:90-157: (b1,a1) <- (b1 b2, a1 b2 + b1 a2), c2 <- a1 a2 */
int mi_tensor(mi_context* ctx, int level, uint64_t* b1, uint64_t* a1, const uint64_t* b2,
              const uint64_t* a2, uint64_t* c2, void* stream);
/* ckks_square_inplace, ckks_mul.cpp:37-88 */
int mi_square(mi_context* ctx, int level, uint64_t* b, uint64_t* a, uint64_t* c2, void* stream);
/* ckks_mul_inplace(ct, ct) / ckks_square_inplace (ckks_mul.cpp:37-157) of
 * `count` listed ciphertext pairs at one level, one launch per 16 pairs: for
 * each k, (b1, a1, c2)[k] as mi_tensor (b2/a2 non-null) or mi_square (b2 and
 * a2 null).  The reference evaluates such products one ciphertext at a time
 * on OpenMP threads; here a batch of ciphertexts that share their metadata
 * (the same schedule applied to independent inputs) is one launch.  No two
 * output polys may overlap. */
int mi_tensor_list(mi_context* ctx, int level, int count, uint64_t* const* b1, uint64_t* const* a1,
                   const uint64_t* const* b2, const uint64_t* const* a2, uint64_t* const* c2, void* stream);
/* mi_tensor_list into new polys: (b_out, a_out, c2)[k] = the tensor of
 * (b1, a1)[k] with (b2, a2)[k], or its square (b2 and a2 null), the inputs
 * read at `level` (their first level+1 limbs) and left unchanged.  It
 * replaces the reference's copy-then-multiply of the power basis
 * (ckks_evalpoly.cpp:908-921: T_a cloned, dropped to the level, then
 * ckks_square_inplace / ckks_mul_inplace), whose product does not depend on
 * the copy.  Outputs pairwise disjoint and overlapping no input. */
int mi_tensor_list_into(mi_context* ctx, int level, int count, const uint64_t* const* b1, const uint64_t* const* a1,
                        const uint64_t* const* b2, const uint64_t* const* a2, uint64_t* const* b_out,
                        uint64_t* const* a_out, uint64_t* const* c2, void* stream);
/* ckks_mul_inplace(ct, pt), ckks_mul.cpp:160-212 (pt_montgomery: plaintext stored x 2^64) */
int mi_mul_plain(mi_context* ctx, int level, uint64_t* polys, int npolys, int64_t poly_stride,
                 const uint64_t* pt, int pt_montgomery, void* stream);
/* ckks_mul_inplace(ct, int64), ckks_mul.cpp:214-272: residues[i] = scalar mod q_i (host array) */
int mi_mul_scalar(mi_context* ctx, int level, uint64_t* polys, int npolys, int64_t poly_stride,
                  const uint64_t* residues, void* stream);
/* fused_multiply_by_real_scalar_then_add_inplace, ckks_evalpoly.cpp:352-398:
 * dst += src * residues[i] per limb i (host residues, the rounded scaled
 * constant mod q_i) */
int mi_mul_scalar_add(mi_context* ctx, int level, uint64_t* dst, const uint64_t* src, int npolys,
                      int64_t poly_stride, const uint64_t* residues, void* stream);
/* zero polynomials (zero_ciphertext_like, ckks_evalpoly.cpp:258-275) */
int mi_memset_zero(mi_context* ctx, void* ptr, size_t bytes, void* stream);
/* ckks_add_inplace / ckks_sub_inplace (ct, ct), ckks_addition.cpp:62-139 */
int mi_add(mi_context* ctx, int level, uint64_t* x, const uint64_t* y, int npolys,
           int64_t poly_stride, void* stream);
int mi_sub(mi_context* ctx, int level, uint64_t* x, const uint64_t* y, int npolys,
           int64_t poly_stride, void* stream);
/* ckks_add_inplace(ct, pt), ckks_addition.cpp:35-60 */
int mi_add_plain(mi_context* ctx, int level, uint64_t* b, const uint64_t* pt, int pt_montgomery,
                 void* stream);
/* ckks_add_integer_constant_inplace, ckks_addition.cpp:89-112 (host residues per limb) */
int mi_add_scalar(mi_context* ctx, int level, uint64_t* b, const uint64_t* residues, void* stream);

/* ---- automorphism (ckks_keyswitch.cpp:37-60, rotation_key_generation.cpp:80-90)
 * dst[i] = src[map_g(i)] in the evaluation domain for Galois element g
 * (5^r mod 2N for a rotation by r, 2N-1 for conjugation). src != dst. */
int mi_automorphism(mi_context* ctx, int limbs, uint64_t galois, const uint64_t* src,
                    uint64_t* dst, int npolys, int64_t poly_stride, void* stream);

/* ---- rescale (ckks_rescale.cpp:68-148) ------------------------------------
 * Divides by q_level with rounding; the first level*N words of each poly
 * hold the result (stride unchanged). */
int mi_rescale(mi_context* ctx, int level, uint64_t* polys, int npolys, int64_t poly_stride,
               void* stream);

/* ---- hybrid key switching (ckks_keyswitch.cpp:1330-1641) ------------------
 * key_term: (level+1)*N eval-domain words (read only).  b_out/a_out are
 * combined with mode 0 = Add, 1 = Assign (CKKSHybridKeySwitchCombine).
 * key_term may alias a_out (rotation).  fold_hybrid_keyswitch=true keys. */
int mi_keyswitch(mi_context* ctx, int level, const uint64_t* key_term, const uint64_t* key,
                 uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, int montgomery,
                 void* stream);
/* mi_keyswitch with no allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The switch's scratch is
 * the library's fixed device array of 2^22 words (32 MiB) per GPU, not the
 * stream workspace; a switch needing more (ks words = (L + parts (L + K) +
 * 2K + 2L) N, e.g. level 29 of a 30 + 10-limb N = 2^16 chain) is refused
 * with MI_EINVAL and nothing runs.  Arguments and results as mi_keyswitch.
 * All calls on one GPU share the array: calls that overlap on one GPU
 * (several streams or host threads) race. */
int mi_keyswitch_noalloc(mi_context* ctx, int level, const uint64_t* key_term, const uint64_t* key,
                         uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, int montgomery,
                         void* stream);

/* One key switch of a ragged batch (mixed levels, keys and combine modes):
 * the reference's OpenMP loops over independent switches
 * (ckks_aes_ctr_walsh_sbox.cpp:238, :302).  Items of one (level, modes) are
 * switched together, up to 32 per launch sequence (each phase of the switch
 * is one launch for all of them); an item whose buffers overlap an earlier
 * item's outputs (or whose outputs overlap its inputs) runs after it, so the
 * result equals switching the items one by one in order. */
typedef struct {
    const uint64_t* key_term;
    const uint64_t* key;
    uint64_t* b_out;
    uint64_t* a_out;
    int32_t level;
    int32_t b_mode;
    int32_t a_mode;
    int32_t reserved;
} mi_keyswitch_item;
int mi_keyswitch_batch(mi_context* ctx, const mi_keyswitch_item* items, int count, int montgomery,
                       void* stream);

/* ckks_relin_hybrid_inplace (ckks_relin.cpp:8-44): (b, a) += KS_rlk(c2) */
int mi_relin(mi_context* ctx, int level, uint64_t* b, uint64_t* a, const uint64_t* c2,
             const uint64_t* key, int montgomery, void* stream);
/* ckks_rotate_inplace (ckks_rotation.cpp:41-96): switch a, then permute b and a */
int mi_rotate(mi_context* ctx, int level, uint64_t* b, uint64_t* a, uint64_t galois,
              const uint64_t* key, int montgomery, void* stream);
/* ckks_conjugate_inplace (ckks_conjugate.cpp:16-61): permute, then switch */
int mi_conjugate(mi_context* ctx, int level, uint64_t* b, uint64_t* a, const uint64_t* key,
                 int montgomery, void* stream);

/* ---- hoisting and QP-basis key switching ----------------------------------
 * Digits: uint64[parts][level+1+np][N] evaluation domain, canonical; QP polys
 * hold level+1 Q limbs then np P limbs.  parts = min(dnum, ceil((level+1)/alpha)).
 *
 * ckks_hoist_ciphertext_a_qp_eval (ckks_keyswitch.cpp:344-373): ModUp of every
 * digit of a (level+1 eval limbs) to the other Q limbs and P. */
int mi_hoist(mi_context* ctx, int level, const uint64_t* a, uint64_t* digits, void* stream);
/* mi_hoist of `count` ciphertexts at one level (the per-member hoists of a
 * batch, e.g. the mask-packed unpack's unpack_segment_from_hoist,
 * ckks_walsh_sbox.cpp:498-509, over every member): the digits of a[i] at
 * digits + i * digits_stride words (stride at least parts*(level+1+np)*N;
 * the digit region must not overlap any input); the ModUps of up to 32
 * ciphertexts run as one launch sequence. */
int mi_hoist_batch(mi_context* ctx, int level, int count, const uint64_t* const* a, uint64_t* digits,
                   int64_t digits_stride, void* stream);
/* ckks_conjugate_inplace (ckks_conjugate.cpp:16-61) of `count` ciphertexts at
 * one level: each automorphism as in mi_conjugate, then the key switches of
 * all of them as one batch (up to 32 per launch sequence); each result
 * equals mi_conjugate's.  The ciphertexts must not share polys. */
int mi_conjugate_batch(mi_context* ctx, int level, int count, uint64_t* const* b, uint64_t* const* a,
                       const uint64_t* key, int montgomery, void* stream);
/* ckks_hybrid_key_switch_hoisted_inplace (ckks_keyswitch.cpp:1072-1328): inner
 * product with the key, folded ModDown, combine (mode 0 = Add, 1 = Assign),
 * then the automorphism g of both outputs (galois 0 = none). */
int mi_keyswitch_hoisted(mi_context* ctx, int level, const uint64_t* digits, const uint64_t* key,
                         uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, uint64_t galois,
                         int montgomery, void* stream);
/* count rotations of one hoisted ciphertext in one launch sequence (the
 * Walsh S-box unpacks one ciphertext through dozens of hoisted rotations,
 * ckks_walsh_sbox.cpp:521-546): for each i, ckks_hybrid_key_switch_hoisted_inplace
 * with b Add and a Assign and the map of galois[i], i.e.
 * (b_out[i], a_out[i]) = g_i(b + KS_i(a)) with digits = mi_hoist of a.
 * keys[i], b_out[i], a_out[i]: device pointers (host arrays); outputs must
 * be pairwise disjoint and overlap neither digits nor b. */
int mi_rotate_hoisted_batch(mi_context* ctx, int level, const uint64_t* digits, const uint64_t* b, int count,
                            const uint64_t* galois, const uint64_t* const* keys, uint64_t* const* b_out,
                            uint64_t* const* a_out, int montgomery, void* stream);
/* mi_rotate_hoisted_batch with its own hoist per rotation: rotation i from
 * the digits at digits + (i mod digits_period) * digits_stride words
 * (digits_period 0: the count; stride 0: one shared hoist; else at least
 * parts*(level+1+np)*N, e.g. mi_hoist_batch's output) with b[i] added,
 * (b_out[i], a_out[i]) = g_i(b[i] + KS_i(a_i)).  Ordered rotation-major over
 * the members of a batch (the mask-packed unpack's unpack_segment_from_hoist,
 * ckks_walsh_sbox.cpp:498-509, for every member and segment: period = the
 * members), a launch's rotations share one key; launches of independent
 * chunks run on the library's side streams too.  Outputs must be pairwise
 * disjoint and overlap neither the digits nor any b. */
int mi_rotate_hoisted_items(mi_context* ctx, int level, int count, const uint64_t* digits, int64_t digits_stride,
                            int digits_period, const uint64_t* const* b, const uint64_t* galois,
                            const uint64_t* const* keys, uint64_t* const* b_out, uint64_t* const* a_out,
                            int montgomery, void* stream);
/* ckks_rotate_inplace (ckks_rotation.cpp:41-96) for `count` independent
 * ciphertexts at one level, as the mask-packed parity input rotates its 29
 * parity sums (add_rotated_segment_term_inplace, ckks_walsh_sbox.cpp:
 * 484-546): each (b[i], a[i]) <- g_i(b[i] + KS_i(a[i]), KS_i(a[i])) with
 * keys[i], in place.  The key switches of up to 32 ciphertexts go through
 * each phase as one launch.  Ciphertexts must not overlap. */
int mi_rotate_batch(mi_context* ctx, int level, int count, uint64_t* const* b, uint64_t* const* a,
                    const uint64_t* galois, const uint64_t* const* keys, int montgomery, void* stream);
/* ckks_hybrid_key_switch_hoisted_qp_rotated (ckks_keyswitch.cpp:675-852):
 * b_qp = g([b ; 0] + sum_d digit_d key_b,d), a_qp = g(sum_d digit_d key_a,d),
 * no ModDown.  b: level+1 eval limbs; b_qp, a_qp: level+1+np limbs. */
int mi_keyswitch_hoisted_qp(mi_context* ctx, int level, const uint64_t* digits, const uint64_t* key,
                            const uint64_t* b, uint64_t* b_qp, uint64_t* a_qp, uint64_t galois,
                            int montgomery, void* stream);
/* ckks_hybrid_key_switch_qp_from_key_term (ckks_keyswitch.cpp:854-1070):
 * ModUp of key_term then the inner product, result kept in the QP basis. */
int mi_keyswitch_qp(mi_context* ctx, int level, const uint64_t* key_term, const uint64_t* key,
                    uint64_t* b_qp, uint64_t* a_qp, int montgomery, void* stream);
/* moddown_qp_poly_to_q_eval (ckks_linear_transform.cpp:581-605) with the
 * folded ModDown (ckks_keyswitch.cpp:609-673): in place; the first level+1
 * limbs of qp hold the Q result, the P limbs are overwritten. */
int mi_moddown_qp(mi_context* ctx, int level, uint64_t* qp, void* stream);

/* ---- slot encoding (encode_slots_custom_basis, ckks_encoding.cpp:368-410;
 * fft_dif, fft.cpp:75-95) ----------------------------------------------------
 * mi_set_encoding_tables: the context's FFT twiddles (N complex doubles as
 * re, im pairs: the conjugates of twiddle_ifft_generation, fft.cpp:37-73,
 * ckks_context.cpp:610-614) and its slot
 * permutation (N/2 entries; perm_table_generation, ckks_context.cpp:56-68).
 * mi_encode_slots: count vectors of N/2 complex doubles (host, re/im pairs)
 * -> count plaintexts of `limbs` limbs (q_0..), poly b at out + b*poly_stride;
 * eval != 0 also applies the forward NTT.  The FFT runs in double with the
 * reference's operation order, so the rounded coefficients equal its own. */
int mi_set_encoding_tables(mi_context* ctx, const double* twiddles, const uint64_t* perm);
int mi_encode_slots(mi_context* ctx, const double* values, int count, double scale, int limbs, int eval,
                    uint64_t* out, int64_t poly_stride, void* stream);
/* ---- decoding (ckks_encoding.cpp:254-329, fft.cpp:37-117) ----------------
 * mi_set_decoding_tables: the context's inverse special FFT twiddles (N
 * complex as 2N doubles) and slot permutation (N/2 entries below N).
 * mi_decode_batch: `count` coefficient-domain plaintexts (canonical residues,
 * `limbs` x N words each, device) -> count x N/2 complex slots (device, 2
 * doubles each; coeff_domain = 1: the coefficient-domain read-out (m[brev i],
 * m[N/2 + brev i]) instead of the FFT).  The scale of each is given exactly
 * as an x87 extended value: significand scale_mant (top bit set) and
 * exponent scale_exp (scale = scale_mant 2^scale_exp).  The CRT uses the
 * longest prefix of the chain whose product fits in 128 bits; the results
 * equal the host decode bit for bit. */
int mi_set_decoding_tables(mi_context* ctx, const double* twiddles, const uint64_t* perm);
int mi_decode_batch(mi_context* ctx, int count, const uint64_t* const* coeffs, int limbs,
                    const uint64_t* scale_mant, const int32_t* scale_exp, int coeff_domain, double* out,
                    void* stream);

/* ---- sampling (ckks_uniform_mask_polys / ckks_gaussian_noise_polys,
 * ckks_encoding.hpp:11-15; the reference draws from std::random_device-seeded
 * mt19937_64, ckks_encoding.cpp:211-223) -------------------------------------
 * Counter-based (Philox4x32-10): the output is a function of (seed, counter)
 * only.  mi_sample_uniform: limbs x N residues uniform in [0, q_b) for basis
 * indices b = basis0.. .  mi_sample_gaussian: one discrete Gaussian integer
 * per coefficient (cumulative table over |x| <= ceil(12.0061 sigma)), written
 * as residues of limbs q_0..q_{limbs-1} (coefficient domain).  Not a
 * cryptographic generator, like the reference's. */
int mi_sample_uniform(mi_context* ctx, uint64_t* out, int limbs, int basis0, uint64_t seed, uint64_t counter,
                      void* stream);
int mi_sample_gaussian(mi_context* ctx, uint64_t* out, int limbs, double sigma, uint64_t seed, uint64_t counter,
                       void* stream);

/* ---- BSGS linear-transform stage, QP evaluator ---------------------------
 * ckks_apply_linear_transform_owned_stage_qp (ckks_linear_transform.cpp:
 * 1174-1407) without its final rescale: a hoisted baby step per needed
 * rotation r < baby_steps, the diagonal inner products in the QP basis, one
 * QP key switch and automorphism per giant step, one ModDown pair at the end.
 * diag_indices[t] >= 0 is term t's diagonal; diags[t] points to its
 * level+1+np limb plaintext (Montgomery form if pt_montgomery); keys[i] is the
 * rotation key for key_rotations[i] (host arrays of device pointers).
 * a_out / b_out: level+1 limbs each, may alias a / b. */
int mi_linear_transform_qp(mi_context* ctx, int level, const uint64_t* a, const uint64_t* b,
                           int nterms, const int64_t* diag_indices, const uint64_t* const* diags,
                           int pt_montgomery, int baby_steps, int slots, int nkeys,
                           const int64_t* key_rotations, const uint64_t* const* keys,
                           int montgomery, uint64_t* a_out, uint64_t* b_out, void* stream);
/* The same stage applied to count ciphertexts at one level: the per-batch
 * OpenMP loops of the reference (ckks_aes_ctr_walsh_sbox.cpp:437-470 over
 * refresh batches) run independent ciphertexts through the same stages.
 * One launch sequence serves up to 32 ciphertexts: each diagonal and key is
 * read once per launch, not once per ciphertext.  a_out[i] / b_out[i] may
 * alias a[i] / b[i], not another ciphertext's buffers. */
int mi_linear_transform_qp_batch(mi_context* ctx, int level, int count, const uint64_t* const* a,
                                 const uint64_t* const* b, int nterms, const int64_t* diag_indices,
                                 const uint64_t* const* diags, int pt_montgomery, int baby_steps, int slots,
                                 int nkeys, const int64_t* key_rotations, const uint64_t* const* keys,
                                 int montgomery, uint64_t* const* a_out, uint64_t* const* b_out,
                                 void* stream);

/* ---- BSGS linear-transform stage, Q evaluator ----------------------------
 * ckks_apply_linear_transform_bsgs (ckks_linear_transform.cpp:818-996) and
 * ckks_apply_linear_transform_owned_stage_q (:998-1172) without the final
 * rescale: every needed baby rotation r < baby_steps is a full hoisted
 * rotation in Q, diags[t] are level+1 limb plaintexts (Montgomery form if
 * pt_montgomery), each giant step's inner sum is rotated by a key switch and
 * its automorphism, and the results are added.  Indices must be distinct and
 * below slots.  Arguments as for mi_linear_transform_qp. */
int mi_linear_transform_q(mi_context* ctx, int level, const uint64_t* a, const uint64_t* b, int nterms,
                          const int64_t* diag_indices, const uint64_t* const* diags, int pt_montgomery,
                          int baby_steps, int slots, int nkeys, const int64_t* key_rotations,
                          const uint64_t* const* keys, int montgomery, uint64_t* a_out, uint64_t* b_out,
                          void* stream);
/* The Q-evaluator stage over count ciphertexts at one level (as
 * mi_linear_transform_qp_batch). */
int mi_linear_transform_q_batch(mi_context* ctx, int level, int count, const uint64_t* const* a,
                                const uint64_t* const* b, int nterms, const int64_t* diag_indices,
                                const uint64_t* const* diags, int pt_montgomery, int baby_steps, int slots,
                                int nkeys, const int64_t* key_rotations, const uint64_t* const* keys,
                                int montgomery, uint64_t* const* a_out, uint64_t* const* b_out, void* stream);

/* ---- bootstrapping pieces --------------------------------------------------
 * modraise_eval_poly_from_q0 (ckks_modraise.cpp:31-75): one q_0 eval limb,
 * centred lift, reduced to q_0..q_target and returned to the evaluation
 * domain (target_level+1 limbs; out may alias poly_q0). */
int mi_modraise(mi_context* ctx, int target_level, const uint64_t* poly_q0, uint64_t* out,
                void* stream);
/* mi_modraise of `count` polys (ckks_modraise_inplace over the polys and
 * members of a ciphertext batch, ckks_modraise.cpp:76-110): out[i] from
 * poly_q0[i]; up to 32 polys per launch sequence; out[i] may alias
 * poly_q0[i], outputs must not overlap each other. */
int mi_modraise_batch(mi_context* ctx, int target_level, int count, const uint64_t* const* poly_q0,
                      uint64_t* const* out, void* stream);
/* ckks_q0p0_key_switch_level0_inplace (ckks_keyswitch.cpp:1661-1816): the
 * level-0 dense-to-sparse switch over {q_0, p_0}; key = [b(q0), b(p0),
 * a(q0), a(p0)], N words each.  (b, a) <- (b + KS_b(a), KS_a(a)). */
int mi_q0p0_keyswitch(mi_context* ctx, uint64_t* b, uint64_t* a, const uint64_t* key,
                      int montgomery, void* stream);
/* mi_q0p0_keyswitch of `count` level-0 ciphertexts with one key (the
 * members of a batch, ckks_key_switch_dense_to_sparse_inplace over each):
 * every phase of up to 32 switches as one launch; ciphertexts must not
 * overlap. */
int mi_q0p0_keyswitch_batch(mi_context* ctx, int count, uint64_t* const* b, uint64_t* const* a,
                            const uint64_t* key, int montgomery, void* stream);

/* ckks_rescale (ckks_rescale.cpp:68-148) of `npolys` listed polys at one
 * level (the polys of one or more ciphertexts), in place; the top limb of
 * every poly goes through one inverse-NTT launch pair and the remainder /
 * finish through one forward pair per 64 polys. */
int mi_rescale_list(mi_context* ctx, int level, int npolys, uint64_t* const* polys, void* stream);
/* The AES-CTR masking loops (mask_inactive_slots_inplace /
 * mask_final_refresh_slots_inplace, ckks_aes_ctr_walsh_sbox.cpp:217-329):
 * every listed poly times one plaintext (ckks_mul_inplace(ct, pt)), then,
 * if rescale, ckks_rescale of all of them; one launch (pair) per 64 polys
 * instead of a loop over ciphertexts. */
int mi_mul_plain_rescale_list(mi_context* ctx, int level, int npolys, uint64_t* const* polys, const uint64_t* pt,
                              int pt_montgomery, int rescale, void* stream);
/* Listed poly pairs, one launch per 64 pairs, over limbs 0..level:
 * x[k] = y[k] (MI_LIST_COPY), x[k] + y[k] (MI_LIST_ADD), x[k] - y[k]
 * (MI_LIST_SUB), x[k] * r (MI_LIST_MUL_SCALAR, y unused) or x[k] + y[k] * r
 * (MI_LIST_MUL_SCALAR_ADD) mod q, r = residues[limb] (host array, each below
 * its modulus).  The per-poly loops of a ciphertext copy
 * (CKKSCiphertext::clone, ckks_ciphertext.hpp:18-108), ckks_add_inplace /
 * ckks_sub_inplace (ckks_addition.cpp:35-139), ckks_mul_inplace(ct, int)
 * (ckks_mul.cpp:37-272) and fused_multiply_by_real_scalar_then_add_inplace
 * (ckks_evalpoly.cpp:352-398) as one launch.  A copy's source and
 * destination must not overlap. */
/* out[k] = sum_t coef[k * nterms + t] * in[k * nterms + t] mod q over limbs
 * 0..level (coef 0: term unused; 1 <= nterms <= 16; coefficients below
 * every modulus; no output may alias an input), 12 outputs per launch: the
 * AES round's lazy-XOR layer (MixColumns and AddRoundKey as sums of S-box
 * output and round-key bit ciphertexts with the reference's integer
 * scale-ups, ckks_aes_ctr_walsh_sbox.cpp:575-641, 741-817) in a few
 * launches.  Every sum is exact mod q, so the result equals the
 * reference's sequence of additions. */
int mi_linear_combine(mi_context* ctx, int level, int nout, uint64_t* const* out, int nterms,
                      const uint64_t* const* in, const uint32_t* coef, void* stream);
/* out[k] = constant[limb] (k < nconst) + sum_t residues[t * (level+1) + limb]
 * * in[t * npolys + k] mod q over limbs 0..level, for the npolys listed
 * output polys (in[t * npolys + k] null: term t has no poly k; 1 <= nterms
 * <= 64; residues and constant below their moduli; outputs overlap neither
 * each other nor any input).  One Chebyshev leaf of the Paterson-Stockmeyer
 * evaluation — the reference's zero ciphertext, its constant and its loop of
 * MulThenAdd over the baby-step powers (evaluate_leaf / MulThenAdd,
 * ckks_evalpoly.cpp:352-398 and the leaf loop that calls it), with every
 * integer scale-up folded into the residues on the host — in one pass per
 * 16 polys instead of a zero fill, a constant add and a read-modify-write
 * per term.  Also the mask-packed parity input's 30 parity sums (masked
 * input bits, coefficients 1) and its packing sum of the 30 rotated terms
 * (ckks_walsh_sbox.cpp:363-383, 497-519).  The sum is exact mod q, so the
 * result equals the reference's sequence of operations. */
int mi_scaled_sum(mi_context* ctx, int level, int npolys, uint64_t* const* out, int nterms,
                  const uint64_t* const* in, const uint64_t* residues, int nconst, const uint64_t* constant,
                  void* stream);
/* Also MI_LIST_ADD_SCALAR: x[k] + r (ckks_add_integer_constant_inplace,
 * ckks_addition.cpp:89-112, on the b polys of a batch; y unused),
 * MI_LIST_ZERO: x[k] = 0 (y and residues unused), MI_LIST_ADD_PLAIN_MONT:
 * x[k] + REDC(y[k]) (ckks_add_inplace(ct, pt) of a Montgomery-form plaintext,
 * ckks_addition.cpp:35-60). */
enum {
    MI_LIST_COPY = 0,
    MI_LIST_ADD = 1,
    MI_LIST_SUB = 2,
    MI_LIST_MUL_SCALAR = 3,
    MI_LIST_MUL_SCALAR_ADD = 4,
    MI_LIST_ADD_SCALAR = 5,
    MI_LIST_ZERO = 6,
    MI_LIST_ADD_PLAIN_MONT = 7
};
int mi_poly_list_op(mi_context* ctx, int op, int level, int npolys, uint64_t* const* x, const uint64_t* const* y,
                    const uint64_t* residues, void* stream);

/* ---- Walsh S-box ----
 * The linear combinations and raw 4+4 products of one output pair of the
 * Walsh S-box (walsh_sum_for_output_pair, ckks_walsh_sbox.cpp:984-1043, up
 * to its relinearisation), fused into one pass over the signs.  All
 * ciphertexts are at `level` (level+1 limbs, limb-major, evaluation domain).
 * low_b/low_a and high_b/high_a: 15 pointers each, the sign ciphertexts of
 * nibble masks 1..15.  cr, ci: int16 tables [16][16] indexed [high mask]
 * [low mask] (the real and imaginary output bits' Walsh coefficients); the
 * column [h][0] is unused: the constant term enters as the rounded scaled
 * integer V (ckks_add_integer_constant_inplace) given as |V| in two words
 * (lo, hi) in const_abs[4h + 2r .. 4h + 2r + 1] and its sign in
 * const_neg[2h + r] (r = 0 real, 1 imaginary).  pt: the encoded i (Montgomery
 * form if pt_montgomery).  With L(c) = sum_m c[m] low_m + V on b and
 * Z_h = L(cr[h]) + pt * L(ci[h]):  acc = Z_0 (b, a) and
 * raw = sum_{h=1..15} Z_h (x) high_h (b, a, c2).  Outputs must not alias
 * the inputs. */
int mi_walsh_product_sum(mi_context* ctx, int level, const uint64_t* const* low_b, const uint64_t* const* low_a,
                         const uint64_t* const* high_b, const uint64_t* const* high_a, const int16_t* cr,
                         const int16_t* ci, const uint64_t* const_abs, const int32_t* const_neg, const uint64_t* pt,
                         int pt_montgomery, uint64_t* acc_b, uint64_t* acc_a, uint64_t* raw_b, uint64_t* raw_a,
                         uint64_t* raw_c2, void* stream);
/* mi_walsh_product_sum for `count` (member, output pair) items of a batch:
 * item i's 15 + 15 sign pointers at low_b[15 i ...] etc., its tables cr[i],
 * ci[i], const_abs[i], const_neg[i], its outputs acc_b[i] ...; one encoded
 * i plaintext for all.  The four output pairs of every member of an S-box
 * batch (the loop over output pairs, ckks_walsh_sbox.cpp:1137-1152)
 * in one call, the launches spread over the library's streams.  Outputs
 * pairwise disjoint and overlapping no input. */
int mi_walsh_product_sum_batch(mi_context* ctx, int level, int count, const uint64_t* const* low_b,
                               const uint64_t* const* low_a, const uint64_t* const* high_b,
                               const uint64_t* const* high_a, const int16_t* const* cr, const int16_t* const* ci,
                               const uint64_t* const* const_abs, const int32_t* const* const_neg, const uint64_t* pt,
                               int pt_montgomery, uint64_t* const* acc_b, uint64_t* const* acc_a,
                               uint64_t* const* raw_b, uint64_t* const* raw_a, uint64_t* const* raw_c2, void* stream);

/* ---- HIP graph capture ----
 * No counterpart in the reference, whose evaluator runs eagerly on host
 * threads; these let a caller record a fixed sequence of the calls above (an
 * S-box, a bootstrap stage, a batch of key switches) once and replay it as
 * one HIP graph.  mi_graph_begin starts capturing `stream` (a created
 * stream, not the null stream); every call made on it until mi_graph_end is
 * recorded, not run.  Calls that synchronise the host (encode, sampling
 * table changes, host copies) cannot be captured, and a call whose stream
 * workspace would have to grow fails: run the sequence once eagerly first.
 * Replays read and write the buffers that were passed during capture.
 * Device blocks that the context allocates or frees during a capture are
 * kept from reuse until every graph of the context is destroyed.  A replay
 * shares the context's workspaces with eager calls: do not replay on one
 * stream while calls run on another. */
typedef struct mi_graph mi_graph;
int mi_graph_begin(mi_context* ctx, void* stream);
int mi_graph_end(mi_context* ctx, void* stream, mi_graph** out);
int mi_graph_launch(mi_context* ctx, mi_graph* graph, void* stream);
int mi_graph_destroy(mi_context* ctx, mi_graph* graph);

#ifdef __cplusplus
}
#endif
#endif
