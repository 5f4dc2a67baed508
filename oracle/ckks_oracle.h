/*
 * CPU oracle for the CKKS evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (TrustworthyComputing/aes-walsh-ckks); every function cites the
 * reference file:line it follows.  It exists to check the HIP library's
 * results.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker -- the product path never links,
 * loads or falls back to it.
 *
 * Conventions (the reference's): an RNS polynomial is a flat uint64 array of
 * limbs*N words, limb-major (limb*N + coeff); evaluation-domain index i holds
 * A(psi^(2*brev(i)+1)); stored values are canonical residues in [0, q).
 * Floating-point HPS estimates are evaluated without FMA contraction
 * (the build uses -ffp-contract=off).
 */
#ifndef CKKS_ORACLE_H
#define CKKS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- modular arithmetic (src/math/modarith) ---- */
uint64_t oc_powmod(uint64_t a, uint64_t e, uint64_t m);
uint64_t oc_mod_inverse(uint64_t a, uint64_t m);
uint64_t oc_compute_shoup(uint64_t w, uint64_t q);
void oc_barrett_const(uint64_t q, uint64_t out_lo_hi[2]);
uint64_t oc_barrett_mul(uint64_t a, uint64_t b, uint64_t q);
uint64_t oc_barrett_reduce_u64(uint64_t a, uint64_t q);
uint64_t oc_shoup_mul(uint64_t a, uint64_t w, uint64_t q);
uint64_t oc_montgomery_neg_inverse(uint64_t q);
uint64_t oc_montgomery_radix_mod_q(uint64_t q);
uint64_t oc_montgomery_mul(uint64_t normal, uint64_t mont, uint64_t q);

/* ---- prime search (src/util/random.cpp) ---- */
int oc_is_prime(uint64_t n);
int oc_find_p_and_q_primes(const uint64_t* p_bits, int np, const uint64_t* q_bits,
                           int nq, uint64_t N, uint64_t* p_out, uint64_t* q_out);

/* ---- twiddles and NTT (src/math/ntt/ntt.cpp) ---- */
uint64_t oc_find_root_of_unity(uint64_t N, uint64_t q);
/* fwd, fwd_shoup, inv: N words; inv_shoup: N+1 words (last = Shoup(N^-1)). */
int oc_ntt_tables(uint64_t N, uint64_t q, uint64_t* fwd, uint64_t* fwd_shoup,
                  uint64_t* inv, uint64_t* inv_shoup, uint64_t* n_inv);
void oc_ntt_forward(uint64_t* a, uint64_t N, uint64_t q, const uint64_t* fwd,
                    const uint64_t* fwd_shoup, int lazy);
void oc_ntt_inverse(uint64_t* a, uint64_t N, uint64_t q, const uint64_t* inv,
                    const uint64_t* inv_shoup, uint64_t n_inv, int lazy);

/* ---- context: per-modulus tables and key-switch precomputes ---- */
typedef struct oc_ctx oc_ctx;
/* q: nq chain moduli (limb 0 .. Lmax), p: np special moduli, dnum digits. */
oc_ctx* oc_ctx_create(uint64_t N, const uint64_t* q, int nq, const uint64_t* p, int np, int dnum);
void oc_ctx_destroy(oc_ctx* ctx);
int oc_ctx_alpha(const oc_ctx* ctx);
int oc_omp_threads(void);

/* NTT over the first `limbs` chain moduli (flat RNS poly). */
void oc_ntt_fwd_q(const oc_ctx* ctx, uint64_t* data, int limbs, int lazy);
void oc_ntt_inv_q(const oc_ctx* ctx, uint64_t* data, int limbs, int lazy);
/* NTT over an arbitrary list of basis indices: idx < nq is q[idx], idx >= nq is p[idx-nq]. */
void oc_ntt_fwd_basis(const oc_ctx* ctx, uint64_t* data, const int* basis, int limbs, int lazy);
void oc_ntt_inv_basis(const oc_ctx* ctx, uint64_t* data, const int* basis, int limbs, int lazy);

/* HPS ModUp of one digit `part` at `level`: src = digit limbs (coefficient
 * domain, alpha_d*N words), dst = (L - alpha_d + K)*N words in the
 * reference's destination order (Q limbs below the digit, Q limbs above it, P). */
int oc_modup(const oc_ctx* ctx, int level, int part, const uint64_t* src, uint64_t* dst);

/* Hybrid key switch (ckks_keyswitch.cpp:1330-1641), fold_hybrid_keyswitch=true.
 * key_term: (level+1)*N eval-domain words (overwritten, like the reference).
 * keys: dnum_active digits laid out [digit][poly b=0,a=1][Lmax+1+K limbs][N].
 * mode 0 = Add, 1 = Assign. montgomery selects Montgomery-form keys. */
int oc_keyswitch(const oc_ctx* ctx, int level, uint64_t* key_term, const uint64_t* keys,
                 uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, int montgomery);

/* Galois automorphisms (rotation_key_generation.cpp:73-90, conjugation :59-61). */
uint64_t oc_rotation_galois_element(uint64_t N, int64_t rotation);
void oc_ntt_automorphism_map(uint64_t N, uint64_t galois, uint64_t* dst_to_src);
void oc_apply_automorphism(uint64_t* data, uint64_t N, int limbs, const uint64_t* dst_to_src);

/* Evaluator ops on ciphertexts stored as polys [b, a, (c2)] each (level+1)*N. */
int oc_rotate(const oc_ctx* ctx, int level, uint64_t* b, uint64_t* a, int64_t rotation,
              const uint64_t* keys, int montgomery);
int oc_conjugate(const oc_ctx* ctx, int level, uint64_t* b, uint64_t* a,
                 const uint64_t* keys, int montgomery);
int oc_relin(const oc_ctx* ctx, int level, uint64_t* b, uint64_t* a, uint64_t* c2,
             const uint64_t* keys, int montgomery);
/* polys: npolys consecutive polys of stride (level+1)*N; after the call the
 * first level*N words of each poly hold the rescaled result. */
int oc_rescale(const oc_ctx* ctx, int level, uint64_t* polys, int npolys);
int oc_tensor(const oc_ctx* ctx, int level, uint64_t* b1, uint64_t* a1, const uint64_t* b2,
              const uint64_t* a2, uint64_t* c2);
int oc_square(const oc_ctx* ctx, int level, uint64_t* b, uint64_t* a, uint64_t* c2, int montgomery);
int oc_mul_plain(const oc_ctx* ctx, int level, uint64_t* polys, int npolys, const uint64_t* pt,
                 int pt_montgomery);
int oc_mul_int(const oc_ctx* ctx, int level, uint64_t* polys, int npolys, int64_t scalar,
               int montgomery);
int oc_add(const oc_ctx* ctx, int level, uint64_t* x, const uint64_t* y, int npolys);
int oc_sub(const oc_ctx* ctx, int level, uint64_t* x, const uint64_t* y, int npolys);
int oc_add_plain(const oc_ctx* ctx, int level, uint64_t* b, const uint64_t* pt, int pt_montgomery);
int oc_add_scalar_residues(const oc_ctx* ctx, int level, uint64_t* b, const uint64_t* residues);

/* Hoisting and QP-basis key switching (ckks_keyswitch.cpp:95-227, 344-373,
 * 675-1328).  digits: [parts][L+K][N] evaluation domain, canonical. */
int oc_hoist(const oc_ctx* ctx, int level, const uint64_t* a_eval, uint64_t* digits);
int oc_keyswitch_hoisted(const oc_ctx* ctx, int level, const uint64_t* digits, const uint64_t* keys,
                         uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, uint64_t galois,
                         int montgomery);
int oc_keyswitch_hoisted_qp_rotated(const oc_ctx* ctx, int level, const uint64_t* digits,
                                    const uint64_t* keys, const uint64_t* b_q_eval, uint64_t* b_qp,
                                    uint64_t* a_qp, uint64_t galois, int montgomery);
int oc_keyswitch_qp_from_key_term(const oc_ctx* ctx, int level, const uint64_t* key_term_eval,
                                  const uint64_t* keys, uint64_t* b_qp, uint64_t* a_qp, int montgomery);
int oc_moddown_qp_to_q(const oc_ctx* ctx, int level, uint64_t* qp);
/* BSGS linear-transform stage on the QP evaluator (ckks_linear_transform.cpp:1174-1407),
 * without the final rescale. */
int oc_lt_stage_qp(const oc_ctx* ctx, int level, const uint64_t* a, const uint64_t* b, int nterms,
                   const int64_t* diag_idx, const uint64_t* diags, int pt_mont, int baby_steps,
                   int slots, const int64_t* rot_amounts, const uint64_t* const* rot_keys, int nkeys,
                   int montgomery, uint64_t* a_out, uint64_t* b_out);
/* ModRaise of one q0 limb to target_level+1 limbs (ckks_modraise.cpp:31-75). */
int oc_modraise(const oc_ctx* ctx, int target_level, const uint64_t* poly_q0_eval, uint64_t* out);
/* Dense-to-sparse switch at level 0 over {q0, p0} (ckks_keyswitch.cpp:1661-1816). */
int oc_q0p0_keyswitch(const oc_ctx* ctx, uint64_t* b, uint64_t* a, const uint64_t* key, int montgomery);

#ifdef __cplusplus
}
#endif
#endif
