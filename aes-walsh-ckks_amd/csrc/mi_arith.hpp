// Device modular arithmetic for 64-bit RNS limbs (q < 2^62) on CDNA4 VALU.
//
// Every helper is an exact residue computation: for canonical inputs the
// canonical output equals the reference's (src/math/modarith/modarith.hpp),
// whatever internal representative a lazy step carries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

typedef uint64_t u64;
typedef uint32_t u32;

__device__ __forceinline__ u64 mulhi(u64 a, u64 b) { return __umul64hi(a, b); }

// 128-bit accumulator (lo, hi) for lazily summed products.
struct U128 {
    u64 lo, hi;
};

// Full 64x64 -> 128-bit product from four 32x32 -> 64 multiply-adds (each
// partial sum fits in 64 bits: (2^32-1)^2 + 2 (2^32-1) = 2^64 - 1).
__device__ __forceinline__ U128 mul128(u64 a, u64 b) {
    const u64 a0 = (u32)a, a1 = a >> 32, b0 = (u32)b, b1 = b >> 32;
    const u64 p00 = a0 * b0;
    const u64 t = a0 * b1 + (p00 >> 32);
    const u64 u = a1 * b0 + (u32)t;
    return U128{(u << 32) | (u32)p00, a1 * b1 + (t >> 32) + (u >> 32)};
}

__device__ __forceinline__ void mac128(U128& acc, u64 a, u64 b) {
    const U128 p = mul128(a, b);
    const u64 s = acc.lo + p.lo;
    acc.hi += p.hi + (s < p.lo ? 1 : 0);
    acc.lo = s;
}

// Shoup product a*w mod q for any a < 2^64, w < q, ws = floor(w*2^64/q):
// result in [0, 2q) (lazy) or [0, q).
__device__ __forceinline__ u64 shoup_lazy(u64 a, u64 w, u64 ws, u64 q) {
    return a * w - __umul64hi(a, ws) * q;
}
// The same lazy product with negq = 2^64 - q: a*w + qhat*negq is a*w - qhat*q
// modulo 2^64, so the subtraction rides in the low word's multiply-add.
// negq must come from memory (a table field): from 0 - q the compiler
// rebuilds the subtraction.
__device__ __forceinline__ u64 shoup_lazy_neg(u64 a, u64 w, u64 ws, u64 negq) {
    return a * w + __umul64hi(a, ws) * negq;
}
__device__ __forceinline__ u64 shoup(u64 a, u64 w, u64 ws, u64 q) {
    u64 r = shoup_lazy(a, w, ws, q);
    return r >= q ? r - q : r;
}

__device__ __forceinline__ u64 add_mod(u64 a, u64 b, u64 q) {
    u64 r = a + b;
    return r >= q ? r - q : r;
}
__device__ __forceinline__ u64 sub_mod(u64 a, u64 b, u64 q) {
    return a >= b ? a - b : a + q - b;
}
__device__ __forceinline__ u64 reduce_4q(u64 x, u64 q) {
    u64 q2 = q << 1;
    if (x >= q2) x -= q2;
    return x >= q ? x - q : x;
}

// Single-word reduction x mod q (x < 2^64) with mu_hi = floor(2^128/q) >> 64.
__device__ __forceinline__ u64 reduce64(u64 x, u64 q, u64 mu_hi) {
    u64 r = x - __umul64hi(x, mu_hi) * q;
    return r >= q ? r - q : r;
}

// Barrett reduction of a 128-bit value z < q * 2^64 with mu = floor(2^128/q)
// split as (mu_lo, mu_hi).  Exact: returns z mod q.
__device__ __forceinline__ u64 barrett128(u64 zl, u64 zh, u64 q, u64 mu_lo, u64 mu_hi) {
    // qhat = floor(z * mu / 2^128) (up to -2), computed from the top words.
    u64 t_hi = __umul64hi(zl, mu_hi);
    u64 a_lo = zh * mu_lo;
    u64 a_hi = __umul64hi(zh, mu_lo);
    u64 b = __umul64hi(zl, mu_lo);
    // middle word: a_lo + t_lo(zl*mu_hi) + b  -> carries into the high word
    u64 m_lo = zl * mu_hi;
    u64 s1 = m_lo + a_lo;
    u64 c1 = s1 < m_lo ? 1 : 0;
    u64 s2 = s1 + b;
    u64 c2 = s2 < s1 ? 1 : 0;
    u64 qhat = zh * mu_hi + t_hi + a_hi + c1 + c2;
    // qhat >= floor(z/q) - 1 because z*(2^128/q - mu)/2^128 < q/2^64 < 1/4,
    // so r < 2q.
    u64 r = zl - qhat * q;
    return r >= q ? r - q : r;
}

// a*b mod q via Barrett (a, b < 2^64 with a*b < q*2^64).
__device__ __forceinline__ u64 mul_mod(u64 a, u64 b, u64 q, u64 mu_lo, u64 mu_hi) {
    return barrett128(a * b, __umul64hi(a, b), q, mu_lo, mu_hi);
}

// Any 128-bit z: z mod q (high word reduced first so barrett128 applies).
__device__ __forceinline__ u64 reduce128(u64 zl, u64 zh, u64 q, u64 mu_lo, u64 mu_hi) {
    return barrett128(zl, reduce64(zh, q, mu_hi), q, mu_lo, mu_hi);
}

// Montgomery REDC of a 128-bit t < q*2^64 with qninv = -q^-1 mod 2^64:
// returns t * 2^-64 mod q in [0, q).
__device__ __forceinline__ u64 redc128(u64 tl, u64 th, u64 q, u64 qninv) {
    u64 m = tl * qninv;
    u64 mq_hi = __umul64hi(m, q);
    // tl + m*q_lo == 0 mod 2^64; carry out is 1 iff tl != 0
    u64 u = th + mq_hi + (tl != 0 ? 1 : 0);
    return u >= q ? u - q : u;
}
// REDC of any 128-bit t: the high word is brought below q first (t keeps its
// residue), so the REDC precondition t < q*2^64 holds.
__device__ __forceinline__ u64 redc128_any(u64 tl, u64 th, u64 q, u64 qninv, u64 mu_hi) {
    return redc128(tl, reduce64(th, q, mu_hi), q, qninv);
}
__device__ __forceinline__ u64 mont_mul(u64 normal, u64 mont, u64 q, u64 qninv) {
    return redc128(normal * mont, __umul64hi(normal, mont), q, qninv);
}

}  // namespace mi
