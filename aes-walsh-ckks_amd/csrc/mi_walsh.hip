// Walsh S-box product sum: the linear combinations of sign ciphertexts and
// the raw 4+4 products of one output pair, fused into one pass over the
// signs (ckks_walsh_sbox.cpp:984-1043, walsh_sum_for_output_pair, up to its
// relinearisation).
//
// For high mask h = 0..15 and low masks m = 1..15, with integer coefficient
// tables cr, ci ([h][m]; m = 0 is the integer constant V added to b):
//   L(c)  = sum_m c[m] * low_m  (+ V on b)                 (:862-905)
//   Z_h   = L(cr[h]) + pt * L(ci[h])                         (:912-958, pt = encoded i)
//   acc   = Z_0                                               (:999-1005)
//   raw   = sum_{h >= 1} Z_h (x) high_h  (b*b, b*a + a*b, a*a) (:1009-1036)
// Every step is exact arithmetic mod q, so the canonical results equal the
// reference's sequence of scaled additions, plaintext products and tensors.
#include "mi_internal.hpp"
#include "mi_arith.hpp"

namespace mi {

constexpr int WALSH_MASKS = 16;

struct WalshArgs {
    const u64* low_b[WALSH_MASKS - 1];  // masks 1..15, L limbs each
    const u64* low_a[WALSH_MASKS - 1];
    const u64* high_b[WALSH_MASKS - 1];
    const u64* high_a[WALSH_MASKS - 1];
    const u64* pt;                      // L limbs, evaluation domain
    u64 *acc_b, *acc_a, *raw_b, *raw_a, *raw_c2;
    int cr[WALSH_MASKS][WALSH_MASKS];  // [h][m] (32-bit: scalar loads)
    int ci[WALSH_MASKS][WALSH_MASKS];
    u64 v_lo[2 * WALSH_MASKS], v_hi[2 * WALSH_MASKS];  // |V| of (h, real) at 2h, (h, imag) at 2h+1
    unsigned v_neg;                                     // sign bit per entry
    int logN, pt_mont;
    const BasisConst* bc;
};

// sum_m c[m] x_m mod q for |c| < 2^15, x_m = x1 2^31 + x0 given as 31-bit
// halves: two signed 32x32+64 multiply-adds per term, |partial sums| < 2^50.
// SMALL (q < 2^43, so x1 < 2^12 and |A1 2^31 + A0| < 2^62.1): the signed
// 64-bit value A1 2^31 + A0 shifted non-negative by off (a multiple of q
// near 2^63) and one 64-bit reduction.  Otherwise both halves shifted by off
// (a multiple of q above 2^50) and one 128-bit reduction of
// (A1 + off) 2^31 + (A0 + off).
template <bool SMALL>
__device__ __forceinline__ u64 walsh_lc(const int* c, const int (&x0)[WALSH_MASKS - 1],
                                        const int (&x1)[WALSH_MASKS - 1], u64 off, const BasisConst& b) {
    long long a0 = 0, a1 = 0;
#pragma unroll
    for (int m = 1; m < WALSH_MASKS; ++m) {
        const int cm = c[m];
        a0 += (long long)cm * x0[m - 1];
        a1 += (long long)cm * x1[m - 1];
    }
    if constexpr (SMALL) {
        const u64 v = (u64)(a1 * (1LL << 31) + a0) + off;
        return reduce64(v, b.q, b.mu_hi);
    } else {
        const u64 u = (u64)(a1 + (long long)off), w = (u64)(a0 + (long long)off);
        const u64 lo = (u << 31) + w;
        const u64 hi = (u >> 33) + (lo < w ? 1 : 0);
        return reduce128(lo, hi, b.q, b.mu_lo, b.mu_hi);
    }
}

// A block holds 64 coefficients in 8 waves: side (b or a) x quarter of the
// 16 high masks (quarter 0: h = 0..3, whose h = 0 is the acc term).  A wave
// forms its side of Z_h (b: L(cr) + const + pt L(ci), a: the same without the
// constants) for its 4 masks and the products that side owns:
// b-side Z_b H_b (raw b) and Z_b H_a, a-side Z_a H_b and Z_a H_a (raw c2);
// the partial sums meet in LDS.
constexpr int WALSH_COEFFS = 64;
constexpr int WALSH_BLOCK = 8 * WALSH_COEFFS;

__global__ void __launch_bounds__(WALSH_BLOCK) k_walsh_product_sum(WalshArgs a) {
    __shared__ u64 s_const[2 * WALSH_MASKS];
    __shared__ u64 s_part[2][2][4][WALSH_COEFFS];  // [side][product][quarter][coefficient]
    const int l = (int)(((long long)blockIdx.x * WALSH_COEFFS) >> a.logN);  // uniform: a block is in one limb
    const int lane = threadIdx.x % WALSH_COEFFS;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WALSH_COEFFS);
    const int side = wave >> 2, quarter = wave & 3;
    const long long i = (long long)blockIdx.x * WALSH_COEFFS + lane;
    const BasisConst b = a.bc[l];
    if (threadIdx.x < 2 * WALSH_MASKS) {
        const int t = threadIdx.x;
        const u64 r = reduce128(a.v_lo[t], a.v_hi[t], b.q, b.mu_lo, b.mu_hi);
        s_const[t] = ((a.v_neg >> t) & 1) && r ? b.q - r : r;
    }
    // a multiple of q above every partial sum's magnitude: q ceil(2^50 / q),
    // or, for the one-word sum, q ceil(2^63 / q) (<= 2^63 + q < 2^64 - 2^62.1)
    const bool small = b.q < (1ull << 43);  // uniform: one limb per block
    const u64 off = small ? b.q * (((1ull << 63) + b.q - 1) / b.q) : b.q * (((1ull << 50) + b.q - 1) / b.q);
    int x0[WALSH_MASKS - 1], x1[WALSH_MASKS - 1];
#pragma unroll
    for (int m = 0; m < WALSH_MASKS - 1; ++m) {
        const u64 x = (side ? a.low_a[m] : a.low_b[m])[i];
        x0[m] = (int)(x & 0x7fffffffu);
        x1[m] = (int)(x >> 31);
    }
    const u64 pt = a.pt[i];
    __syncthreads();
    // at most 4 canonical products per 128-bit sum here
    U128 s_h0{0, 0}, s_h1{0, 0};  // side's Z times H_b, times H_a
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        const int h = 4 * quarter + k;
        u64 r = small ? walsh_lc<true>(a.cr[h], x0, x1, off, b) : walsh_lc<false>(a.cr[h], x0, x1, off, b);
        u64 im = small ? walsh_lc<true>(a.ci[h], x0, x1, off, b) : walsh_lc<false>(a.ci[h], x0, x1, off, b);
        if (side == 0) {
            r = add_mod(r, s_const[2 * h], b.q);
            im = add_mod(im, s_const[2 * h + 1], b.q);
        }
        const u64 p = a.pt_mont ? mont_mul(im, pt, b.q, b.qninv) : mul_mod(im, pt, b.q, b.mu_lo, b.mu_hi);
        const u64 z = add_mod(r, p, b.q);
        if (h == 0) {
            (side ? a.acc_a : a.acc_b)[i] = z;
            continue;
        }
        mac128(s_h0, z, a.high_b[h - 1][i]);
        mac128(s_h1, z, a.high_a[h - 1][i]);
    }
    s_part[side][0][quarter][lane] = reduce128(s_h0.lo, s_h0.hi, b.q, b.mu_lo, b.mu_hi);
    s_part[side][1][quarter][lane] = reduce128(s_h1.lo, s_h1.hi, b.q, b.mu_lo, b.mu_hi);
    __syncthreads();
    auto sum4 = [&](int sd, int pr) {
        return add_mod(add_mod(s_part[sd][pr][0][lane], s_part[sd][pr][1][lane], b.q),
                       add_mod(s_part[sd][pr][2][lane], s_part[sd][pr][3][lane], b.q), b.q);
    };
    if (wave == 0) a.raw_b[i] = sum4(0, 0);                                  // Z_b H_b
    else if (wave == 1) a.raw_a[i] = add_mod(sum4(0, 1), sum4(1, 0), b.q);   // Z_b H_a + Z_a H_b
    else if (wave == 2) a.raw_c2[i] = sum4(1, 1);                            // Z_a H_a
}

void walsh_product_sum(const Context& c, int level, const u64* const* low_b, const u64* const* low_a,
                       const u64* const* high_b, const u64* const* high_a, const short* cr, const short* ci,
                       const u64* v_abs, const int* v_neg, const u64* pt, bool pt_mont, u64* acc_b, u64* acc_a,
                       u64* raw_b, u64* raw_a, u64* raw_c2, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid level");
    WalshArgs a{};
    for (int m = 0; m < WALSH_MASKS - 1; ++m) {
        if (!low_b[m] || !low_a[m] || !high_b[m] || !high_a[m]) throw Error(-1, "null sign ciphertext");
        a.low_b[m] = low_b[m];
        a.low_a[m] = low_a[m];
        a.high_b[m] = high_b[m];
        a.high_a[m] = high_a[m];
    }
    for (int h = 0; h < WALSH_MASKS; ++h)
        for (int m = 0; m < WALSH_MASKS; ++m) {
            a.cr[h][m] = cr[h * WALSH_MASKS + m];
            a.ci[h][m] = ci[h * WALSH_MASKS + m];
            if (a.cr[h][m] == -32768 || a.ci[h][m] == -32768) throw Error(-1, "coefficient out of range");
        }
    for (int t = 0; t < 2 * WALSH_MASKS; ++t) {
        a.v_lo[t] = v_abs[2 * t];
        a.v_hi[t] = v_abs[2 * t + 1];
        if (v_neg[t]) a.v_neg |= 1u << t;
    }
    a.pt = pt;
    a.pt_mont = pt_mont ? 1 : 0;
    a.acc_b = acc_b;
    a.acc_a = acc_a;
    a.raw_b = raw_b;
    a.raw_a = raw_a;
    a.raw_c2 = raw_c2;
    a.logN = c.logN;
    a.bc = c.d_bc;
    const long long words = (long long)(level + 1) << c.logN;  // a multiple of WALSH_COEFFS
    hipLaunchKernelGGL(k_walsh_product_sum, dim3((unsigned)(words / WALSH_COEFFS)), dim3(WALSH_BLOCK), 0, s, a);
    MI_HIP(hipGetLastError());
}

}  // namespace mi
