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

constexpr int WALSH_THREADS = 256;
constexpr int WALSH_MASKS = 16;

struct WalshArgs {
    const u64* low_b[WALSH_MASKS - 1];  // masks 1..15, L limbs each
    const u64* low_a[WALSH_MASKS - 1];
    const u64* high_b[WALSH_MASKS - 1];
    const u64* high_a[WALSH_MASKS - 1];
    const u64* pt;                      // L limbs, evaluation domain
    u64 *acc_b, *acc_a, *raw_b, *raw_a, *raw_c2;
    short cr[WALSH_MASKS][WALSH_MASKS];  // [h][m]
    short ci[WALSH_MASKS][WALSH_MASKS];
    u64 v_lo[2 * WALSH_MASKS], v_hi[2 * WALSH_MASKS];  // |V| of (h, real) at 2h, (h, imag) at 2h+1
    unsigned v_neg;                                     // sign bit per entry
    int logN, pt_mont;
    const BasisConst* bc;
};

// sum_m c[m] x_m mod q for |c| < 2^15, x_m = x1 2^31 + x0 given as 31-bit
// halves: two signed 32x32+64 multiply-adds per term, |partial sums| < 2^50,
// shifted non-negative by off (a multiple of q above 2^50), then one 128-bit
// reduction of (A1 + off) 2^31 + (A0 + off).
__device__ __forceinline__ u64 walsh_lc(const short* c, const int (&x0)[WALSH_MASKS - 1],
                                        const int (&x1)[WALSH_MASKS - 1], u64 off, const BasisConst& b) {
    long long a0 = 0, a1 = 0;
#pragma unroll
    for (int m = 1; m < WALSH_MASKS; ++m) {
        const int cm = c[m];
        a0 += (long long)cm * x0[m - 1];
        a1 += (long long)cm * x1[m - 1];
    }
    const u64 u = (u64)(a1 + (long long)off), w = (u64)(a0 + (long long)off);
    const u64 lo = (u << 31) + w;
    const u64 hi = (u >> 33) + (lo < w ? 1 : 0);
    return reduce128(lo, hi, b.q, b.mu_lo, b.mu_hi);
}

// A block holds 64 coefficients; its 4 waves each take 4 of the 16 high
// masks (wave 0: h = 0..3, whose h = 0 is the acc term) for all 64, so a pair
// launches 4 blocks per 256 coefficients.  The waves' partial sums of the raw
// products meet in LDS.
constexpr int WALSH_COEFFS = WALSH_THREADS / 4;

__global__ void __launch_bounds__(WALSH_THREADS) k_walsh_product_sum(WalshArgs a) {
    __shared__ u64 s_const[2 * WALSH_MASKS];
    __shared__ u64 s_part[3][4][WALSH_COEFFS];
    const int l = (int)(((long long)blockIdx.x * WALSH_COEFFS) >> a.logN);  // uniform: a block is in one limb
    const int lane = threadIdx.x % WALSH_COEFFS;
    const int quarter = __builtin_amdgcn_readfirstlane(threadIdx.x / WALSH_COEFFS);  // one wave per quarter
    const long long i = (long long)blockIdx.x * WALSH_COEFFS + lane;
    const BasisConst b = a.bc[l];
    if (threadIdx.x < 2 * WALSH_MASKS) {
        const int t = threadIdx.x;
        const u64 r = reduce128(a.v_lo[t], a.v_hi[t], b.q, b.mu_lo, b.mu_hi);
        s_const[t] = ((a.v_neg >> t) & 1) && r ? b.q - r : r;
    }
    // q * ceil(2^50 / q): a multiple of q above every partial sum's magnitude
    const u64 off = b.q * (((1ull << 50) + b.q - 1) / b.q);
    int lb0[WALSH_MASKS - 1], lb1[WALSH_MASKS - 1], la0[WALSH_MASKS - 1], la1[WALSH_MASKS - 1];
#pragma unroll
    for (int m = 0; m < WALSH_MASKS - 1; ++m) {
        const u64 xb = a.low_b[m][i], xa = a.low_a[m][i];
        lb0[m] = (int)(xb & 0x7fffffffu);
        lb1[m] = (int)(xb >> 31);
        la0[m] = (int)(xa & 0x7fffffffu);
        la1[m] = (int)(xa >> 31);
    }
    const u64 pt = a.pt[i];
    __syncthreads();
    auto z_pair = [&](int h, u64& zb, u64& za) {
        const u64 rb = add_mod(walsh_lc(a.cr[h], lb0, lb1, off, b), s_const[2 * h], b.q);
        const u64 ra = walsh_lc(a.cr[h], la0, la1, off, b);
        const u64 ib = add_mod(walsh_lc(a.ci[h], lb0, lb1, off, b), s_const[2 * h + 1], b.q);
        const u64 ia = walsh_lc(a.ci[h], la0, la1, off, b);
        const u64 pb = a.pt_mont ? mont_mul(ib, pt, b.q, b.qninv) : mul_mod(ib, pt, b.q, b.mu_lo, b.mu_hi);
        const u64 pa = a.pt_mont ? mont_mul(ia, pt, b.q, b.qninv) : mul_mod(ia, pt, b.q, b.mu_lo, b.mu_hi);
        zb = add_mod(rb, pb, b.q);
        za = add_mod(ra, pa, b.q);
    };
    // at most 4 canonical products per 128-bit sum here
    U128 sb{0, 0}, sab{0, 0}, sba{0, 0}, sc{0, 0};
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        const int h = 4 * quarter + k;
        u64 zb, za;
        z_pair(h, zb, za);
        if (h == 0) {
            a.acc_b[i] = zb;
            a.acc_a[i] = za;
            continue;
        }
        const u64 hb = a.high_b[h - 1][i], ha = a.high_a[h - 1][i];
        mac128(sb, zb, hb);
        mac128(sab, zb, ha);
        mac128(sba, za, hb);
        mac128(sc, za, ha);
    }
    s_part[0][quarter][lane] = reduce128(sb.lo, sb.hi, b.q, b.mu_lo, b.mu_hi);
    s_part[1][quarter][lane] = add_mod(reduce128(sab.lo, sab.hi, b.q, b.mu_lo, b.mu_hi),
                                       reduce128(sba.lo, sba.hi, b.q, b.mu_lo, b.mu_hi), b.q);
    s_part[2][quarter][lane] = reduce128(sc.lo, sc.hi, b.q, b.mu_lo, b.mu_hi);
    __syncthreads();
    if (quarter < 3) {  // wave w sums raw poly w over the quarters
        const u64 v = add_mod(add_mod(s_part[quarter][0][lane], s_part[quarter][1][lane], b.q),
                              add_mod(s_part[quarter][2][lane], s_part[quarter][3][lane], b.q), b.q);
        u64* out = quarter == 0 ? a.raw_b : quarter == 1 ? a.raw_a : a.raw_c2;
        out[i] = v;
    }
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
    hipLaunchKernelGGL(k_walsh_product_sum, dim3((unsigned)(words / WALSH_COEFFS)), dim3(WALSH_THREADS), 0, s,
                       a);
    MI_HIP(hipGetLastError());
}

}  // namespace mi
