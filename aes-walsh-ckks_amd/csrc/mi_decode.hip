// Decoding on the device: CRT interpolation of the coefficient-domain
// plaintext, the division by the scale and the inverse special FFT
// (crt_interpolate, ckks_encoding.cpp:254-329; fft_dit, fft.cpp:37-117).
//
// Every step reproduces the host decode bit for bit:
//   * CRT over the first `use` limbs whose product P fits in 128 bits: the
//     same modular sum (x_k * phat_k^-1 mod q_k) * phat_k, accumulated mod P
//     by compare-and-subtract, centred, converted to double with round to
//     nearest even (the conversion of an unsigned 128-bit integer);
//   * the division (double)((long double)v / scale): the host computes the
//     quotient rounded to a 64-bit significand (x87 extended precision) and
//     rounds that to double; here the 64-bit significand comes from an exact
//     128-by-64-bit integer division of the significands, rounded to nearest
//     even, then rounded again to 53 bits the same way;
//   * the FFT stages in double with the host's operation order (complex
//     product (ac - bd, ad + bc), then t + u and t - u; no contraction).
#include "mi_arith.hpp"
#include "mi_internal.hpp"

namespace mi {

namespace {

constexpr int DEC_THREADS = 256;

struct U128v {
    u64 lo, hi;
};
__device__ __forceinline__ bool ge128(U128v a, U128v b) { return a.hi > b.hi || (a.hi == b.hi && a.lo >= b.lo); }
__device__ __forceinline__ U128v sub128(U128v a, U128v b) {
    U128v r{a.lo - b.lo, a.hi - b.hi - (a.lo < b.lo ? 1 : 0)};
    return r;
}
__device__ __forceinline__ U128v add128(U128v a, U128v b) {
    U128v r{a.lo + b.lo, a.hi + b.hi};
    r.hi += r.lo < a.lo ? 1 : 0;
    return r;
}

// round to nearest even of an unsigned 128-bit integer to double
__device__ __forceinline__ double u128_to_double(U128v a) {
    if (a.hi == 0) return (double)a.lo;  // u64 -> f64 is correctly rounded
    const int lz = __clzll(a.hi);
    // top 64 bits, the bits shifted out folded into a sticky lowest bit
    const int sh = 64 - lz;  // bits of a.lo below the top 64
    const u64 top = (a.hi << lz) | (lz ? (a.lo >> (64 - lz)) : 0);
    const u64 below = lz ? (a.lo << lz) : a.lo;
    const u64 t = top | (below ? 1 : 0);
    return ldexp((double)t, sh);
}

// floor((n_hi 2^64 + n_lo) / d) and the remainder, for n_hi < d (quotient < 2^64)
__device__ __forceinline__ u64 div128by64(u64 n_hi, u64 n_lo, u64 d, u64& rem) {
    u64 q = 0, r = n_hi;
#pragma unroll 1
    for (int i = 63; i >= 0; --i) {
        const bool carry = (r >> 63) != 0;
        r = (r << 1) | ((n_lo >> i) & 1);
        if (carry || r >= d) {
            r -= d;
            q |= 1ull << i;
        }
    }
    rem = r;
    return q;
}

// (double)((long double)v / S), S = mS 2^eS with 2^63 <= mS < 2^64, as the
// x87 extended-precision division followed by the conversion to double
__device__ double div_by_extended(double v, u64 mS, int eS) {
    if (v == 0.0) return 0.0;
    const bool neg = v < 0.0;
    int ev;
    const double fv = frexp(neg ? -v : v, &ev);  // fv in [0.5, 1)
    const u64 mv = (u64)ldexp(fv, 53);           // 2^52 <= mv < 2^53, v = mv 2^(ev-53)
    // Q = floor(mv 2^75 / mS), 2^63 < Q < 2^65; mv 2^75 as 128 bits is
    // (high word mv 2^11, low word 0), and mv 2^11 < 2^64
    const u64 hi = mv << 11;
    // the quotient can reach 2^65: divide the top word first
    const u64 q1 = hi / mS;                       // 0 or 1 (hi < 2^64 <= 2 mS)
    const u64 r1 = hi - q1 * mS;
    u64 rem;
    const u64 q0 = div128by64(r1, 0, mS, rem);    // floor((r1 2^64) / mS)
    // Q = q1 2^64 + q0; value = Q 2^(ev - 53 - 75 - eS) plus rem / mS of an ulp
    int e = ev - 53 - 75 - eS;
    u64 sig;  // the quotient rounded to a 64-bit significand
    if (q1) {
        // 65 bits: the lowest (q0 bit 0) is dropped, rem below it; nearest even
        const u64 low = q0 & 1;
        sig = (q0 >> 1) | (1ull << 63);
        e += 1;
        if (low && (rem != 0 || (sig & 1))) {
            ++sig;
            if (sig == 0) {  // carried out of 64 bits
                sig = 1ull << 63;
                e += 1;
            }
        }
    } else {
        // 64 bits: round with the remainder (compare 2 rem with mS)
        sig = q0;
        const bool half_or_more = rem >= mS - rem;  // 2 rem >= mS without overflow
        const bool exact_half = rem == mS - rem;
        if (half_or_more && (!exact_half || (sig & 1))) {
            ++sig;
            if (sig == 0) {
                sig = 1ull << 63;
                e += 1;
            }
        }
    }
    // round the 64-bit significand to 53 bits, ties to even
    const u64 drop = sig & 0x7ff;
    u64 m53 = sig >> 11;
    if (drop > 0x400 || (drop == 0x400 && (m53 & 1))) {
        ++m53;
        if (m53 == (1ull << 53)) {
            m53 >>= 1;
            e += 1;
        }
    }
    const double r = ldexp((double)m53, e + 11);
    return neg ? -r : r;
}

struct CrtArgs {
    const u64* coeffs[64];  // per ciphertext: limbs x N canonical coefficients
    double* out;            // per ciphertext 2N doubles (complex m)
    u64 mS[64];
    int eS[64];
    int use;                // CRT limbs
    u64 q[4], mu_lo[4], mu_hi[4], phat_inv[4];
    U128v phat[4], P, half_p;
    int logN;
};

__global__ void __launch_bounds__(DEC_THREADS) k_dec_crt(CrtArgs a) {
    const u64 N = 1ull << a.logN;
    const int ct = blockIdx.y;
    const u64 j = (u64)blockIdx.x * DEC_THREADS + threadIdx.x;
    if (j >= N) return;
    U128v acc{0, 0};
    for (int k = 0; k < a.use; ++k) {
        const u64 x = a.coeffs[ct][(u64)k * N + j];
        const u64 al = mul_mod(x, a.phat_inv[k], a.q[k], a.mu_lo[k], a.mu_hi[k]);
        // term = al * phat_k (< P < 2^128, exact in 128 bits)
        U128v term{al * a.phat[k].lo, __umul64hi(al, a.phat[k].lo) + al * a.phat[k].hi};
        const U128v rem = sub128(a.P, acc);
        acc = ge128(term, rem) ? sub128(term, rem) : add128(acc, term);
    }
    // centred: acc > P/2 is negative
    const double v = ge128(a.half_p, acc) ? u128_to_double(acc) : -1.0 * u128_to_double(sub128(a.P, acc));
    double* m = a.out + (size_t)ct * 2 * N;
    m[2 * j] = div_by_extended(v, a.mS[ct], a.eS[ct]);
    m[2 * j + 1] = 0.0;
}

// one radix-2 stage of fft_dit: m groups of span 2k (tw[m + i] for group i)
__global__ void __launch_bounds__(DEC_THREADS)
k_dec_fft_stage(double* __restrict__ x, const double* __restrict__ tw, int logN, int m, int k, long long total) {
    const long long t = (long long)blockIdx.x * DEC_THREADS + threadIdx.x;
    if (t >= total) return;
    const long long N = 1ll << logN;
    const long long half = N >> 1;
    const long long ct = t / half, b = t - ct * half;
    const long long i = b / k, j = b - i * k;
    double* v = x + ct * 2 * N;
    const long long i0 = (i << 1) * k + j, i1 = i0 + k;
    const double wr = tw[2 * (m + i)], wi = tw[2 * (m + i) + 1];
    const double tr = v[2 * i0], ti = v[2 * i0 + 1];
    const double ar = v[2 * i1], ai = v[2 * i1 + 1];
    const double ur = __dsub_rn(__dmul_rn(ar, wr), __dmul_rn(ai, wi));
    const double ui = __dadd_rn(__dmul_rn(ar, wi), __dmul_rn(ai, wr));
    v[2 * i0] = __dadd_rn(tr, ur);
    v[2 * i0 + 1] = __dadd_rn(ti, ui);
    v[2 * i1] = __dsub_rn(tr, ur);
    v[2 * i1 + 1] = __dsub_rn(ti, ui);
}

// slots: out[i] = m[perm[i]]; coefficients: out[i] = (m[brev i].re, m[N/2 + brev i].re)
__global__ void __launch_bounds__(DEC_THREADS)
k_dec_out(const double* __restrict__ x, const int* __restrict__ perm, double* __restrict__ out, int logN,
          int coeff_domain, long long total) {
    const long long t = (long long)blockIdx.x * DEC_THREADS + threadIdx.x;
    if (t >= total) return;
    const long long N = 1ll << logN, half = N >> 1;
    const long long ct = t / half, i = t - ct * half;
    const double* v = x + ct * 2 * N;
    double* o = out + ct * N + 2 * i;
    if (coeff_domain) {
        const long long ci = (long long)(__brev((unsigned)i) >> (32 - (logN - 1)));
        o[0] = v[2 * ci];
        o[1] = v[2 * (half + ci)];
    } else {
        const long long s = perm[i];
        o[0] = v[2 * s];
        o[1] = v[2 * s + 1];
    }
}

unsigned dec_blocks(long long work) { return (unsigned)((work + DEC_THREADS - 1) / DEC_THREADS); }

}  // namespace

void set_decoding_tables(Context& c, const double* twiddles, const u64* perm) {
    const int N = (int)c.N;
    std::vector<int> p(N / 2);
    for (int i = 0; i < N / 2; ++i) {
        if (perm[i] >= (u64)N) throw Error(-1, "slot permutation entry out of range");
        p[i] = (int)perm[i];
    }
    if (!c.d_dec_tw) MI_HIP(hipMalloc(&c.d_dec_tw, sizeof(double) * 2 * N));
    if (!c.d_dec_perm) MI_HIP(hipMalloc(&c.d_dec_perm, sizeof(int) * N / 2));
    MI_HIP(hipMemcpy(c.d_dec_tw, twiddles, sizeof(double) * 2 * N, hipMemcpyHostToDevice));
    MI_HIP(hipMemcpy(c.d_dec_perm, p.data(), sizeof(int) * (N / 2), hipMemcpyHostToDevice));
}

void decode_batch(Context& c, int count, const u64* const* coeffs, int limbs, const u64* scale_mant,
                  const int* scale_exp, bool coeff_domain, double* out, hipStream_t s) {
    if (!c.d_dec_tw) throw Error(-2, "decoding tables are not set (mi_set_decoding_tables)");
    if (count < 1 || limbs < 1 || limbs > c.nq) throw Error(-1, "invalid count or limb count");
    const u64 N = c.N;
    // the CRT limbs: the longest prefix of q_0.. whose product fits in 128 bits
    unsigned __int128 P = 1;
    int use = 0;
    for (int i = 0; i < limbs && use < 4; ++i) {
        if (P > (~(unsigned __int128)0) / c.q[i]) break;
        P *= c.q[i];
        ++use;
    }
    CrtArgs base{};
    base.use = use;
    base.logN = c.logN;
    base.P = U128v{(u64)P, (u64)(P >> 64)};
    const unsigned __int128 hp = P / 2;
    base.half_p = U128v{(u64)hp, (u64)(hp >> 64)};
    for (int k = 0; k < use; ++k) {
        const u64 q = c.q[k];
        const unsigned __int128 ph = P / q;
        base.phat[k] = U128v{(u64)ph, (u64)(ph >> 64)};
        u64 r = (u64)(ph % q), e = q - 2, acc = 1;  // phat^-1 mod q by Fermat
        while (e) {
            if (e & 1) acc = (u64)((unsigned __int128)acc * r % q);
            r = (u64)((unsigned __int128)r * r % q);
            e >>= 1;
        }
        base.phat_inv[k] = acc;
        base.q[k] = q;
        base.mu_lo[k] = c.h_bc[k].mu_lo;
        base.mu_hi[k] = c.h_bc[k].mu_hi;
    }
    double* work = (double*)c.workspace(s, sizeof(double) * 2 * N * (size_t)std::min(count, 64));
    for (int c0 = 0; c0 < count; c0 += 64) {
        const int n = std::min(64, count - c0);
        CrtArgs a = base;
        for (int i = 0; i < n; ++i) {
            a.coeffs[i] = coeffs[c0 + i];
            a.mS[i] = scale_mant[c0 + i];
            a.eS[i] = scale_exp[c0 + i];
            if ((a.mS[i] >> 63) == 0) throw Error(-1, "scale significand must be normalised (top bit set)");
        }
        a.out = work;
        hipLaunchKernelGGL(k_dec_crt, dim3((unsigned)(N / DEC_THREADS), n), dim3(DEC_THREADS), 0, s, a);
        MI_HIP(hipGetLastError());
        const long long total = (long long)n * (long long)(N >> 1);
        if (!coeff_domain)
            for (int m = 1, k = (int)(N >> 1); m < (int)N; m <<= 1, k >>= 1) {
                hipLaunchKernelGGL(k_dec_fft_stage, dim3(dec_blocks(total)), dim3(DEC_THREADS), 0, s, work,
                                   c.d_dec_tw, c.logN, m, k, total);
                MI_HIP(hipGetLastError());
            }
        hipLaunchKernelGGL(k_dec_out, dim3(dec_blocks(total)), dim3(DEC_THREADS), 0, s, work, c.d_dec_perm,
                           out + (size_t)c0 * N, c.logN, coeff_domain ? 1 : 0, total);
        MI_HIP(hipGetLastError());
    }
}

}  // namespace mi
