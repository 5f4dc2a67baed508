// Slot encoding on the device (encode_slots_custom_basis,
// ckks_encoding.cpp:368-410, with fft_dif of fft.cpp:75-95):
//   ev[perm[i]] = z_i, ev[N-1-perm[i]] = conj(z_i)     (slot placement)
//   ev = FFT_dif(ev) with the context's twiddles      (N-point, in place)
//   pt_l[j] = llrint(Re(ev[j]) * (scale / N)) mod q_l (signed residues)
// then the forward NTT of every limb for an evaluation-domain plaintext.
// Every double operation is a separately rounded IEEE multiply, add or
// subtract in the reference's order (no contraction), so the rounded
// integers - and hence the residues - equal the host restatement's bit for bit.
#include "mi_ntt_kernels.hpp"

#include <cmath>
#include <mutex>
#include <vector>

namespace mi {

constexpr int ENC_THREADS = 256;

// inv[j] = i + 1: ev[j] = z_i; -(i + 1): ev[j] = conj(z_i); 0: ev[j] = 0
__global__ void __launch_bounds__(ENC_THREADS)
k_enc_place(const double* __restrict__ z, const int* __restrict__ inv, double* __restrict__ ev, int N,
            long long total) {
    const long long t = (long long)blockIdx.x * ENC_THREADS + threadIdx.x;
    if (t >= total) return;
    const long long b = t / N;
    const int j = (int)(t - b * N);
    const int s = inv[j];
    double re = 0.0, im = 0.0;
    if (s != 0) {
        const int i = (s > 0 ? s : -s) - 1;
        const double* zi = z + (b * (N >> 1) + i) * 2;
        re = zi[0];
        im = s > 0 ? zi[1] : -zi[1];
    }
    ev[t * 2] = re;
    ev[t * 2 + 1] = im;
}

// One decimation-in-frequency stage: block pairs (2i k + j, 2i k + j + k)
// with twiddle tw[m + i]; x0 = u + v, x1 = (u - v) w.
__global__ void __launch_bounds__(ENC_THREADS)
k_fft_dif_stage(double* __restrict__ ev, const double* __restrict__ tw, int N, int m, int k, long long total) {
    const long long t = (long long)blockIdx.x * ENC_THREADS + threadIdx.x;
    if (t >= total) return;
    const int half = N >> 1;
    const long long b = t / half;
    const int r = (int)(t - b * half);
    const int i = r / k, j = r - i * k;
    double* x0 = ev + (b * N + 2LL * i * k + j) * 2;
    double* x1 = x0 + 2LL * k;
    const double ur = x0[0], ui = x0[1], vr = x1[0], vi = x1[1];
    const double wr = tw[2 * (m + i)], wi = tw[2 * (m + i) + 1];
    x0[0] = __dadd_rn(ur, vr);
    x0[1] = __dadd_rn(ui, vi);
    const double dr = __dsub_rn(ur, vr), di = __dsub_rn(ui, vi);
    x1[0] = __dsub_rn(__dmul_rn(dr, wr), __dmul_rn(di, wi));
    x1[1] = __dadd_rn(__dmul_rn(dr, wi), __dmul_rn(di, wr));
}

// Rounded real parts as signed residues of every limb.
__global__ void __launch_bounds__(ENC_THREADS)
k_enc_round(const double* __restrict__ ev, const BasisConst* __restrict__ bc, u64* __restrict__ out, int N,
            int limbs, double k, long long stride, long long total) {
    const long long t = (long long)blockIdx.x * ENC_THREADS + threadIdx.x;
    if (t >= total) return;
    const long long per = (long long)limbs * N;
    const long long b = t / per;
    const long long r = t - b * per;
    const int l = (int)(r / N), j = (int)(r - (long long)l * N);
    const long long v = __double2ll_rn(__dmul_rn(ev[(b * N + j) * 2], k));
    const u64 q = bc[l].q;
    u64 res;
    if (v >= 0) {
        res = (u64)v % q;
    } else {
        const u64 m = (u64)(-(v + 1)) % q;  // (-v - 1) mod q, no overflow at INT64_MIN
        res = q - 1 - m;                    // (-v) mod q subtracted from q, in [0, q)
    }
    out[b * stride + r] = res;
}

static unsigned enc_blocks(long long work) { return (unsigned)((work + ENC_THREADS - 1) / ENC_THREADS); }

void set_encoding_tables(Context& c, const double* twiddles, const u64* perm) {
    const int N = (int)c.N;
    std::vector<int> inv(N, 0);
    for (int i = 0; i < N / 2; ++i) {
        const u64 d = perm[i];
        if (d >= (u64)N) throw Error(-1, "slot permutation entry out of range");
        inv[d] = i + 1;
        inv[N - 1 - d] = -(i + 1);
    }
    if (!c.d_enc_tw) MI_HIP(hipMalloc(&c.d_enc_tw, sizeof(double) * 2 * N));
    if (!c.d_enc_inv) MI_HIP(hipMalloc(&c.d_enc_inv, sizeof(int) * N));
    MI_HIP(hipMemcpy(c.d_enc_tw, twiddles, sizeof(double) * 2 * N, hipMemcpyHostToDevice));
    MI_HIP(hipMemcpy(c.d_enc_inv, inv.data(), sizeof(int) * N, hipMemcpyHostToDevice));
}

void encode_slots(Context& c, const double* values, int count, double scale, int limbs, bool eval, u64* out,
                  long long stride, hipStream_t s) {
    if (!c.d_enc_tw) throw Error(-2, "encoding tables are not set (mi_set_encoding_tables)");
    if (count < 1 || limbs < 1 || limbs > c.nq) throw Error(-1, "invalid count or limb count");
    if (!(scale > 0.0)) throw Error(-1, "scale must be positive");
    if (stride < (long long)limbs * (long long)c.N) throw Error(-1, "poly stride smaller than limbs * N");
    const int N = (int)c.N;
    const size_t zwords = (size_t)count * N;   // count x N/2 complex
    const size_t evwords = (size_t)count * 2 * N;
    double* z = (double*)c.workspace(s, sizeof(double) * (zwords + evwords));
    double* ev = z + zwords;
    MI_HIP(hipMemcpyAsync(z, values, sizeof(double) * zwords, hipMemcpyHostToDevice, s));
    const long long total = (long long)count * N;
    hipLaunchKernelGGL(k_enc_place, dim3(enc_blocks(total)), dim3(ENC_THREADS), 0, s, z, c.d_enc_inv, ev, N,
                       total);
    MI_HIP(hipGetLastError());
    const long long bf = (long long)count * (N >> 1);
    for (int m = N >> 1, k = 1; m > 0; m >>= 1, k <<= 1) {
        hipLaunchKernelGGL(k_fft_dif_stage, dim3(enc_blocks(bf)), dim3(ENC_THREADS), 0, s, ev, c.d_enc_tw, N, m,
                           k, bf);
        MI_HIP(hipGetLastError());
    }
    const long long rt = (long long)count * limbs * N;
    hipLaunchKernelGGL(k_enc_round, dim3(enc_blocks(rt)), dim3(ENC_THREADS), 0, s, ev, c.d_bc, out, N, limbs,
                       scale / (double)N, stride, rt);
    MI_HIP(hipGetLastError());
    if (eval) ntt_forward(c, LimbBatch{out, stride, limbs, 0, count * limbs}, s);
    // the staged values live in the stream's workspace: finish before the
    // next call can reuse it from the host side
    MI_HIP(hipStreamSynchronize(s));
}

}  // namespace mi

namespace mi {

// ---- sampling for encryption and key generation ----------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): a counter-based generator, so every
// coefficient's sample is a pure function of (seed, counter, limb, index)
// and any launch order gives the same polynomial.
struct Philox {
    u32 v[4];
};
__device__ __forceinline__ Philox philox4x32_10(u32 c0, u32 c1, u32 c2, u32 c3, u32 k0, u32 k1) {
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const u64 p0 = (u64)0xD2511F53u * c0, p1 = (u64)0xCD9E8D57u * c2;
        const u32 hi0 = (u32)(p0 >> 32), lo0 = (u32)p0, hi1 = (u32)(p1 >> 32), lo1 = (u32)p1;
        const u32 n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return Philox{{c0, c1, c2, c3}};
}

// One block of 128 random bits for (seed, counter, limb, index).
__device__ __forceinline__ Philox draw(u64 seed, u64 counter, u32 limb, u32 index) {
    return philox4x32_10(index, limb, (u32)counter, (u32)(counter >> 32), (u32)seed, (u32)(seed >> 32));
}

// Uniform residues: floor(r * q / 2^128) of a 128-bit r (bias below 2^-66).
__global__ void __launch_bounds__(ENC_THREADS)
k_sample_uniform(u64* __restrict__ out, const BasisConst* __restrict__ bc, int logN, u64 seed, u64 counter,
                 long long total) {
    const long long t = (long long)blockIdx.x * ENC_THREADS + threadIdx.x;
    if (t >= total) return;
    const int l = (int)(t >> logN);
    const u32 i = (u32)(t & (((long long)1 << logN) - 1));
    const Philox r = draw(seed, counter, (u32)l, i);
    const u64 lo = ((u64)r.v[1] << 32) | r.v[0], hi = ((u64)r.v[3] << 32) | r.v[2];
    const u64 q = bc[l].q;
    const U128 a = mul128(hi, q);
    const u64 c = __umul64hi(lo, q);
    const u64 s = a.lo + c;
    out[t] = a.hi + (s < c ? 1 : 0);
}

// Discrete Gaussian by inversion of the cumulative table over [-bound, bound]
// (u in [0, 1) with 53 random bits); one integer per coefficient, written as
// a residue of every limb.
__global__ void __launch_bounds__(ENC_THREADS)
k_sample_gaussian(u64* __restrict__ out, const BasisConst* __restrict__ bc, const double* __restrict__ cdf,
                  int ncdf, int bound, int limbs, int logN, u64 seed, u64 counter) {
    const long long N = (long long)1 << logN;
    const long long i = (long long)blockIdx.x * ENC_THREADS + threadIdx.x;
    if (i >= N) return;
    const Philox r = draw(seed, counter, 0xffffffffu, (u32)i);
    const u64 bits = ((u64)r.v[1] << 32) | r.v[0];
    const double u = (double)(bits >> 11) * 0x1.0p-53;
    int lo = 0, hi = ncdf - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    const long long e = (long long)lo - bound;
    for (int l = 0; l < limbs; ++l) {
        const u64 q = bc[l].q;
        out[(long long)l * N + i] = e >= 0 ? (u64)e : q - (u64)(-e);
    }
}

void sample_uniform(Context& c, u64* out, int limbs, int basis0, u64 seed, u64 counter, hipStream_t s) {
    if (limbs < 1 || basis0 < 0 || basis0 + limbs > c.nb()) throw Error(-1, "limb range outside the basis");
    const long long total = (long long)limbs * (long long)c.N;
    hipLaunchKernelGGL(k_sample_uniform, dim3(enc_blocks(total)), dim3(ENC_THREADS), 0, s, out, c.d_bc + basis0,
                       c.logN, seed, counter, total);
    MI_HIP(hipGetLastError());
}

void sample_gaussian(Context& c, u64* out, int limbs, double sigma, u64 seed, u64 counter, hipStream_t s) {
    if (limbs < 1 || limbs > c.nb()) throw Error(-1, "invalid limb count");
    if (!(sigma > 0.0) || sigma > 1e6) throw Error(-1, "sigma must be in (0, 1e6]");
    std::lock_guard<std::mutex> g(c.gauss_mu);
    if (c.gauss_sigma != sigma) {
        // the sampler's table: weights exp(-x^2 / 2 sigma^2) over |x| <= ceil(12.0061 sigma)
        const int bound = (int)std::ceil(12.00610553538285 * sigma);
        std::vector<double> cdf;
        const double var2 = 2.0 * sigma * sigma;
        double total = 0.0;
        for (long long x = -bound; x <= bound; ++x) {
            total += std::exp(-(double)(x * x) / var2);
            cdf.push_back(total);
        }
        for (double& v : cdf) v /= total;
        if (c.d_gauss_cdf) MI_HIP(hipStreamSynchronize(s));
        (void)hipFree(c.d_gauss_cdf);
        c.d_gauss_cdf = nullptr;
        MI_HIP(hipMalloc(&c.d_gauss_cdf, sizeof(double) * cdf.size()));
        MI_HIP(hipMemcpy(c.d_gauss_cdf, cdf.data(), sizeof(double) * cdf.size(), hipMemcpyHostToDevice));
        c.gauss_sigma = sigma;
        c.gauss_bound = bound;
        c.gauss_n = (int)cdf.size();
    }
    hipLaunchKernelGGL(k_sample_gaussian, dim3(enc_blocks((long long)c.N)), dim3(ENC_THREADS), 0, s, out, c.d_bc,
                       c.d_gauss_cdf, c.gauss_n, c.gauss_bound, limbs, c.logN, seed, counter);
    MI_HIP(hipGetLastError());
}

}  // namespace mi
