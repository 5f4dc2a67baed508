// Element-wise evaluator kernels, automorphisms and the fused rescale.
//
//   tensor / square / x plaintext / x int   ckks_mul.cpp:37-272
//   add / sub / + plaintext / + constant    ckks_addition.cpp:35-139
//   NTT-domain automorphism gather          ckks_keyswitch.cpp:37-60 with the
//                                           map of rotation_key_generation.cpp:80-90
//   rescale by the top modulus              ckks_rescale.cpp:68-148
// All products are exact (Barrett or Montgomery REDC), so canonical outputs
// equal the reference's.
#include "mi_ntt_kernels.hpp"

namespace mi {

constexpr int EW_THREADS = 256;

static unsigned ew_blocks(long long work) {
    long long b = (work + EW_THREADS - 1) / EW_THREADS;
    return (unsigned)(b > 0 ? b : 1);
}

// Element-wise launches over polys of `limbs` limbs: blockIdx.y is the poly,
// blockIdx.x walks its limbs * N words.  N is a multiple of EW_THREADS, so a
// block lies inside one limb: the limb index, its basis constants and its
// scalar are uniform over the block (scalar loads, no per-thread division).
static dim3 ew_grid(const Context& c, int limbs, int npolys) {
    if (npolys > 65535) throw Error(-1, "too many polynomials in one element-wise launch");
    return dim3((unsigned)(((long long)limbs << c.logN) / EW_THREADS), (unsigned)npolys);
}
__device__ __forceinline__ int ew_limb(int logN) { return (int)(((long long)blockIdx.x * EW_THREADS) >> logN); }

__global__ void __launch_bounds__(EW_THREADS)
k_ew_binary(const BasisConst* __restrict__ bc, int op, u64* x, const u64* __restrict__ y,
            int logN, long long xstride, long long ystride) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const long long k = blockIdx.y;
    const BasisConst b = bc[ew_limb(logN)];  // bc is pre-offset to the first limb's basis
    u64* px = x + k * xstride + r;
    const u64 vy = y[k * ystride + r];
    const u64 vx = *px;
    u64 o;
    switch (op) {
        case EW_ADD: o = add_mod(vx, vy, b.q); break;
        case EW_SUB: o = sub_mod(vx, vy, b.q); break;
        case EW_MUL: o = mul_mod(vx, vy, b.q, b.mu_lo, b.mu_hi); break;
        default: o = mont_mul(vx, vy, b.q, b.qninv); break;
    }
    *px = o;
}

void ew_binary(const Context& c, int op, u64* x, const u64* y, int limbs, int npolys,
               long long xstride, long long ystride, hipStream_t s, int basis0) {
    if (npolys <= 0 || limbs <= 0) return;
    hipLaunchKernelGGL(k_ew_binary, ew_grid(c, limbs, npolys), dim3(EW_THREADS), 0, s, c.d_bc + basis0, op, x,
                       y, c.logN, xstride, ystride);
    MI_HIP(hipGetLastError());
}

// Zero fill: whole 8-byte words at an 8-byte aligned base, bytes for the
// unaligned head and the tail.
__global__ void __launch_bounds__(EW_THREADS) k_fill_zero_words(u64* p, long long words) {
    const long long i = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    if (i < words) p[i] = 0;
}
__global__ void __launch_bounds__(64) k_fill_zero_bytes(unsigned char* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

void fill_zero(void* ptr, size_t bytes, hipStream_t s) {
    unsigned char* p = (unsigned char*)ptr;
    const size_t head = (8 - ((uintptr_t)p & 7)) & 7;
    if (head >= bytes) {
        hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(64), 0, s, p, (int)bytes);
        MI_HIP(hipGetLastError());
        return;
    }
    if (head) hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(64), 0, s, p, (int)head);
    const long long words = (long long)((bytes - head) / 8);
    const int tail = (int)((bytes - head) % 8);
    if (words)
        hipLaunchKernelGGL(k_fill_zero_words, dim3(ew_blocks(words)), dim3(EW_THREADS), 0, s, (u64*)(p + head),
                           words);
    if (tail) hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(64), 0, s, p + head + 8 * words, tail);
    MI_HIP(hipGetLastError());
}

struct Scalars {
    u64 v[MAX_LIMBS];
};

__global__ void __launch_bounds__(EW_THREADS)
k_ew_scalar(const BasisConst* __restrict__ bc, int op, u64* x, Scalars sc, int logN, long long stride) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const long long k = blockIdx.y;
    const int l = ew_limb(logN);
    const BasisConst b = bc[l];
    u64* px = x + k * stride + r;
    const u64 s = sc.v[l];
    *px = op == EW_ADD ? add_mod(*px, s, b.q) : mul_mod(*px, s, b.q, b.mu_lo, b.mu_hi);
}

// dst[k][l] += src[k][l] * scalar[l]
// (fused_multiply_by_real_scalar_then_add_inplace, ckks_evalpoly.cpp:352-398)
__global__ void __launch_bounds__(EW_THREADS)
k_ew_scalar_mac(const BasisConst* __restrict__ bc, u64* dst, const u64* __restrict__ src, Scalars sc,
                int logN, long long dstride, long long sstride) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const long long k = blockIdx.y;
    const int l = ew_limb(logN);
    const BasisConst b = bc[l];
    u64* pd = dst + k * dstride + r;
    *pd = add_mod(*pd, mul_mod(src[k * sstride + r], sc.v[l], b.q, b.mu_lo, b.mu_hi), b.q);
}

void ew_scalar_mac(const Context& c, u64* dst, const u64* src, const u64* scalars_host, int limbs,
                   int npolys, long long dstride, long long sstride, hipStream_t s) {
    if (limbs > MAX_LIMBS) throw Error(-1, "too many limbs for a scalar operation");
    Scalars sc{};
    for (int l = 0; l < limbs; ++l) {
        if (scalars_host[l] >= c.h_bc[l].q) throw Error(-1, "scalar residue is not reduced");
        sc.v[l] = scalars_host[l];
    }
    if (npolys <= 0 || limbs <= 0) return;
    hipLaunchKernelGGL(k_ew_scalar_mac, ew_grid(c, limbs, npolys), dim3(EW_THREADS), 0, s, c.d_bc, dst, src,
                       sc, c.logN, dstride, sstride);
    MI_HIP(hipGetLastError());
}

void ew_scalar(const Context& c, int op, u64* x, const u64* scalars_host, int limbs, int npolys,
               long long stride, hipStream_t s, int basis0) {
    if (limbs > MAX_LIMBS) throw Error(-1, "too many limbs for a scalar operation");
    Scalars sc{};
    for (int l = 0; l < limbs; ++l) {
        if (scalars_host[l] >= c.h_bc[basis0 + l].q) throw Error(-1, "scalar residue is not reduced");
        sc.v[l] = scalars_host[l];
    }
    if (npolys <= 0 || limbs <= 0) return;
    hipLaunchKernelGGL(k_ew_scalar, ew_grid(c, limbs, npolys), dim3(EW_THREADS), 0, s, c.d_bc + basis0, op, x,
                       sc, c.logN, stride);
    MI_HIP(hipGetLastError());
}

// ckks_mul.cpp:127-154: c2 = a1 a2, b1 <- b1 b2, a1 <- a1 b2 + b1 a2
__global__ void __launch_bounds__(EW_THREADS)
k_tensor(const BasisConst* __restrict__ bc, int logN, long long total, u64* b1, u64* a1,
         const u64* __restrict__ b2, const u64* __restrict__ a2, u64* c2) {
    long long i = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= total) return;
    const BasisConst b = bc[ew_limb(logN)];
    const u64 vb1 = b1[i], va1 = a1[i], vb2 = b2[i], va2 = a2[i];
    c2[i] = mul_mod(va1, va2, b.q, b.mu_lo, b.mu_hi);
    b1[i] = mul_mod(vb1, vb2, b.q, b.mu_lo, b.mu_hi);
    a1[i] = add_mod(mul_mod(va1, vb2, b.q, b.mu_lo, b.mu_hi), mul_mod(vb1, va2, b.q, b.mu_lo, b.mu_hi), b.q);
}

void tensor(const Context& c, int limbs, u64* b1, u64* a1, const u64* b2, const u64* a2, u64* c2,
            hipStream_t s) {
    long long total = (long long)limbs * (long long)c.N;
    hipLaunchKernelGGL(k_tensor, dim3(ew_blocks(total)), dim3(EW_THREADS), 0, s, c.d_bc, c.logN,
                       total, b1, a1, b2, a2, c2);
    MI_HIP(hipGetLastError());
}

// ckks_mul.cpp:69-83: c2 = a^2, b <- b^2, a <- 2ab
__global__ void __launch_bounds__(EW_THREADS)
k_square(const BasisConst* __restrict__ bc, int logN, long long total, u64* b, u64* a, u64* c2) {
    long long i = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= total) return;
    const BasisConst k = bc[ew_limb(logN)];
    const u64 vb = b[i], va = a[i];
    c2[i] = mul_mod(va, va, k.q, k.mu_lo, k.mu_hi);
    b[i] = mul_mod(vb, vb, k.q, k.mu_lo, k.mu_hi);
    const u64 ab = mul_mod(va, vb, k.q, k.mu_lo, k.mu_hi);
    a[i] = add_mod(ab, ab, k.q);
}

void square(const Context& c, int limbs, u64* b, u64* a, u64* c2, hipStream_t s) {
    long long total = (long long)limbs * (long long)c.N;
    hipLaunchKernelGGL(k_square, dim3(ew_blocks(total)), dim3(EW_THREADS), 0, s, c.d_bc, c.logN,
                       total, b, a, c2);
    MI_HIP(hipGetLastError());
}

// The tensor product (or square, when b2 is null) of listed ciphertext
// pairs: blockIdx.y is the pair, so a batch of ciphertexts at one level is
// one launch per TL_MAX pairs.  The first operand is read from (s1b, s1a)
// and the product written to (b1, a1, c2): the same polys in place, or new
// ones.
constexpr int TL_MAX = 16;
struct TensorList {
    const u64* s1b[TL_MAX];
    const u64* s1a[TL_MAX];
    u64* b1[TL_MAX];
    u64* a1[TL_MAX];
    const u64* b2[TL_MAX];
    const u64* a2[TL_MAX];
    u64* c2[TL_MAX];
};
template <bool SQUARE>
__global__ void __launch_bounds__(EW_THREADS) k_tensor_list(const BasisConst* __restrict__ bc, TensorList l, int logN) {
    const long long i = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const int k = blockIdx.y;
    const BasisConst b = bc[ew_limb(logN)];
    const u64 vb1 = l.s1b[k][i], va1 = l.s1a[k][i];
    if constexpr (SQUARE) {
        l.c2[k][i] = mul_mod(va1, va1, b.q, b.mu_lo, b.mu_hi);
        l.b1[k][i] = mul_mod(vb1, vb1, b.q, b.mu_lo, b.mu_hi);
        const u64 ab = mul_mod(va1, vb1, b.q, b.mu_lo, b.mu_hi);
        l.a1[k][i] = add_mod(ab, ab, b.q);
    } else {
        const u64 vb2 = l.b2[k][i], va2 = l.a2[k][i];
        l.c2[k][i] = mul_mod(va1, va2, b.q, b.mu_lo, b.mu_hi);
        l.b1[k][i] = mul_mod(vb1, vb2, b.q, b.mu_lo, b.mu_hi);
        l.a1[k][i] = add_mod(mul_mod(va1, vb2, b.q, b.mu_lo, b.mu_hi), mul_mod(vb1, va2, b.q, b.mu_lo, b.mu_hi), b.q);
    }
}

void tensor_list(const Context& c, int level, int count, u64* const* b1, u64* const* a1, const u64* const* b2,
                 const u64* const* a2, u64* const* c2, hipStream_t s, const u64* const* s1b,
                 const u64* const* s1a) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid level");
    const bool sq = b2 == nullptr;
    for (int k0 = 0; k0 < count; k0 += TL_MAX) {
        const int n = std::min(TL_MAX, count - k0);
        TensorList l{};
        for (int k = 0; k < n; ++k) {
            l.s1b[k] = s1b ? s1b[k0 + k] : b1[k0 + k];
            l.s1a[k] = s1a ? s1a[k0 + k] : a1[k0 + k];
            l.b1[k] = b1[k0 + k];
            l.a1[k] = a1[k0 + k];
            l.b2[k] = sq ? nullptr : b2[k0 + k];
            l.a2[k] = sq ? nullptr : a2[k0 + k];
            l.c2[k] = c2[k0 + k];
        }
        if (sq)
            hipLaunchKernelGGL(k_tensor_list<true>, ew_grid(c, level + 1, n), dim3(EW_THREADS), 0, s, c.d_bc, l,
                               c.logN);
        else
            hipLaunchKernelGGL(k_tensor_list<false>, ew_grid(c, level + 1, n), dim3(EW_THREADS), 0, s, c.d_bc, l,
                               c.logN);
        MI_HIP(hipGetLastError());
    }
}

// b += REDC(1 * pt): Montgomery plaintext back to normal form (ckks_encoding.cpp:67-78)
__global__ void __launch_bounds__(EW_THREADS)
k_add_plain_mont(const BasisConst* __restrict__ bc, int logN, long long total, u64* b,
                 const u64* __restrict__ pt) {
    long long i = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    if (i >= total) return;
    const BasisConst k = bc[ew_limb(logN)];
    b[i] = add_mod(b[i], redc128(pt[i], 0, k.q, k.qninv), k.q);
}

void add_plain_mont(const Context& c, int limbs, u64* b, const u64* pt, hipStream_t s) {
    long long total = (long long)limbs * (long long)c.N;
    hipLaunchKernelGGL(k_add_plain_mont, dim3(ew_blocks(total)), dim3(EW_THREADS), 0, s, c.d_bc,
                       c.logN, total, b, pt);
    MI_HIP(hipGetLastError());
}

// dst[brev(d)] = src[brev(((2d+1) g mod 2N - 1) / 2)]
__device__ __forceinline__ u32 brev_n(u32 x, int logN) { return __brev(x) >> (32 - logN); }

__global__ void __launch_bounds__(EW_THREADS)
k_automorphism(int logN, u64 galois, const u64* __restrict__ src, u64* __restrict__ dst, long long sstride,
               long long dstride) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const long long k = blockIdx.y;
    const long long limb_off = (long long)ew_limb(logN) << logN;
    const u32 idx = (u32)(r & (((long long)1 << logN) - 1));
    const u32 mask2n = (2u << logN) - 1;
    const u32 d = brev_n(idx, logN);
    const u32 e = (u32)(((u64)(2 * d + 1) * galois) & mask2n);
    const u32 sidx = brev_n((e - 1) >> 1, logN);
    dst[k * dstride + limb_off + idx] = src[k * sstride + limb_off + sidx];
}

// dst (+)= automorphism(src) over Q limbs 0..limbs-1 (canonical residues)
__global__ void __launch_bounds__(EW_THREADS)
k_automorphism_acc(const BasisConst* __restrict__ bc, int logN, u64 galois, const u64* __restrict__ src,
                   u64* __restrict__ dst, long long sstride, long long dstride, int accumulate) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const long long k = blockIdx.y;
    const int l = ew_limb(logN);
    const long long limb_off = (long long)l << logN;
    const u32 idx = (u32)(r & (((long long)1 << logN) - 1));
    const u32 mask2n = (2u << logN) - 1;
    const u32 d = brev_n(idx, logN);
    const u32 e = (u32)(((u64)(2 * d + 1) * galois) & mask2n);
    const u32 sidx = brev_n((e - 1) >> 1, logN);
    const u64 v = src[k * sstride + limb_off + sidx];
    u64* p = dst + k * dstride + limb_off + idx;
    *p = accumulate ? add_mod(*p, v, bc[l].q) : v;
}

void automorphism_acc(const Context& c, int limbs, u64 galois, const u64* src, u64* dst, int npolys,
                      long long sstride, long long dstride, bool accumulate, hipStream_t s) {
    if ((galois & 1) == 0) throw Error(-1, "Galois element must be odd");
    if (npolys <= 0 || limbs <= 0) return;
    hipLaunchKernelGGL(k_automorphism_acc, ew_grid(c, limbs, npolys), dim3(EW_THREADS), 0, s, c.d_bc, c.logN,
                       galois % (2 * c.N), src, dst, sstride, dstride, accumulate ? 1 : 0);
    MI_HIP(hipGetLastError());
}

void automorphism(const Context& c, int limbs, u64 galois, const u64* src, u64* dst, int npolys,
                  long long sstride, long long dstride, hipStream_t s) {
    if (src == dst) throw Error(-1, "source and destination must differ");
    if ((galois & 1) == 0) throw Error(-1, "Galois element must be odd");
    if (npolys <= 0 || limbs <= 0) return;
    hipLaunchKernelGGL(k_automorphism, ew_grid(c, limbs, npolys), dim3(EW_THREADS), 0, s, c.logN,
                       galois % (2 * c.N), src, dst, sstride, dstride);
    MI_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- rescale
// Polys of one launch are named by a pointer list (up to RS_MAX_POLYS).
constexpr int RS_MAX_POLYS = 64;

// In-place NTT of one limb (basis `basis`) of each listed poly.
struct TopLimbJob {
    u64* p[RS_MAX_POLYS];
    int basis;
    struct View {
        int basis;
        u64* p;
        __device__ u64 load(int i) const { return p[i]; }
        __device__ void store(int i, u64 v) const { p[i] = v; }
    };
    __device__ View view(int k) const { return View{basis, p[k]}; }
    __device__ bool skip(int) const { return false; }
};

// Column pass input: rem_i = ((x_l + floor(p/2)) mod p) mod q_i from the
// coefficient-domain top limb; output to scratch.
struct RescaleInJob {
    const u64* polys[RS_MAX_POLYS];
    u64* scratch;     // [npolys][level][N]
    int level, npolys;
    u64 N;
    u64 p, half;
    const BasisConst* bc;
    FastDiv per_limb;  // npolys
    struct View {
        int basis;
        const u64* top;
        u64* out;
        u64 p, half, q, mu_hi;
        __device__ u64 load(int i) const {
            u64 s = top[i] + half;
            if (s >= p) s -= p;
            return reduce64(s, q, mu_hi);
        }
        __device__ void store(int i, u64 v) const { out[i] = v; }
    };
    __device__ View view(int k) const {
        const int i = per_limb.div(k), poly = k - i * npolys;  // modulus-major
        const BasisConst b = bc[i];
        return View{i, polys[poly] + (long long)level * N, scratch + ((long long)poly * level + i) * N, p, half,
                    b.q, b.mu_hi};
    }
    __device__ bool skip(int) const { return false; }
};

// Row pass output: x_i <- (x_i + NTT(half) - NTT(rem)) * p^-1 mod q_i
struct RescaleOutJob {
    u64* polys[RS_MAX_POLYS];
    const u64* scratch;
    int level, npolys;
    u64 N;
    const u64* half_ntt;
    const u64* inv_p;
    const u64* inv_p_shoup;
    const BasisConst* bc;
    FastDiv per_limb;  // npolys
    struct View {
        int basis;
        const u64* src;
        u64* x;
        const u64* half;
        u64 q, ip, ips;
        __device__ u64 load(int i) const { return src[i]; }
        __device__ u64 pre(int i) const { return add_mod(x[i], half[i], q); }
        __device__ void store_pre(int i, u64 v, u64 num) const { x[i] = shoup(sub_mod(num, v, q), ip, ips, q); }
    };
    __device__ View view(int k) const {
        const int i = per_limb.div(k), poly = k - i * npolys;  // modulus-major
        return View{i, scratch + ((long long)poly * level + i) * N, polys[poly] + (long long)i * N,
                    half_ntt + (long long)i * N, bc[i].q, inv_p[i], inv_p_shoup[i]};
    }
    __device__ bool skip(int) const { return false; }
};

void rescale_list(Context& c, int level, u64* const* polys, int npolys, hipStream_t s) {
    if (level < 1 || level >= c.nq) throw Error(-1, "invalid level");
    if (npolys <= 0) return;
    const u64 N = c.N;
    const u64 p = c.q[level];
    const RescaleTables& rt = c.rescale[level];
    for (int k0 = 0; k0 < npolys; k0 += RS_MAX_POLYS) {
        const int n = std::min(RS_MAX_POLYS, npolys - k0);
        // 1. top limb of every poly to the coefficient domain (in place)
        TopLimbJob top{};
        top.basis = level;
        for (int k = 0; k < n; ++k) top.p[k] = polys[k0 + k] + (long long)level * N;
        run_ntt_inv(c, top, top, n, s);
        // 2. fused remainder -> NTT -> finish
        u64* scratch = (u64*)c.workspace(s, sizeof(u64) * (size_t)n * level * N);
        RescaleInJob in{};
        RescaleOutJob out{};
        for (int k = 0; k < n; ++k) {
            in.polys[k] = polys[k0 + k];
            out.polys[k] = polys[k0 + k];
        }
        in.scratch = scratch;
        in.level = out.level = level;
        in.npolys = out.npolys = n;
        in.N = out.N = N;
        in.p = p;
        in.half = p >> 1;
        in.bc = out.bc = c.d_bc;
        in.per_limb = out.per_limb = fast_div(n);
        out.scratch = scratch;
        out.half_ntt = rt.half_ntt;
        out.inv_p = rt.inv_p;
        out.inv_p_shoup = rt.inv_p_shoup;
        run_ntt_fwd(c, in, out, n * level, s);
    }
}

void rescale(Context& c, int level, u64* polys, int npolys, long long stride, hipStream_t s) {
    if (level < 1 || level >= c.nq) throw Error(-1, "invalid level");
    std::vector<u64*> list(npolys > 0 ? npolys : 0);
    for (int k = 0; k < npolys; ++k) list[k] = polys + (long long)k * stride;
    rescale_list(c, level, list.data(), npolys, s);
}

// x[k] *= pt (limb-wise, Barrett or Montgomery plaintext) for listed polys
struct PolyList {
    u64* p[RS_MAX_POLYS];
};
__global__ void __launch_bounds__(EW_THREADS)
k_mul_plain_list(const BasisConst* __restrict__ bc, PolyList list, const u64* __restrict__ pt, int logN, int mont) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const BasisConst b = bc[ew_limb(logN)];
    u64* px = list.p[blockIdx.y] + r;
    *px = mont ? mont_mul(*px, pt[r], b.q, b.qninv) : mul_mod(*px, pt[r], b.q, b.mu_lo, b.mu_hi);
}

// x[k] = y[k] (op 0), x[k] + y[k] (op 1) or x[k] - y[k] (op 2) mod q for
// listed poly pairs: one launch per RS_MAX_POLYS pairs.
struct PolyPairList {
    u64* x[RS_MAX_POLYS];
    const u64* y[RS_MAX_POLYS];
};
// per-limb scalar residues of ops 3 (x *= r) and 4 (x += y * r)
struct LimbResidues {
    u64 r[MAX_LIMBS];
};
// ops 5 (x += r), 6 (x = 0) and 7 (x += REDC(y): a Montgomery-form plaintext
// added in normal form) complete the list: a ciphertext batch's constant
// additions, zero fills and plaintext additions are one launch each.
template <int OP>
__global__ void __launch_bounds__(EW_THREADS)
k_pair_list(const BasisConst* __restrict__ bc, PolyPairList l, LimbResidues res, int logN) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    u64* px = l.x[blockIdx.y] + r;
    if constexpr (OP == 6) {
        *px = 0;
        return;
    }
    if constexpr (OP == 3 || OP == 5) {
        const int limb = ew_limb(logN);
        const BasisConst b = bc[limb];
        *px = OP == 3 ? mul_mod(*px, res.r[limb], b.q, b.mu_lo, b.mu_hi) : add_mod(*px, res.r[limb], b.q);
        return;
    }
    const u64 v = l.y[blockIdx.y][r];
    if constexpr (OP == 0) {
        *px = v;
    } else if constexpr (OP == 4) {
        const int limb = ew_limb(logN);
        const BasisConst b = bc[limb];
        *px = add_mod(*px, mul_mod(v, res.r[limb], b.q, b.mu_lo, b.mu_hi), b.q);
    } else if constexpr (OP == 7) {
        const BasisConst b = bc[ew_limb(logN)];
        *px = add_mod(*px, redc128(v, 0, b.q, b.qninv), b.q);
    } else {
        const u64 q = bc[ew_limb(logN)].q;
        *px = OP == 1 ? add_mod(*px, v, q) : sub_mod(*px, v, q);
    }
}

void pair_list_op(const Context& c, int op, int level, u64* const* x, const u64* const* y, int npolys,
                  const u64* residues, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid level");
    if (op < 0 || op > 7) throw Error(-1, "invalid list operation");
    const bool uses_y = op <= 2 || op == 4 || op == 7;
    LimbResidues res{};
    if (op == 3 || op == 4 || op == 5)
        for (int i = 0; i <= level; ++i) {
            if (residues[i] >= c.h_bc[i].q) throw Error(-1, "scalar residue not below its modulus");
            res.r[i] = residues[i];
        }
    for (int k0 = 0; k0 < npolys; k0 += RS_MAX_POLYS) {
        const int n = std::min(RS_MAX_POLYS, npolys - k0);
        PolyPairList l{};
        for (int k = 0; k < n; ++k) {
            l.x[k] = x[k0 + k];
            l.y[k] = uses_y ? y[k0 + k] : nullptr;
        }
        const dim3 g = ew_grid(c, level + 1, n);
        const dim3 b(EW_THREADS);
        switch (op) {
            case 0: hipLaunchKernelGGL(k_pair_list<0>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
            case 1: hipLaunchKernelGGL(k_pair_list<1>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
            case 2: hipLaunchKernelGGL(k_pair_list<2>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
            case 3: hipLaunchKernelGGL(k_pair_list<3>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
            case 4: hipLaunchKernelGGL(k_pair_list<4>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
            case 5: hipLaunchKernelGGL(k_pair_list<5>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
            case 6: hipLaunchKernelGGL(k_pair_list<6>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
            default: hipLaunchKernelGGL(k_pair_list<7>, g, b, 0, s, c.d_bc, l, res, c.logN); break;
        }
        MI_HIP(hipGetLastError());
    }
}

// out[k] = sum_t coef[k][t] * in[k][t] mod q for listed output polys (coef 0:
// term unused); LC_ROWS outputs of up to LC_TERMS terms per launch.
constexpr int LC_ROWS = 12;
constexpr int LC_TERMS = 16;
struct LinCombArgs {
    u64* out[LC_ROWS];
    const u64* in[LC_ROWS][LC_TERMS];
    unsigned coef[LC_ROWS][LC_TERMS];
};
__global__ void __launch_bounds__(EW_THREADS)
k_lincomb(const BasisConst* __restrict__ bc, LinCombArgs a, int nterms, int logN) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const int k = blockIdx.y;
    const BasisConst b = bc[ew_limb(logN)];
    u64 acc = 0;
    for (int t = 0; t < nterms; ++t) {
        const unsigned cf = a.coef[k][t];
        if (cf == 0) continue;
        acc = add_mod(acc, mul_mod(a.in[k][t][r], (u64)cf, b.q, b.mu_lo, b.mu_hi), b.q);
    }
    a.out[k][r] = acc;
}

void linear_combine(const Context& c, int level, int nout, u64* const* out, int nterms, const u64* const* in,
                    const unsigned* coef, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid level");
    if (nterms < 1 || nterms > LC_TERMS) throw Error(-1, "1..16 terms per output");
    for (int l = 0; l <= level; ++l)
        for (int i = 0; i < nout * nterms; ++i)
            if ((u64)coef[i] >= c.h_bc[l].q) throw Error(-1, "coefficient not below the modulus");
    for (int k0 = 0; k0 < nout; k0 += LC_ROWS) {
        const int n = std::min(LC_ROWS, nout - k0);
        LinCombArgs a{};
        for (int k = 0; k < n; ++k) {
            a.out[k] = out[k0 + k];
            for (int t = 0; t < nterms; ++t) {
                a.in[k][t] = in[(size_t)(k0 + k) * nterms + t];
                a.coef[k][t] = coef[(size_t)(k0 + k) * nterms + t];
            }
        }
        hipLaunchKernelGGL(k_lincomb, ew_grid(c, level + 1, n), dim3(EW_THREADS), 0, s, c.d_bc, a, nterms, c.logN);
        MI_HIP(hipGetLastError());
    }
}

// out[k] = cst[limb] (k < nconst) + sum_t r[t][limb] * in[t][k] mod q: a
// whole Chebyshev leaf (its constant and its scaled baby-step powers), a
// parity sum or a packing sum in one pass per chunk of polys.  in[t][k]
// null: term t has no poly k.  The sum of up to SS_TERMS canonical products
// is exact in 128 bits (8 q^2 < 2^128 for q < 2^62) and reduced once.
// Kernel arguments stay below 4 KB: a launch takes as many terms as SS_RES
// residue words hold (at most SS_TERMS) and as many polys as SS_PTRS input
// pointers allow for them; later term chunks accumulate into the output.
constexpr int SS_MAXP = 64;
constexpr int SS_PTRS = 128;
constexpr int SS_TERMS = 8;
constexpr int SS_RES = 256;
struct ScaledSumArgs {
    u64* out[SS_MAXP];
    const u64* in[SS_PTRS];  // [term][poly of the launch]
    u64 r[SS_RES];           // [term][limb], then the constant's residues at [nterms][limb]
    int nterms, limbs, nconst, accumulate, np;
};
__global__ void __launch_bounds__(EW_THREADS)
k_scaled_sum(const BasisConst* __restrict__ bc, ScaledSumArgs a, int logN) {
    const long long r = (long long)blockIdx.x * EW_THREADS + threadIdx.x;
    const int k = blockIdx.y;
    const int limb = ew_limb(logN);
    const BasisConst b = bc[limb];
    u64* po = a.out[k] + r;
    U128 acc{0, 0};
    if (a.accumulate) acc.lo = *po;
    else if (k < a.nconst) acc.lo = a.r[a.nterms * a.limbs + limb];
    for (int t = 0; t < a.nterms; ++t) {
        const u64* pi = a.in[t * a.np + k];
        if (pi) mac128(acc, pi[r], a.r[t * a.limbs + limb]);
    }
    *po = reduce128(acc.lo, acc.hi, b.q, b.mu_lo, b.mu_hi);
}

void scaled_sum(const Context& c, int level, int npolys, u64* const* out, int nterms, const u64* const* in,
                const u64* residues, int nconst, const u64* constant, hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid level");
    if (npolys < 1 || nterms < 1 || nterms > 64) throw Error(-1, "1..64 terms over at least one poly");
    if (nconst < 0 || nconst > npolys) throw Error(-1, "constant polys out of range");
    const int limbs = level + 1;
    for (int l = 0; l < limbs; ++l) {
        for (int t = 0; t < nterms; ++t)
            if (residues[(size_t)t * limbs + l] >= c.h_bc[l].q) throw Error(-1, "scalar residue not below its modulus");
        if (nconst && constant[l] >= c.h_bc[l].q) throw Error(-1, "constant residue not below its modulus");
    }
    // terms per launch: the residues of the chunk and the constant fit SS_RES words
    const int per = std::min(SS_TERMS, SS_RES / limbs - 1);
    if (per < 1) throw Error(-1, "too many limbs for a scaled sum");
    for (int t0 = 0; t0 < nterms; t0 += per) {
        const int m = std::min(per, nterms - t0);
        const int np = std::min(SS_MAXP, SS_PTRS / m);
        for (int k0 = 0; k0 < npolys; k0 += np) {
            const int n = std::min(np, npolys - k0);
            ScaledSumArgs a{};
            a.nterms = m;
            a.limbs = limbs;
            a.np = n;
            a.accumulate = t0 > 0;
            a.nconst = std::max(0, std::min(n, nconst - k0));
            for (int k = 0; k < n; ++k) {
                a.out[k] = out[k0 + k];
                for (int t = 0; t < m; ++t) a.in[t * n + k] = in[(size_t)(t0 + t) * npolys + k0 + k];
            }
            for (int t = 0; t < m; ++t)
                for (int l = 0; l < limbs; ++l) a.r[t * limbs + l] = residues[(size_t)(t0 + t) * limbs + l];
            if (a.nconst)
                for (int l = 0; l < limbs; ++l) a.r[m * limbs + l] = constant[l];
            hipLaunchKernelGGL(k_scaled_sum, ew_grid(c, limbs, n), dim3(EW_THREADS), 0, s, c.d_bc, a, c.logN);
            MI_HIP(hipGetLastError());
        }
    }
}

void mul_plain_list(const Context& c, int level, u64* const* polys, int npolys, const u64* pt, bool mont,
                    hipStream_t s) {
    if (level < 0 || level >= c.nq) throw Error(-1, "invalid level");
    for (int k0 = 0; k0 < npolys; k0 += RS_MAX_POLYS) {
        const int n = std::min(RS_MAX_POLYS, npolys - k0);
        PolyList list{};
        for (int k = 0; k < n; ++k) list.p[k] = polys[k0 + k];
        hipLaunchKernelGGL(k_mul_plain_list, ew_grid(c, level + 1, n), dim3(EW_THREADS), 0, s, c.d_bc, list, pt,
                           c.logN, mont ? 1 : 0);
        MI_HIP(hipGetLastError());
    }
}

}  // namespace mi
