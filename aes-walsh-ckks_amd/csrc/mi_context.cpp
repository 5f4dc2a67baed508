#include <cstdlib>
// Context construction: host precompute of the device-resident tables.
//
// Everything here is exact modular arithmetic on 64-bit moduli (128-bit
// intermediates), so each table entry equals the reference's value:
//   twiddles      ntt.cpp:28-148  (psi = first a >= 2 of order 2N)
//   ModUp tables  relinearization_key_generation.cpp:45-129,
//                 ckks_context.cpp:468-535 (digit partition 38-47)
//   folded P^-1   ckks_context.cpp:537-560
//   rescale       ckks_context.cpp:565-599
#include <algorithm>
#include <cstring>
#include <thread>

#include "mi_internal.hpp"

namespace mi {

typedef unsigned __int128 u128;

static u64 mulm(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }
static u64 powm(u64 a, u64 e, u64 m) {
    u64 r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = mulm(r, a, m);
        a = mulm(a, a, m);
        e >>= 1;
    }
    return r;
}
static u64 invm(u64 a, u64 m) {
    __int128 t = 0, nt = 1, r = m, nr = a % m;
    while (nr) {
        __int128 qq = r / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (r != 1) throw Error(-1, "modulus is not invertible");
    if (t < 0) t += m;
    return (u64)t;
}
static u64 shoup_of(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }
static u64 brev_u(u64 x, int bits) {
    u64 r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

// Twiddles of one modulus: table[m + i] = psi^(step*(2*brev_m(i)+1)) with
// step = N/(2m) (forward) and the same with psi^-1 (inverse).
static void make_twiddles(u64 N, int logN, u64 q, u64* fwd, u64* fws, u64* inv, u64* ivs,
                          BasisConst& bc) {
    if ((q - 1) % (2 * N) != 0) throw Error(-1, "modulus is not 1 mod 2N");
    const u64 e = (q - 1) / (2 * N);
    u64 psi = 0;
    for (u64 a = 2; a + 2 <= q; ++a) {
        u64 r = powm(a, e, q);
        if (powm(r, N, q) == q - 1) { psi = r; break; }
    }
    if (!psi) throw Error(-1, "no primitive 2N-th root of unity");
    std::vector<u64> pw(2 * N);
    pw[0] = 1;
    for (u64 i = 1; i < 2 * N; ++i) pw[i] = mulm(pw[i - 1], psi, q);
    fwd[0] = 1;
    inv[0] = 1;
    for (u64 m = 1; m < N; m <<= 1) {
        int sb = 0;
        while ((1ull << sb) < m) ++sb;
        u64 step = N / (2 * m);
        for (u64 i = 0; i < m; ++i) {
            u64 ex = step * (2 * brev_u(i, sb) + 1);
            fwd[m + i] = pw[ex];
            inv[m + i] = pw[(2 * N - ex) % (2 * N)];
        }
    }
    for (u64 i = 0; i < N; ++i) {
        fws[i] = shoup_of(fwd[i], q);
        ivs[i] = shoup_of(inv[i], q);
    }
    u128 mu = (~(u128)0) / q;
    bc.q = q;
    bc.mu_lo = (u64)mu;
    bc.mu_hi = (u64)(mu >> 64);
    u64 inv_q = 1;
    for (int i = 0; i < 6; ++i) inv_q *= 2 - q * inv_q;
    bc.qninv = 0 - inv_q;
    bc.n_inv = invm(N % q, q);
    bc.n_inv_shoup = shoup_of(bc.n_inv, q);
    bc.r_mod_q = (u64)((((u128)1) << 64) % q);
    bc.inv_d = 1.0 / (double)q;
    bc.negq = (u64)0 - q;
    bc.w1_ninv = (u64)(((u128)inv[1] * bc.n_inv) % q);
    bc.w1_ninv_shoup = shoup_of(bc.w1_ninv, q);
    (void)logN;
}

void* Context::workspace(hipStream_t s, size_t bytes) {
    std::lock_guard<std::mutex> g(ws_mu);
    Workspace& w = ws[s];
    if (w.bytes < bytes) {
        // Under stream capture the buffer can be neither synchronised nor
        // reallocated: the captured calls must have run once before.
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        MI_HIP(hipStreamIsCapturing(s, &cap));
        if (cap != hipStreamCaptureStatusNone)
            throw Error(-2, "stream workspace holds " + std::to_string(w.bytes) + " bytes, the call needs " +
                                std::to_string(bytes) +
                                "; run the captured calls once outside the capture first so the workspace is sized");
        const size_t nb = std::max(bytes, w.bytes + w.bytes / 2);
        int graphs = 0;
        {
            std::lock_guard<std::mutex> gp(pool_mu);
            graphs = live_graphs;
        }
        if (w.ptr) {
            if (graphs > 0) {
                retired.push_back(w.ptr);  // a graph may still replay into it
            } else {
                MI_HIP(hipStreamSynchronize(s));
                MI_HIP(hipFree(w.ptr));
            }
            w.ptr = nullptr;
            w.bytes = 0;
        }
        MI_HIP(hipMalloc(&w.ptr, nb));
        w.bytes = nb;
    }
    return w.ptr;
}

int Context::ntt_variant(int count) const {
    if (ntt_forced) return ntt_forced;
    return tune.ntt_variant[tune_bucket(count)];
}

long Context::probe_open(hipStream_t s, long long limbs) const {
    std::lock_guard<std::mutex> g(probe_mu);
    if (!probe.on) return -1;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MI_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return -1;  // graphs are timed as a whole
    while (probe.events.size() < probe.used + 2) {
        hipEvent_t e;
        MI_HIP(hipEventCreate(&e));
        probe.events.push_back(e);
    }
    const long slot = (long)probe.used;
    probe.used += 2;
    probe.bytes += 16.0 * (double)N * (double)limbs;
    probe.limbs += limbs;
    MI_HIP(hipEventRecord(probe.events[slot], s));
    return slot;
}

void Context::probe_close(long slot, hipStream_t s) const {
    if (slot < 0) return;
    std::lock_guard<std::mutex> g(probe_mu);
    MI_HIP(hipEventRecord(probe.events[slot + 1], s));
}

const std::vector<hipStream_t>& Context::side_streams() {
    std::lock_guard<std::mutex> g(side_mu);
    if (side.empty()) {
        for (int i = 0; i < 4; ++i) {
            hipStream_t st = nullptr;
            MI_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            side.push_back(st);
        }
    }
    return side;
}

struct Blob {
    std::vector<unsigned char> host;
    size_t add(const void* p, size_t bytes) {
        size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + bytes);
        if (bytes) std::memcpy(host.data() + off, p, bytes);
        return off;
    }
};

Context* context_create(u64 N, const u64* q, int nq, const u64* p, int np, int dnum) {
    if (N < (1u << 10) || N > (1u << 16) || (N & (N - 1)))
        throw Error(-1, "N must be a power of two in [2^10, 2^16]");
    if (nq < 1 || np < 1 || dnum < 1 || dnum > nq)
        throw Error(-1, "need nq >= 1, np >= 1, 1 <= dnum <= nq");
    for (int i = 0; i < nq + np; ++i) {
        u64 m = i < nq ? q[i] : p[i - nq];
        if (m < 3 || m >= (1ull << 62)) throw Error(-1, "moduli must be < 2^62");
        for (int j = 0; j < i; ++j)
            if ((j < nq ? q[j] : p[j - nq]) == m) throw Error(-1, "moduli must be distinct");
    }
    auto* c = new Context();
    c->N = N;
    while ((1ull << c->logN) < N) ++c->logN;
    c->n1 = c->logN / 2;
    c->n2 = c->logN - c->n1;
    c->nq = nq;
    c->np = np;
    c->dnum = dnum;
    c->alpha = (nq + dnum - 1) / dnum;
    c->q.assign(q, q + nq);
    c->p.assign(p, p + np);
    MI_HIP(hipGetDevice(&c->device));
    noalloc_setup(c->device);
    if (const char* e = std::getenv("MICKKS_FUSE_MAC")) c->fuse_mac = std::string(e) != "0";
    if (const char* e = std::getenv("MICKKS_MULTI_STREAM")) c->multi_stream = std::string(e) != "0";
    if (const char* e = std::getenv("MICKKS_NTT_VARIANT")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= NTT_VARIANTS) c->ntt_forced = v;
    }
    const int nb = nq + np;
    std::vector<u64> fwd(nb * N), fws(nb * N), inv(nb * N), ivs(nb * N);
    c->h_bc.resize(nb);
    {
        std::vector<std::thread> th;
        unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        for (unsigned w = 0; w < hw; ++w)
            th.emplace_back([&, w] {
                for (int b = (int)w; b < nb; b += (int)hw) {
                    u64 m = b < nq ? q[b] : p[b - nq];
                    make_twiddles(N, c->logN, m, &fwd[b * N], &fws[b * N], &inv[b * N],
                                  &ivs[b * N], c->h_bc[b]);
                }
            });
        for (auto& t : th) t.join();
    }
    const size_t tb = sizeof(u64) * nb * N;
    MI_HIP(hipMalloc(&c->d_fwd, tb));
    MI_HIP(hipMalloc(&c->d_fwd_shoup, tb));
    MI_HIP(hipMalloc(&c->d_inv, tb));
    MI_HIP(hipMalloc(&c->d_inv_shoup, tb));
    MI_HIP(hipMalloc(&c->d_bc, sizeof(BasisConst) * nb));
    MI_HIP(hipMemcpy(c->d_fwd, fwd.data(), tb, hipMemcpyHostToDevice));
    MI_HIP(hipMemcpy(c->d_fwd_shoup, fws.data(), tb, hipMemcpyHostToDevice));
    MI_HIP(hipMemcpy(c->d_inv, inv.data(), tb, hipMemcpyHostToDevice));
    MI_HIP(hipMemcpy(c->d_inv_shoup, ivs.data(), tb, hipMemcpyHostToDevice));
    MI_HIP(hipMemcpy(c->d_bc, c->h_bc.data(), sizeof(BasisConst) * nb, hipMemcpyHostToDevice));
    MI_HIP(hipMalloc(&c->d_zero, sizeof(u64) * N));
    MI_HIP(hipMemset(c->d_zero, 0, sizeof(u64) * N));
    c->table_bytes = 4 * tb + sizeof(BasisConst) * nb;

    // ---- ModUp / ModDown / rescale tables packed in one blob ----
    Blob blob;
    struct MU { size_t qi, qis, si, qmd, ac, db; int ns, nd, s0, s1; };
    std::vector<std::vector<MU>> mu(nq);
    for (int level = 0; level < nq; ++level) {
        const int L = level + 1;
        const int parts = c->parts(level);
        for (int part = 0; part < parts; ++part) {
            const int s0 = part * c->alpha, s1 = std::min(s0 + c->alpha, L);
            const int ns = s1 - s0;
            std::vector<u64> dst;
            std::vector<int> dbasis;
            for (int i = 0; i < s0; ++i) { dst.push_back(q[i]); dbasis.push_back(i); }
            for (int i = s1; i < L; ++i) { dst.push_back(q[i]); dbasis.push_back(i); }
            for (int i = 0; i < np; ++i) { dst.push_back(p[i]); dbasis.push_back(nq + i); }
            const int nd = (int)dst.size();
            std::vector<u64> qhi(ns), qhis(ns), qmd((size_t)nd * ns), ac((size_t)nd * (ns + 1));
            std::vector<double> sinv(ns);
            for (int i = 0; i < ns; ++i) {
                const u64 qi = q[s0 + i];
                sinv[i] = 1.0 / (double)qi;
                u64 h = 1;
                for (int j = 0; j < ns; ++j)
                    if (j != i) h = mulm(h, q[s0 + j] % qi, qi);
                qhi[i] = invm(h, qi);
                qhis[i] = shoup_of(qhi[i], qi);
                for (int t = 0; t < nd; ++t) {
                    u64 v = 1;
                    for (int j = 0; j < ns; ++j)
                        if (j != i) v = mulm(v, q[s0 + j] % dst[t], dst[t]);
                    // Montgomery form (x 2^64): the kernel's sum ends in one REDC
                    qmd[(size_t)t * ns + i] = (u64)(((unsigned __int128)v << 64) % dst[t]);
                }
            }
            for (int t = 0; t < nd; ++t) {
                u64 prod = 1;
                for (int i = 0; i < ns; ++i) prod = mulm(prod, q[s0 + i] % dst[t], dst[t]);
                for (int a = 0; a <= ns; ++a) ac[(size_t)t * (ns + 1) + a] = mulm(prod, (u64)a, dst[t]);
            }
            MU m;
            m.ns = ns; m.nd = nd; m.s0 = s0; m.s1 = s1;
            m.qi = blob.add(qhi.data(), 8 * ns);
            m.qis = blob.add(qhis.data(), 8 * ns);
            m.si = blob.add(sinv.data(), 8 * ns);
            m.qmd = blob.add(qmd.data(), 8 * qmd.size());
            m.ac = blob.add(ac.data(), 8 * ac.size());
            m.db = blob.add(dbasis.data(), sizeof(int) * nd);
            mu[level].push_back(m);
        }
    }
    std::vector<size_t> md_off(nq);
    for (int level = 0; level < nq; ++level) {
        const int L = level + 1;
        std::vector<u64> pinv((size_t)L * np);
        for (int i = 0; i < L; ++i)
            for (int j = 0; j < np; ++j)  // Montgomery form, as above
                pinv[(size_t)i * np + j] = (u64)(((unsigned __int128)invm(p[j] % q[i], q[i]) << 64) % q[i]);
        md_off[level] = blob.add(pinv.data(), 8 * pinv.size());
    }
    struct RS { size_t ip, ips, hm; };
    std::vector<RS> rs(nq);
    for (int d = 1; d < nq; ++d) {
        std::vector<u64> ip(d), ips(d), hm(d);
        const u64 pp = q[d], half = pp >> 1;
        for (int i = 0; i < d; ++i) {
            ip[i] = invm(pp % q[i], q[i]);
            ips[i] = shoup_of(ip[i], q[i]);
            hm[i] = half % q[i];
        }
        rs[d].ip = blob.add(ip.data(), 8 * d);
        rs[d].ips = blob.add(ips.data(), 8 * d);
        rs[d].hm = blob.add(hm.data(), 8 * d);
    }
    c->blob_bytes = blob.host.size();
    MI_HIP(hipMalloc(&c->d_blob, std::max<size_t>(c->blob_bytes, 256)));
    MI_HIP(hipMemcpy(c->d_blob, blob.host.data(), c->blob_bytes, hipMemcpyHostToDevice));
    c->table_bytes += c->blob_bytes;
    auto P = [&](size_t off) { return (const unsigned char*)c->d_blob + off; };
    c->modup.resize(nq);
    for (int level = 0; level < nq; ++level)
        for (const MU& m : mu[level]) {
            ModUpTables t;
            t.ns = m.ns; t.nd = m.nd; t.s0 = m.s0; t.s1 = m.s1;
            t.qhat_inv = (const u64*)P(m.qi);
            t.qhat_inv_shoup = (const u64*)P(m.qis);
            t.src_inv = (const double*)P(m.si);
            t.qhat_mod_dst = (const u64*)P(m.qmd);
            t.alpha_corr = (const u64*)P(m.ac);
            t.dst_basis = (const int*)P(m.db);
            c->modup[level].push_back(t);
        }
    c->moddown.resize(nq);
    for (int level = 0; level < nq; ++level) c->moddown[level].pinv = (const u64*)P(md_off[level]);

    // Rescale: NTT of the polynomial whose every coefficient is floor(q_d/2) mod q_i.
    c->rescale.resize(nq);
    size_t half_words = 0;
    for (int d = 1; d < nq; ++d) half_words += (size_t)d * N;
    u64* d_half = nullptr;
    if (half_words) {
        MI_HIP(hipMalloc(&d_half, 8 * half_words));
        c->table_bytes += 8 * half_words;
        std::vector<u64> h(half_words);
        size_t off = 0;
        for (int d = 1; d < nq; ++d) {
            const u64 half = q[d] >> 1;
            for (int i = 0; i < d; ++i)
                std::fill(h.begin() + off + (size_t)i * N, h.begin() + off + (size_t)(i + 1) * N, half % q[i]);
            off += (size_t)d * N;
        }
        MI_HIP(hipMemcpy(d_half, h.data(), 8 * half_words, hipMemcpyHostToDevice));
        off = 0;
        for (int d = 1; d < nq; ++d) {
            ntt_forward(*c, LimbBatch{d_half + off, (long long)d * (long long)N, d, 0, d}, nullptr);
            c->rescale[d].half_ntt = d_half + off;
            c->rescale[d].inv_p = (const u64*)P(rs[d].ip);
            c->rescale[d].inv_p_shoup = (const u64*)P(rs[d].ips);
            c->rescale[d].half_mod = (const u64*)P(rs[d].hm);
            off += (size_t)d * N;
        }
        MI_HIP(hipDeviceSynchronize());
    }
    c->d_half = d_half;
    // level-0 dense-to-sparse switch: p_0^-1 mod q_0 (ckks_keyswitch.cpp:1767-1768)
    c->q0p0_inv = invm(p[0] % q[0], q[0]);
    c->q0p0_inv_shoup = shoup_of(c->q0p0_inv, q[0]);
    try {
        tune_context(*c);
    } catch (const Error&) {
        // tuning could not run (e.g. no memory for its scratch): the defaults
        c->tune = TunePlan{};
    }
    return c;
}

// Pool blocks are rounded to 4 KiB; a cached block is reused for a request of
// at least half its size.  Frees are not stream-ordered: the caller frees a
// block only after the work using it is ordered before any later use (one
// stream, or a synchronisation).
void* Context::pool_alloc(size_t bytes) {
    const size_t rounded = (std::max<size_t>(bytes, 1) + 4095) & ~size_t(4095);
    std::lock_guard<std::mutex> g(pool_mu);
    if (pool_limit && pool_live_bytes + rounded > pool_limit)
        throw Error(-2, "device memory limit reached (" + std::to_string(pool_live_bytes) + " of " +
                            std::to_string(pool_limit) + " bytes live)");
    void* p = nullptr;
    size_t got = rounded;
    auto it = pool_free.lower_bound(rounded);
    if (it != pool_free.end() && it->first <= 2 * rounded) {
        p = it->second;
        got = it->first;
        pool_cached_bytes -= got;
        pool_free.erase(it);
    } else {
        hipError_t e = hipMalloc(&p, rounded);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            for (auto& kv : pool_free) (void)hipFree(kv.second);
            pool_free.clear();
            pool_cached_bytes = 0;
            MI_HIP(hipMalloc(&p, rounded));
        }
    }
    pool_live[p] = got;
    if (capturing > 0) captured_blocks.insert(p);
    pool_live_bytes += got;
    pool_peak_bytes = std::max(pool_peak_bytes, pool_live_bytes);
    return p;
}

void Context::pool_release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(pool_mu);
    auto it = pool_live.find(p);
    if (it == pool_live.end()) throw Error(-1, "pointer was not allocated by this context");
    pool_live_bytes -= it->second;
    if (capturing > 0 || captured_blocks.count(p)) {
        graph_held.emplace_back(it->second, p);
    } else {
        pool_cached_bytes += it->second;
        pool_free.emplace(it->second, p);
    }
    pool_live.erase(it);
}

void Context::capture_begin() {
    std::lock_guard<std::mutex> g(pool_mu);
    ++capturing;
}

void Context::capture_end(bool made_graph) {
    std::lock_guard<std::mutex> g(pool_mu);
    --capturing;
    if (made_graph) ++live_graphs;
    if (capturing > 0 || live_graphs > 0) return;
    for (auto& h : graph_held) {
        pool_cached_bytes += h.first;
        pool_free.emplace(h.first, h.second);
    }
    graph_held.clear();
    captured_blocks.clear();
}

void Context::graph_destroyed() {
    {
        std::lock_guard<std::mutex> g(pool_mu);
        --live_graphs;
        if (capturing > 0 || live_graphs > 0) return;
        for (auto& h : graph_held) {
            pool_cached_bytes += h.first;
            pool_free.emplace(h.first, h.second);
        }
        graph_held.clear();
        captured_blocks.clear();
    }
    std::lock_guard<std::mutex> g(ws_mu);
    for (void* r : retired) (void)hipFree(r);
    retired.clear();
}

void Context::pool_trim() {
    std::lock_guard<std::mutex> g(pool_mu);
    for (auto& kv : pool_free) (void)hipFree(kv.second);
    pool_free.clear();
    pool_cached_bytes = 0;
}

void context_destroy(Context* c) {
    if (!c) return;
    (void)hipDeviceSynchronize();
    for (auto& kv : c->pool_free) (void)hipFree(kv.second);
    for (auto& kv : c->pool_live) (void)hipFree(kv.first);
    for (auto& kv : c->graph_held) (void)hipFree(kv.second);
    (void)hipFree(c->d_fwd);
    (void)hipFree(c->d_fwd_shoup);
    (void)hipFree(c->d_inv);
    (void)hipFree(c->d_inv_shoup);
    (void)hipFree(c->d_bc);
    (void)hipFree(c->d_zero);
    (void)hipFree(c->d_blob);
    (void)hipFree(c->d_half);
    (void)hipFree(c->d_enc_tw);
    (void)hipFree(c->d_enc_inv);
    (void)hipFree(c->d_dec_tw);
    (void)hipFree(c->d_dec_perm);
    (void)hipFree(c->d_gauss_cdf);
    for (auto& kv : c->ws)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (void* r : c->retired) (void)hipFree(r);
    for (hipStream_t st : c->side) (void)hipStreamDestroy(st);
    for (hipEvent_t e : c->probe.events) (void)hipEventDestroy(e);
    delete c;
}

}  // namespace mi
