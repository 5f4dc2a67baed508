#include <algorithm>
// extern "C" entry points of libmickks.so (declared in include/mickks.h).
#include <cstring>
#include <string>

#include "../../include/mickks.h"
#include "mi_internal.hpp"

#include <vector>

struct mi_context {
    mi::Context* c;
};

struct mi_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {
thread_local std::string g_err;

int fail(const char* fn, int code, const std::string& msg) {
    g_err = std::string(fn) + ": " + msg;
    return code;
}

template <class F>
int guard(const char* fn, F&& f) {
    try {
        f();
        return MI_OK;
    } catch (const mi::Error& e) {
        return fail(fn, e.code, e.what());
    } catch (const std::exception& e) {
        return fail(fn, MI_ERUNTIME, e.what());
    }
}

inline hipStream_t S(void* s) { return (hipStream_t)s; }

void check_ctx(const mi_context* ctx) {
    if (!ctx || !ctx->c) throw mi::Error(MI_EINVAL, "null context");
}
void check_level(const mi::Context& c, int level) {
    if (level < 0 || level >= c.nq) throw mi::Error(MI_EINVAL, "invalid level");
}
void check_ptr(const void* p) {
    if (!p) throw mi::Error(MI_EINVAL, "null buffer");
}
void check_galois(uint64_t g, bool zero_ok) {
    if (g == 0 && zero_ok) return;
    if ((g & 1) == 0) throw mi::Error(MI_EINVAL, "Galois element must be odd");
}
}  // namespace

extern "C" {

const char* mi_last_error(void) { return g_err.c_str(); }

int mi_context_create(mi_context** out, uint64_t N, const uint64_t* q, int nq, const uint64_t* p,
                      int np, int dnum) {
    return guard("mi_context_create", [&] {
        if (!out || !q || !p) throw mi::Error(MI_EINVAL, "null argument");
        *out = nullptr;
        mi::Context* c = mi::context_create(N, q, nq, p, np, dnum);
        *out = new mi_context{c};
    });
}

int mi_context_destroy(mi_context* ctx) {
    return guard("mi_context_destroy", [&] {
        if (!ctx) return;
        mi::context_destroy(ctx->c);
        delete ctx;
    });
}

int mi_context_info(const mi_context* ctx, int64_t* info) {
    return guard("mi_context_info", [&] {
        check_ctx(ctx);
        const mi::Context& c = *ctx->c;
        info[0] = (int64_t)c.N;
        info[1] = c.nq;
        info[2] = c.np;
        info[3] = c.dnum;
        info[4] = c.alpha;
        info[5] = c.device;
    });
}

int64_t mi_context_device_bytes(const mi_context* ctx) {
    if (!ctx || !ctx->c) return -1;
    int64_t b = (int64_t)ctx->c->table_bytes;
    for (auto& kv : ctx->c->ws) b += (int64_t)kv.second.bytes;
    return b;
}

int mi_device_memory_info(mi_context* ctx, int64_t* free_bytes, int64_t* total_bytes) {
    return guard("mi_device_memory_info", [&] {
        check_ctx(ctx);
        check_ptr(free_bytes);
        check_ptr(total_bytes);
        int dev = -1;
        MI_HIP(hipGetDevice(&dev));
        if (dev != ctx->c->device) MI_HIP(hipSetDevice(ctx->c->device));
        size_t f = 0, t = 0;
        MI_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = (int64_t)f;
        *total_bytes = (int64_t)t;
    });
}

int mi_probe_begin(mi_context* ctx) {
    return guard("mi_probe_begin", [&] {
        check_ctx(ctx);
        mi::Context& c = *ctx->c;
        std::lock_guard<std::mutex> g(c.probe_mu);
        if (!c.probe.on) c.probe.saved_multi_stream = c.multi_stream;
        c.multi_stream = false;
        c.probe.on = true;
        c.probe.used = 0;
        c.probe.bytes = 0.0;
        c.probe.limbs = 0;
    });
}

int mi_probe_end(mi_context* ctx, double* stats) {
    return guard("mi_probe_end", [&] {
        check_ctx(ctx);
        check_ptr(stats);
        mi::Context& c = *ctx->c;
        std::lock_guard<std::mutex> g(c.probe_mu);
        if (c.probe.on) c.multi_stream = c.probe.saved_multi_stream;
        c.probe.on = false;
        double ms = 0.0;
        for (size_t i = 0; i + 1 < c.probe.used; i += 2) {
            MI_HIP(hipEventSynchronize(c.probe.events[i + 1]));
            float t = 0.0f;
            MI_HIP(hipEventElapsedTime(&t, c.probe.events[i], c.probe.events[i + 1]));
            ms += t;
        }
        stats[0] = ms;
        stats[1] = (double)(c.probe.used / 2);
        stats[2] = c.probe.bytes;
        stats[3] = (double)c.probe.limbs;
    });
}

int mi_tune_report(mi_context* ctx, char* buf, size_t len) {
    return guard("mi_tune_report", [&] {
        check_ctx(ctx);
        check_ptr(buf);
        const mi::TunePlan& t = ctx->c->tune;
        std::string r = "{\"device\": \"" + ctx->c->tune_key + "\", \"tuned\": " + (t.tuned ? "true" : "false") +
                        ", \"from_cache\": " + (t.from_cache ? "true" : "false") + ", \"ntt_variant\": [";
        for (int b = 0; b < mi::TUNE_BUCKETS; ++b) r += (b ? ", " : "") + std::to_string(t.ntt_variant[b]);
        r += "], \"ntt_bucket_limbs\": [0, 4, 16, 64, 256, 1024], \"keyswitch_group\": " +
             std::to_string(t.ks_group) + ", \"ntt_forced\": " + std::to_string(ctx->c->ntt_forced) + "}";
        if (r.size() + 1 > len) throw mi::Error(MI_EINVAL, "buffer too small");
        std::memcpy(buf, r.c_str(), r.size() + 1);
    });
}

int mi_tune_set(mi_context* ctx, const int* ntt_variant, int keyswitch_group) {
    return guard("mi_tune_set", [&] {
        check_ctx(ctx);
        check_ptr(ntt_variant);
        mi::TunePlan t = ctx->c->tune;
        for (int b = 0; b < mi::TUNE_BUCKETS; ++b) {
            if (ntt_variant[b] < 1 || ntt_variant[b] > mi::NTT_VARIANTS)
                throw mi::Error(MI_EINVAL, "NTT variant must be 1, 2 or 3");
            t.ntt_variant[b] = ntt_variant[b];
        }
        if (keyswitch_group != 8 && keyswitch_group != 16 && keyswitch_group != 32)
            throw mi::Error(MI_EINVAL, "key-switch group must be 8, 16 or 32");
        t.ks_group = keyswitch_group;
        ctx->c->tune = t;
    });
}

int mi_alloc(mi_context* ctx, size_t bytes, void** out) {
    return guard("mi_alloc", [&] {
        check_ctx(ctx);
        check_ptr(out);
        *out = ctx->c->pool_alloc(bytes);
    });
}

int mi_free(mi_context* ctx, void* ptr) {
    return guard("mi_free", [&] {
        check_ctx(ctx);
        ctx->c->pool_release(ptr);
    });
}

int mi_memory_trim(mi_context* ctx) {
    return guard("mi_memory_trim", [&] {
        check_ctx(ctx);
        ctx->c->pool_trim();
    });
}

int mi_memory_stats(mi_context* ctx, int64_t* stats) {
    return guard("mi_memory_stats", [&] {
        check_ctx(ctx);
        check_ptr(stats);
        mi::Context& c = *ctx->c;
        std::lock_guard<std::mutex> g(c.pool_mu);
        stats[0] = (int64_t)c.pool_live_bytes;
        stats[1] = (int64_t)c.pool_peak_bytes;
        stats[2] = (int64_t)c.pool_cached_bytes;
        stats[3] = (int64_t)c.pool_limit;
    });
}

int mi_set_memory_limit(mi_context* ctx, int64_t bytes) {
    return guard("mi_set_memory_limit", [&] {
        check_ctx(ctx);
        if (bytes < 0) throw mi::Error(MI_EINVAL, "negative limit");
        std::lock_guard<std::mutex> g(ctx->c->pool_mu);
        ctx->c->pool_limit = (size_t)bytes;
    });
}

int mi_memcpy(mi_context* ctx, void* dst, const void* src, size_t bytes, int kind, void* stream) {
    return guard("mi_memcpy", [&] {
        check_ctx(ctx);
        if (bytes == 0) return;
        check_ptr(dst); check_ptr(src);
        hipMemcpyKind k;
        switch (kind) {
            case MI_COPY_H2D: k = hipMemcpyHostToDevice; break;
            case MI_COPY_D2H: k = hipMemcpyDeviceToHost; break;
            case MI_COPY_D2D: k = hipMemcpyDeviceToDevice; break;
            default: throw mi::Error(MI_EINVAL, "invalid copy kind");
        }
        MI_HIP(hipMemcpyAsync(dst, src, bytes, k, S(stream)));
        if (kind != MI_COPY_D2D) MI_HIP(hipStreamSynchronize(S(stream)));
    });
}

int mi_stream_create(mi_context* ctx, void** out) {
    return guard("mi_stream_create", [&] {
        check_ctx(ctx);
        if (!out) throw mi::Error(-1, "null output");
        hipStream_t st = nullptr;
        MI_HIP(hipStreamCreate(&st));  // default priority, synchronises with the null stream
        *out = st;
    });
}

int mi_stream_destroy(mi_context* ctx, void* stream) {
    return guard("mi_stream_destroy", [&] {
        check_ctx(ctx);
        if (!stream) throw mi::Error(-1, "null stream");
        MI_HIP(hipStreamSynchronize(S(stream)));
        MI_HIP(hipStreamDestroy(S(stream)));
    });
}

int mi_stream_synchronize(mi_context* ctx, void* stream) {
    return guard("mi_stream_synchronize", [&] {
        check_ctx(ctx);
        MI_HIP(hipStreamSynchronize(S(stream)));
    });
}

int mi_poly_op(mi_context* ctx, int op, uint64_t* x, const uint64_t* y, int limbs, int basis0,
               int npolys, int64_t poly_stride, void* stream) {
    return guard("mi_poly_op", [&] {
        check_ctx(ctx);
        check_ptr(x); check_ptr(y);
        const mi::Context& c = *ctx->c;
        if (limbs < 1 || basis0 < 0 || basis0 + limbs > c.nq + c.np)
            throw mi::Error(MI_EINVAL, "invalid limb range");
        if (npolys < 1 || poly_stride < (int64_t)limbs * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid poly count or stride");
        if (op < MI_OP_ADD || op > MI_OP_MUL_MONT) throw mi::Error(MI_EINVAL, "invalid op");
        mi::ew_binary(c, op, x, y, limbs, npolys, poly_stride, poly_stride, S(stream), basis0);
    });
}

int mi_poly_scalar(mi_context* ctx, int op, uint64_t* x, const uint64_t* residues, int limbs,
                   int basis0, int npolys, int64_t poly_stride, void* stream) {
    return guard("mi_poly_scalar", [&] {
        check_ctx(ctx);
        check_ptr(x); check_ptr(residues);
        const mi::Context& c = *ctx->c;
        if (limbs < 1 || basis0 < 0 || basis0 + limbs > c.nq + c.np)
            throw mi::Error(MI_EINVAL, "invalid limb range");
        if (npolys < 1 || poly_stride < (int64_t)limbs * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid poly count or stride");
        if (op != MI_OP_ADD && op != MI_OP_MUL) throw mi::Error(MI_EINVAL, "invalid op");
        mi::ew_scalar(c, op, x, residues, limbs, npolys, poly_stride, S(stream), basis0);
    });
}

int mi_ntt_forward(mi_context* ctx, uint64_t* data, int npolys, int limbs, int64_t poly_stride,
                   int basis0, void* stream) {
    return guard("mi_ntt_forward", [&] {
        check_ctx(ctx);
        check_ptr(data);
        const mi::Context& c = *ctx->c;
        if (npolys < 0 || limbs < 1 || basis0 < 0 || basis0 + limbs > c.nb() ||
            poly_stride < (int64_t)limbs * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid batch shape");
        mi::ntt_forward(c, mi::LimbBatch{data, poly_stride, limbs, basis0, npolys * limbs}, S(stream));
    });
}

int mi_ntt_inverse(mi_context* ctx, uint64_t* data, int npolys, int limbs, int64_t poly_stride,
                   int basis0, void* stream) {
    return guard("mi_ntt_inverse", [&] {
        check_ctx(ctx);
        check_ptr(data);
        const mi::Context& c = *ctx->c;
        if (npolys < 0 || limbs < 1 || basis0 < 0 || basis0 + limbs > c.nb() ||
            poly_stride < (int64_t)limbs * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid batch shape");
        mi::ntt_inverse(c, mi::LimbBatch{data, poly_stride, limbs, basis0, npolys * limbs}, S(stream));
    });
}

int mi_tensor(mi_context* ctx, int level, uint64_t* b1, uint64_t* a1, const uint64_t* b2,
              const uint64_t* a2, uint64_t* c2, void* stream) {
    return guard("ckks_mul_inplace(ct,ct)", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(b1); check_ptr(a1); check_ptr(b2); check_ptr(a2); check_ptr(c2);
        mi::tensor(*ctx->c, level + 1, b1, a1, b2, a2, c2, S(stream));
    });
}

static void check_disjoint(std::vector<const uint64_t*> lo, long long W, const char* what);

int mi_tensor_list(mi_context* ctx, int level, int count, uint64_t* const* b1, uint64_t* const* a1,
                   const uint64_t* const* b2, const uint64_t* const* a2, uint64_t* const* c2, void* stream) {
    return guard(b2 ? "ckks_mul_inplace(ct,ct)" : "ckks_square_inplace", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (count < 1 || !b1 || !a1 || !c2 || (!b2 != !a2)) throw mi::Error(MI_EINVAL, "invalid ciphertext list");
        const long long W = (long long)(level + 1) * (long long)ctx->c->N;
        for (int k = 0; k < count; ++k) {
            check_ptr(b1[k]); check_ptr(a1[k]); check_ptr(c2[k]);
            if (b2) {
                check_ptr(b2[k]);
                check_ptr(a2[k]);
            }
            // outputs are written in place: no two outputs of the list may overlap
            const uint64_t* outs[3] = {b1[k], a1[k], c2[k]};
            for (int j = 0; j <= k; ++j) {
                const uint64_t* other[3] = {b1[j], a1[j], c2[j]};
                for (int u = 0; u < 3; ++u)
                    for (int v = 0; v < 3; ++v) {
                        if (j == k && v <= u) continue;
                        if (outs[u] < other[v] + W && other[v] < outs[u] + W)
                            throw mi::Error(MI_EINVAL, "output polys overlap");
                    }
            }
        }
        mi::tensor_list(*ctx->c, level, count, b1, a1, b2, a2, c2, S(stream));
    });
}

int mi_tensor_list_into(mi_context* ctx, int level, int count, const uint64_t* const* b1, const uint64_t* const* a1,
                        const uint64_t* const* b2, const uint64_t* const* a2, uint64_t* const* b_out,
                        uint64_t* const* a_out, uint64_t* const* c2, void* stream) {
    return guard(b2 ? "ckks_mul(ct,ct)" : "ckks_square", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (count < 1 || !b1 || !a1 || !b_out || !a_out || !c2 || (!b2 != !a2))
            throw mi::Error(MI_EINVAL, "invalid ciphertext list");
        const long long W = (long long)(level + 1) * (long long)ctx->c->N;
        std::vector<const uint64_t*> outs, ins;
        for (int k = 0; k < count; ++k) {
            check_ptr(b1[k]); check_ptr(a1[k]);
            check_ptr(b_out[k]); check_ptr(a_out[k]); check_ptr(c2[k]);
            outs.insert(outs.end(), {b_out[k], a_out[k], c2[k]});
            ins.insert(ins.end(), {b1[k], a1[k]});
            if (b2) {
                check_ptr(b2[k]);
                check_ptr(a2[k]);
                ins.insert(ins.end(), {b2[k], a2[k]});
            }
        }
        check_disjoint(outs, W, "output polys overlap");
        for (const uint64_t* o : outs)
            for (const uint64_t* x : ins)
                if (o < x + W && x < o + W) throw mi::Error(MI_EINVAL, "an output overlaps an input");
        mi::tensor_list(*ctx->c, level, count, b_out, a_out, b2, a2, c2, S(stream), b1, a1);
    });
}

int mi_square(mi_context* ctx, int level, uint64_t* b, uint64_t* a, uint64_t* c2, void* stream) {
    return guard("ckks_square_inplace", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(b); check_ptr(a); check_ptr(c2);
        mi::square(*ctx->c, level + 1, b, a, c2, S(stream));
    });
}

int mi_mul_plain(mi_context* ctx, int level, uint64_t* polys, int npolys, int64_t poly_stride,
                 const uint64_t* pt, int pt_montgomery, void* stream) {
    return guard("ckks_mul_inplace(ct,pt)", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(polys); check_ptr(pt);
        const mi::Context& c = *ctx->c;
        if (npolys < 1 || poly_stride < (int64_t)(level + 1) * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid ciphertext shape");
        mi::ew_binary(c, pt_montgomery ? mi::EW_MUL_MONT : mi::EW_MUL, polys, pt, level + 1, npolys,
                      poly_stride, 0, S(stream));
    });
}

int mi_mul_scalar(mi_context* ctx, int level, uint64_t* polys, int npolys, int64_t poly_stride,
                  const uint64_t* residues, void* stream) {
    return guard("ckks_mul_inplace(ct,int)", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(polys); check_ptr(residues);
        const mi::Context& c = *ctx->c;
        if (npolys < 1 || poly_stride < (int64_t)(level + 1) * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid ciphertext shape");
        mi::ew_scalar(c, mi::EW_MUL, polys, residues, level + 1, npolys, poly_stride, S(stream));
    });
}

int mi_mul_scalar_add(mi_context* ctx, int level, uint64_t* dst, const uint64_t* src, int npolys,
                      int64_t poly_stride, const uint64_t* residues, void* stream) {
    return guard("fused_multiply_by_real_scalar_then_add", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(dst); check_ptr(src); check_ptr(residues);
        const mi::Context& c = *ctx->c;
        if (npolys < 1 || poly_stride < (int64_t)(level + 1) * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid ciphertext shape");
        mi::ew_scalar_mac(c, dst, src, residues, level + 1, npolys, poly_stride, poly_stride, S(stream));
    });
}

int mi_memset_zero(mi_context* ctx, void* ptr, size_t bytes, void* stream) {
    return guard("mi_memset_zero", [&] {
        check_ctx(ctx);
        if (bytes == 0) return;
        check_ptr(ptr);
        mi::fill_zero(ptr, bytes, S(stream));
    });
}

int mi_add(mi_context* ctx, int level, uint64_t* x, const uint64_t* y, int npolys,
           int64_t poly_stride, void* stream) {
    return guard("ckks_add_inplace(ct,ct)", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(x); check_ptr(y);
        const mi::Context& c = *ctx->c;
        if (npolys < 1 || poly_stride < (int64_t)(level + 1) * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid ciphertext shape");
        mi::ew_binary(c, mi::EW_ADD, x, y, level + 1, npolys, poly_stride, poly_stride, S(stream));
    });
}

int mi_sub(mi_context* ctx, int level, uint64_t* x, const uint64_t* y, int npolys,
           int64_t poly_stride, void* stream) {
    return guard("ckks_sub_inplace(ct,ct)", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(x); check_ptr(y);
        const mi::Context& c = *ctx->c;
        if (npolys < 1 || poly_stride < (int64_t)(level + 1) * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid ciphertext shape");
        mi::ew_binary(c, mi::EW_SUB, x, y, level + 1, npolys, poly_stride, poly_stride, S(stream));
    });
}

int mi_add_plain(mi_context* ctx, int level, uint64_t* b, const uint64_t* pt, int pt_montgomery,
                 void* stream) {
    return guard("ckks_add_inplace(ct,pt)", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(b); check_ptr(pt);
        const mi::Context& c = *ctx->c;
        if (pt_montgomery)
            mi::add_plain_mont(c, level + 1, b, pt, S(stream));
        else
            mi::ew_binary(c, mi::EW_ADD, b, pt, level + 1, 1, 0, 0, S(stream));
    });
}

int mi_add_scalar(mi_context* ctx, int level, uint64_t* b, const uint64_t* residues, void* stream) {
    return guard("ckks_add_integer_constant_inplace", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(b); check_ptr(residues);
        mi::ew_scalar(*ctx->c, mi::EW_ADD, b, residues, level + 1, 1, 0, S(stream));
    });
}

int mi_automorphism(mi_context* ctx, int limbs, uint64_t galois, const uint64_t* src, uint64_t* dst,
                    int npolys, int64_t poly_stride, void* stream) {
    return guard("ckks_apply_ntt_automorphism", [&] {
        check_ctx(ctx);
        check_ptr(src); check_ptr(dst);
        const mi::Context& c = *ctx->c;
        if (limbs < 1 || limbs > c.nb() || npolys < 1 || poly_stride < (int64_t)limbs * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid batch shape");
        mi::automorphism(c, limbs, galois, src, dst, npolys, poly_stride, poly_stride, S(stream));
    });
}

int mi_rescale(mi_context* ctx, int level, uint64_t* polys, int npolys, int64_t poly_stride,
               void* stream) {
    return guard("ckks_rescale", [&] {
        check_ctx(ctx);
        check_ptr(polys);
        mi::Context& c = *ctx->c;
        if (level < 1 || level >= c.nq) throw mi::Error(MI_EINVAL, "invalid level");
        if (npolys < 1 || poly_stride < (int64_t)(level + 1) * (int64_t)c.N)
            throw mi::Error(MI_EINVAL, "invalid ciphertext buffer size");
        mi::rescale(c, level, polys, npolys, poly_stride, S(stream));
    });
}

int mi_keyswitch(mi_context* ctx, int level, const uint64_t* key_term, const uint64_t* key,
                 uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, int montgomery,
                 void* stream) {
    return guard("ckks_hybrid_keyswitch", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(key_term); check_ptr(key); check_ptr(b_out); check_ptr(a_out);
        if ((b_mode != 0 && b_mode != 1) || (a_mode != 0 && a_mode != 1))
            throw mi::Error(MI_EINVAL, "invalid combine mode");
        mi::keyswitch(*ctx->c, mi::KsItem{key_term, key, b_out, a_out, level, b_mode, a_mode},
                      montgomery != 0, S(stream));
    });
}

int mi_keyswitch_noalloc(mi_context* ctx, int level, const uint64_t* key_term, const uint64_t* key,
                         uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, int montgomery,
                         void* stream) {
    return guard("mi_keyswitch_noalloc", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(key_term); check_ptr(key); check_ptr(b_out); check_ptr(a_out);
        if ((b_mode != 0 && b_mode != 1) || (a_mode != 0 && a_mode != 1))
            throw mi::Error(MI_EINVAL, "invalid combine mode");
        mi::keyswitch_noalloc(*ctx->c, mi::KsItem{key_term, key, b_out, a_out, level, b_mode, a_mode},
                              montgomery != 0, S(stream));
    });
}

int mi_keyswitch_batch(mi_context* ctx, const mi_keyswitch_item* items, int count, int montgomery,
                       void* stream) {
    return guard("ckks_hybrid_keyswitch", [&] {
        check_ctx(ctx);
        if (count < 0 || (count > 0 && !items)) throw mi::Error(MI_EINVAL, "invalid batch");
        for (int i = 0; i < count; ++i) {
            const mi_keyswitch_item& it = items[i];
            check_level(*ctx->c, it.level);
            check_ptr(it.key_term); check_ptr(it.key); check_ptr(it.b_out); check_ptr(it.a_out);
        }
        std::vector<mi::KsItem> v;
        v.reserve(count);
        for (int i = 0; i < count; ++i) {
            const mi_keyswitch_item& it = items[i];
            if (it.b_mode < 0 || it.b_mode > 1 || it.a_mode < 0 || it.a_mode > 1)
                throw mi::Error(MI_EINVAL, "combine mode must be 0 (Add) or 1 (Assign)");
            v.push_back(mi::KsItem{it.key_term, it.key, it.b_out, it.a_out, it.level, it.b_mode, it.a_mode});
        }
        if (count > 0) mi::keyswitch_batch(*ctx->c, v.data(), count, montgomery != 0, S(stream));
    });
}

int mi_relin(mi_context* ctx, int level, uint64_t* b, uint64_t* a, const uint64_t* c2,
             const uint64_t* key, int montgomery, void* stream) {
    return guard("ckks_relin", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(b); check_ptr(a); check_ptr(c2); check_ptr(key);
        mi::keyswitch(*ctx->c, mi::KsItem{c2, key, b, a, level, 0, 0}, montgomery != 0, S(stream));
    });
}

int mi_rotate(mi_context* ctx, int level, uint64_t* b, uint64_t* a, uint64_t galois,
              const uint64_t* key, int montgomery, void* stream) {
    return guard("ckks_rotate", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(b); check_ptr(a); check_ptr(key);
        mi::Context& c = *ctx->c;
        const int L = level + 1;
        const long long words = (long long)L * (long long)c.N;
        mi::keyswitch(c, mi::KsItem{a, key, b, a, level, 0, 1}, montgomery != 0, S(stream));
        uint64_t* tmp = (uint64_t*)c.workspace(S(stream), sizeof(uint64_t) * 2 * words);
        mi::automorphism(c, L, galois, b, tmp, 1, words, words, S(stream));
        mi::automorphism(c, L, galois, a, tmp + words, 1, words, words, S(stream));
        MI_HIP(hipMemcpyAsync(b, tmp, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, S(stream)));
        MI_HIP(hipMemcpyAsync(a, tmp + words, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, S(stream)));
    });
}

int mi_conjugate(mi_context* ctx, int level, uint64_t* b, uint64_t* a, const uint64_t* key,
                 int montgomery, void* stream) {
    return guard("ckks_conjugate", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(b); check_ptr(a); check_ptr(key);
        mi::Context& c = *ctx->c;
        const int L = level + 1;
        const long long words = (long long)L * (long long)c.N;
        const uint64_t g = 2 * c.N - 1;
        uint64_t* tmp = (uint64_t*)c.workspace(S(stream), sizeof(uint64_t) * 2 * words);
        mi::automorphism(c, L, g, b, tmp, 1, words, words, S(stream));
        mi::automorphism(c, L, g, a, tmp + words, 1, words, words, S(stream));
        MI_HIP(hipMemcpyAsync(b, tmp, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, S(stream)));
        MI_HIP(hipMemcpyAsync(a, tmp + words, sizeof(uint64_t) * words, hipMemcpyDeviceToDevice, S(stream)));
        mi::keyswitch(c, mi::KsItem{a, key, b, a, level, 0, 1}, montgomery != 0, S(stream));
    });
}

int mi_conjugate_batch(mi_context* ctx, int level, int count, uint64_t* const* b, uint64_t* const* a,
                       const uint64_t* key, int montgomery, void* stream) {
    return guard("ckks_conjugate", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(key);
        if (count < 1 || !b || !a) throw mi::Error(MI_EINVAL, "invalid ciphertext list");
        const long long W = (long long)(level + 1) * (long long)ctx->c->N;
        for (int i = 0; i < count; ++i) {
            check_ptr(b[i]);
            check_ptr(a[i]);
            for (int j = 0; j < i; ++j)  // conjugated in place: the ciphertexts must be distinct
                for (const uint64_t* x : {b[i], a[i]})
                    for (const uint64_t* y : {b[j], a[j]})
                        if (x < y + W && y < x + W) throw mi::Error(MI_EINVAL, "ciphertexts of a batch overlap");
            if (b[i] < a[i] + W && a[i] < b[i] + W) throw mi::Error(MI_EINVAL, "b and a overlap");
        }
        // as mi_conjugate: the automorphism first, then the key switch of the new a
        mi::conjugate_batch(*ctx->c, level, count, b, a, key, montgomery != 0, S(stream));
    });
}

int mi_hoist(mi_context* ctx, int level, const uint64_t* a, uint64_t* digits, void* stream) {
    return guard("ckks_hoist_ciphertext_a_qp_eval", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(a); check_ptr(digits);
        mi::hoist(*ctx->c, level, a, digits, S(stream));
    });
}

// Spans [lo, lo + W) of a batch's outputs must be pairwise disjoint (chunks
// of a batch may run concurrently); sorted, so O(n log n) for large batches.
static void check_disjoint(std::vector<const uint64_t*> lo, long long W, const char* what) {
    std::sort(lo.begin(), lo.end());
    for (size_t i = 1; i < lo.size(); ++i)
        if (lo[i] < lo[i - 1] + W) throw mi::Error(MI_EINVAL, what);
}

int mi_hoist_batch(mi_context* ctx, int level, int count, const uint64_t* const* a, uint64_t* digits,
                   int64_t digits_stride, void* stream) {
    return guard("ckks_hoist_ciphertext_a_qp_eval", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(digits);
        if (count < 1 || !a) throw mi::Error(MI_EINVAL, "invalid batch");
        const mi::Context& c = *ctx->c;
        const long long dw = (long long)c.parts(level) * (level + 1 + c.np) * (long long)c.N;
        if (digits_stride < dw) throw mi::Error(MI_EINVAL, "digit stride smaller than one ciphertext's digits");
        const long long W = (long long)(level + 1) * (long long)c.N;
        const uint64_t* d_end = digits + (count - 1) * digits_stride + dw;
        for (int i = 0; i < count; ++i) {
            check_ptr(a[i]);
            if (a[i] < d_end && digits < a[i] + W) throw mi::Error(MI_EINVAL, "digits overlap an input");
        }
        mi::hoist_batch(*ctx->c, level, count, a, digits, digits_stride, S(stream));
    });
}

int mi_rotate_batch(mi_context* ctx, int level, int count, uint64_t* const* b, uint64_t* const* a,
                    const uint64_t* galois, const uint64_t* const* keys, int montgomery, void* stream) {
    return guard("ckks_rotate", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (count < 1 || !b || !a || !galois || !keys) throw mi::Error(MI_EINVAL, "invalid batch");
        const long long W = (long long)(level + 1) * (long long)ctx->c->N;
        for (int i = 0; i < count; ++i) {
            check_galois(galois[i], false);
            check_ptr(b[i]); check_ptr(a[i]); check_ptr(keys[i]);
            for (int j = 0; j < i; ++j)  // rotated in place: the ciphertexts must be distinct
                for (const uint64_t* x : {b[i], a[i]})
                    for (const uint64_t* y : {b[j], a[j]})
                        if (x < y + W && y < x + W) throw mi::Error(MI_EINVAL, "ciphertexts of a batch overlap");
            if (b[i] < a[i] + W && a[i] < b[i] + W) throw mi::Error(MI_EINVAL, "b and a overlap");
        }
        mi::rotate_batch(*ctx->c, level, count, b, a, galois, keys, montgomery != 0, S(stream));
    });
}

int mi_rotate_hoisted_batch(mi_context* ctx, int level, const uint64_t* digits, const uint64_t* b, int count,
                            const uint64_t* galois, const uint64_t* const* keys, uint64_t* const* b_out,
                            uint64_t* const* a_out, int montgomery, void* stream) {
    return guard("ckks_hybrid_key_switch_hoisted", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(digits); check_ptr(b);
        if (count < 1 || !galois || !keys || !b_out || !a_out) throw mi::Error(MI_EINVAL, "invalid batch");
        const long long W = (long long)(level + 1) * (long long)ctx->c->N;
        std::vector<const uint64_t*> outs;
        for (int i = 0; i < count; ++i) {
            check_galois(galois[i], false);
            check_ptr(keys[i]); check_ptr(b_out[i]); check_ptr(a_out[i]);
            for (const uint64_t* o : {b_out[i], a_out[i]}) {
                if (o < b + W && b < o + W) throw mi::Error(MI_EINVAL, "an output overlaps b");
                outs.push_back(o);
            }
        }
        check_disjoint(outs, W, "outputs overlap");
        mi::rotate_hoisted_batch(*ctx->c, level, digits, b, count, galois, keys, b_out, a_out, montgomery != 0,
                                 S(stream));
    });
}

int mi_rotate_hoisted_items(mi_context* ctx, int level, int count, const uint64_t* digits, int64_t digits_stride,
                            int digits_period, const uint64_t* const* b, const uint64_t* galois,
                            const uint64_t* const* keys, uint64_t* const* b_out, uint64_t* const* a_out,
                            int montgomery, void* stream) {
    return guard("ckks_hybrid_key_switch_hoisted", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(digits);
        if (count < 1 || !b || !galois || !keys || !b_out || !a_out) throw mi::Error(MI_EINVAL, "invalid batch");
        const mi::Context& c = *ctx->c;
        const long long dw = (long long)c.parts(level) * (level + 1 + c.np) * (long long)c.N;
        if (digits_stride != 0 && digits_stride < dw)
            throw mi::Error(MI_EINVAL, "digit stride smaller than one ciphertext's digits");
        const long long W = (long long)(level + 1) * (long long)c.N;
        if (digits_period < 0) throw mi::Error(MI_EINVAL, "negative digit period");
        const int period = digits_period == 0 || digits_period > count ? count : digits_period;
        const uint64_t* d_end = digits + (digits_stride ? (long long)(period - 1) * digits_stride : 0) + dw;
        std::vector<const uint64_t*> outs, ins;
        for (int i = 0; i < count; ++i) {
            check_galois(galois[i], false);
            check_ptr(b[i]); check_ptr(keys[i]); check_ptr(b_out[i]); check_ptr(a_out[i]);
            for (const uint64_t* o : {b_out[i], a_out[i]}) {
                if (o < d_end && digits < o + W) throw mi::Error(MI_EINVAL, "an output overlaps the digits");
                outs.push_back(o);
            }
            ins.push_back(b[i]);
        }
        check_disjoint(outs, W, "outputs overlap");
        // no output may overlap an input b (chunks run concurrently)
        std::sort(ins.begin(), ins.end());
        ins.erase(std::unique(ins.begin(), ins.end()), ins.end());
        for (const uint64_t* o : outs) {
            auto it = std::upper_bound(ins.begin(), ins.end(), o);
            if (it != ins.end() && *it < o + W) throw mi::Error(MI_EINVAL, "an output overlaps an input");
            if (it != ins.begin() && o < *(it - 1) + W) throw mi::Error(MI_EINVAL, "an output overlaps an input");
        }
        mi::rotate_hoisted_items(*ctx->c, level, digits, digits_stride, period, b, count, galois, keys, b_out,
                                 a_out, montgomery != 0, S(stream));
    });
}

int mi_keyswitch_hoisted(mi_context* ctx, int level, const uint64_t* digits, const uint64_t* key,
                         uint64_t* b_out, uint64_t* a_out, int b_mode, int a_mode, uint64_t galois,
                         int montgomery, void* stream) {
    return guard("ckks_hybrid_key_switch_hoisted", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(digits); check_ptr(key); check_ptr(b_out); check_ptr(a_out);
        if ((b_mode != 0 && b_mode != 1) || (a_mode != 0 && a_mode != 1))
            throw mi::Error(MI_EINVAL, "invalid combine mode");
        check_galois(galois, true);
        mi::keyswitch_hoisted(*ctx->c, level, digits, key, b_out, a_out, b_mode, a_mode, galois,
                              montgomery != 0, S(stream));
    });
}

int mi_keyswitch_hoisted_qp(mi_context* ctx, int level, const uint64_t* digits, const uint64_t* key,
                            const uint64_t* b, uint64_t* b_qp, uint64_t* a_qp, uint64_t galois,
                            int montgomery, void* stream) {
    return guard("ckks_hybrid_key_switch_hoisted_qp_rotated", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(digits); check_ptr(key); check_ptr(b); check_ptr(b_qp); check_ptr(a_qp);
        check_galois(galois, false);
        mi::keyswitch_hoisted_qp(*ctx->c, level, digits, key, b, b_qp, a_qp, galois, montgomery != 0,
                                 S(stream));
    });
}

int mi_keyswitch_qp(mi_context* ctx, int level, const uint64_t* key_term, const uint64_t* key,
                    uint64_t* b_qp, uint64_t* a_qp, int montgomery, void* stream) {
    return guard("ckks_hybrid_key_switch_qp_from_key_term", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(key_term); check_ptr(key); check_ptr(b_qp); check_ptr(a_qp);
        mi::keyswitch_qp(*ctx->c, level, key_term, key, b_qp, a_qp, montgomery != 0, S(stream));
    });
}

int mi_moddown_qp(mi_context* ctx, int level, uint64_t* qp, void* stream) {
    return guard("moddown_qp_poly_to_q_eval", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(qp);
        mi::moddown_qp(*ctx->c, level, qp, S(stream));
    });
}

int mi_linear_transform_qp(mi_context* ctx, int level, const uint64_t* a, const uint64_t* b,
                           int nterms, const int64_t* diag_indices, const uint64_t* const* diags,
                           int pt_montgomery, int baby_steps, int slots, int nkeys,
                           const int64_t* key_rotations, const uint64_t* const* keys,
                           int montgomery, uint64_t* a_out, uint64_t* b_out, void* stream) {
    return guard("ckks_apply_linear_transform_owned_stage_qp", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (nkeys < 0) throw mi::Error(MI_EINVAL, "invalid key count");
        static_assert(sizeof(long long) == sizeof(int64_t), "int64 width");
        mi::LtStage st{level, nterms, (const long long*)diag_indices, diags, pt_montgomery,
                       baby_steps, slots, nkeys, (const long long*)key_rotations, keys, montgomery};
        mi::linear_transform_qp(*ctx->c, st, a, b, a_out, b_out, S(stream));
    });
}

int mi_linear_transform_qp_batch(mi_context* ctx, int level, int count, const uint64_t* const* a,
                                 const uint64_t* const* b, int nterms, const int64_t* diag_indices,
                                 const uint64_t* const* diags, int pt_montgomery, int baby_steps, int slots,
                                 int nkeys, const int64_t* key_rotations, const uint64_t* const* keys,
                                 int montgomery, uint64_t* const* a_out, uint64_t* const* b_out,
                                 void* stream) {
    return guard("ckks_apply_linear_transform_owned_stage_qp", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (count < 1) throw mi::Error(MI_EINVAL, "batch needs at least one ciphertext");
        if (nkeys < 0) throw mi::Error(MI_EINVAL, "invalid key count");
        if (!a || !b || !a_out || !b_out) throw mi::Error(MI_EINVAL, "null pointer argument");
        mi::LtStage st{level, nterms, (const long long*)diag_indices, diags, pt_montgomery,
                       baby_steps, slots, nkeys, (const long long*)key_rotations, keys, montgomery};
        mi::linear_transform_qp_batch(*ctx->c, st, count, a, b, a_out, b_out, S(stream));
    });
}

int mi_linear_transform_q_batch(mi_context* ctx, int level, int count, const uint64_t* const* a,
                                const uint64_t* const* b, int nterms, const int64_t* diag_indices,
                                const uint64_t* const* diags, int pt_montgomery, int baby_steps, int slots,
                                int nkeys, const int64_t* key_rotations, const uint64_t* const* keys,
                                int montgomery, uint64_t* const* a_out, uint64_t* const* b_out, void* stream) {
    return guard("ckks_apply_linear_transform_bsgs", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (count < 1) throw mi::Error(MI_EINVAL, "batch needs at least one ciphertext");
        if (nkeys < 0) throw mi::Error(MI_EINVAL, "invalid key count");
        if (!a || !b || !a_out || !b_out) throw mi::Error(MI_EINVAL, "null pointer argument");
        mi::LtStage st{level, nterms, (const long long*)diag_indices, diags, pt_montgomery,
                       baby_steps, slots, nkeys, (const long long*)key_rotations, keys, montgomery};
        mi::linear_transform_q_batch(*ctx->c, st, count, a, b, a_out, b_out, S(stream));
    });
}

int mi_linear_transform_q(mi_context* ctx, int level, const uint64_t* a, const uint64_t* b, int nterms,
                          const int64_t* diag_indices, const uint64_t* const* diags, int pt_montgomery,
                          int baby_steps, int slots, int nkeys, const int64_t* key_rotations,
                          const uint64_t* const* keys, int montgomery, uint64_t* a_out, uint64_t* b_out,
                          void* stream) {
    return mi_linear_transform_q_batch(ctx, level, 1, &a, &b, nterms, diag_indices, diags, pt_montgomery,
                                       baby_steps, slots, nkeys, key_rotations, keys, montgomery, &a_out,
                                       &b_out, stream);
}

int mi_set_encoding_tables(mi_context* ctx, const double* twiddles, const uint64_t* perm) {
    return guard("ckks_context", [&] {
        check_ctx(ctx);
        if (!twiddles || !perm) throw mi::Error(MI_EINVAL, "null pointer argument");
        mi::set_encoding_tables(*ctx->c, twiddles, perm);
    });
}

int mi_set_decoding_tables(mi_context* ctx, const double* twiddles, const uint64_t* perm) {
    return guard("ckks_decrypt", [&] {
        check_ctx(ctx);
        check_ptr(twiddles);
        check_ptr(perm);
        mi::set_decoding_tables(*ctx->c, twiddles, perm);
    });
}

int mi_decode_batch(mi_context* ctx, int count, const uint64_t* const* coeffs, int limbs,
                    const uint64_t* scale_mant, const int32_t* scale_exp, int coeff_domain, double* out,
                    void* stream) {
    return guard("ckks_decrypt", [&] {
        check_ctx(ctx);
        if (count < 1 || !coeffs || !scale_mant || !scale_exp) throw mi::Error(MI_EINVAL, "invalid decode batch");
        check_ptr(out);
        for (int i = 0; i < count; ++i) check_ptr(coeffs[i]);
        mi::decode_batch(*ctx->c, count, coeffs, limbs, scale_mant, (const int*)scale_exp, coeff_domain != 0, out,
                         S(stream));
    });
}

int mi_encode_slots(mi_context* ctx, const double* values, int count, double scale, int limbs, int eval,
                    uint64_t* out, int64_t poly_stride, void* stream) {
    return guard("ckks_encode", [&] {
        check_ctx(ctx);
        if (!values || !out) throw mi::Error(MI_EINVAL, "null pointer argument");
        mi::encode_slots(*ctx->c, values, count, scale, limbs, eval != 0, out, (long long)poly_stride, S(stream));
    });
}

int mi_sample_uniform(mi_context* ctx, uint64_t* out, int limbs, int basis0, uint64_t seed, uint64_t counter,
                      void* stream) {
    return guard("ckks_uniform_mask_polys", [&] {
        check_ctx(ctx);
        check_ptr(out);
        mi::sample_uniform(*ctx->c, out, limbs, basis0, seed, counter, S(stream));
    });
}

int mi_sample_gaussian(mi_context* ctx, uint64_t* out, int limbs, double sigma, uint64_t seed, uint64_t counter,
                       void* stream) {
    return guard("ckks_gaussian_noise_polys", [&] {
        check_ctx(ctx);
        check_ptr(out);
        mi::sample_gaussian(*ctx->c, out, limbs, sigma, seed, counter, S(stream));
    });
}

int mi_modraise(mi_context* ctx, int target_level, const uint64_t* poly_q0, uint64_t* out,
                void* stream) {
    return guard("modraise_eval_poly_from_q0", [&] {
        check_ctx(ctx);
        check_ptr(poly_q0); check_ptr(out);
        mi::modraise(*ctx->c, target_level, poly_q0, out, S(stream));
    });
}

int mi_modraise_batch(mi_context* ctx, int target_level, int count, const uint64_t* const* poly_q0,
                      uint64_t* const* out, void* stream) {
    return guard("modraise_eval_poly_from_q0", [&] {
        check_ctx(ctx);
        if (count < 1 || !poly_q0 || !out) throw mi::Error(MI_EINVAL, "invalid batch");
        if (target_level < 1 || target_level >= ctx->c->nq) throw mi::Error(MI_EINVAL, "invalid target level");
        const long long W = (long long)(target_level + 1) * (long long)ctx->c->N;
        for (int i = 0; i < count; ++i) {
            check_ptr(poly_q0[i]); check_ptr(out[i]);
            for (int j = 0; j < i; ++j)
                if (out[i] < out[j] + W && out[j] < out[i] + W) throw mi::Error(MI_EINVAL, "outputs overlap");
        }
        mi::modraise_batch(*ctx->c, target_level, count, poly_q0, out, S(stream));
    });
}

int mi_q0p0_keyswitch_batch(mi_context* ctx, int count, uint64_t* const* b, uint64_t* const* a,
                            const uint64_t* key, int montgomery, void* stream) {
    return guard("ckks_q0p0_key_switch_level0_inplace", [&] {
        check_ctx(ctx);
        check_ptr(key);
        if (count < 1 || !b || !a) throw mi::Error(MI_EINVAL, "invalid batch");
        const long long W = (long long)ctx->c->N;
        for (int i = 0; i < count; ++i) {
            check_ptr(b[i]); check_ptr(a[i]);
            for (int j = 0; j < i; ++j)  // switched in place: the ciphertexts must be distinct
                for (const uint64_t* x : {b[i], a[i]})
                    for (const uint64_t* y : {b[j], a[j]})
                        if (x < y + W && y < x + W) throw mi::Error(MI_EINVAL, "ciphertexts of a batch overlap");
            if (b[i] < a[i] + W && a[i] < b[i] + W) throw mi::Error(MI_EINVAL, "b and a overlap");
        }
        mi::q0p0_keyswitch_batch(*ctx->c, count, b, a, key, montgomery != 0, S(stream));
    });
}

int mi_q0p0_keyswitch(mi_context* ctx, uint64_t* b, uint64_t* a, const uint64_t* key,
                      int montgomery, void* stream) {
    return guard("ckks_q0p0_key_switch_level0_inplace", [&] {
        check_ctx(ctx);
        check_ptr(b); check_ptr(a); check_ptr(key);
        mi::q0p0_keyswitch(*ctx->c, b, a, key, montgomery != 0, S(stream));
    });
}

int mi_rescale_list(mi_context* ctx, int level, int npolys, uint64_t* const* polys, void* stream) {
    return guard("ckks_rescale", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (npolys < 1 || !polys) throw mi::Error(MI_EINVAL, "invalid poly list");
        for (int k = 0; k < npolys; ++k) check_ptr(polys[k]);
        mi::rescale_list(*ctx->c, level, polys, npolys, S(stream));
    });
}

int mi_mul_plain_rescale_list(mi_context* ctx, int level, int npolys, uint64_t* const* polys, const uint64_t* pt,
                              int pt_montgomery, int rescale, void* stream) {
    return guard("ckks_mul_inplace", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        check_ptr(pt);
        if (npolys < 1 || !polys) throw mi::Error(MI_EINVAL, "invalid poly list");
        for (int k = 0; k < npolys; ++k) check_ptr(polys[k]);
        if (rescale && level < 1) throw mi::Error(MI_EINVAL, "invalid level");
        mi::mul_plain_list(*ctx->c, level, polys, npolys, pt, pt_montgomery != 0, S(stream));
        if (rescale) mi::rescale_list(*ctx->c, level, polys, npolys, S(stream));
    });
}

int mi_poly_list_op(mi_context* ctx, int op, int level, int npolys, uint64_t* const* x, const uint64_t* const* y,
                    const uint64_t* residues, void* stream) {
    static const char* const names[] = {"ckks_ciphertext_clone",   "ckks_add_inplace(ct,ct)",
                                        "ckks_sub_inplace(ct,ct)", "ckks_mul_inplace(ct,int)",
                                        "ckks_evalchebyshev",      "ckks_add_integer_constant_inplace",
                                        "ckks_zero_polys",         "ckks_add_inplace(ct,pt)"};
    return guard(op >= 0 && op <= 7 ? names[op] : "mi_poly_list_op", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (op < MI_LIST_COPY || op > MI_LIST_ADD_PLAIN_MONT) throw mi::Error(MI_EINVAL, "invalid list operation");
        const bool uses_y = op <= MI_LIST_SUB || op == MI_LIST_MUL_SCALAR_ADD || op == MI_LIST_ADD_PLAIN_MONT;
        const bool uses_r = op == MI_LIST_MUL_SCALAR || op == MI_LIST_MUL_SCALAR_ADD || op == MI_LIST_ADD_SCALAR;
        if (npolys < 1 || !x || (uses_y && !y)) throw mi::Error(MI_EINVAL, "invalid poly list");
        if (uses_r && !residues) throw mi::Error(MI_EINVAL, "null residues");
        for (int k = 0; k < npolys; ++k) {
            check_ptr(x[k]);
            if (uses_y) check_ptr(y[k]);
            if (op == MI_LIST_COPY && x[k] == y[k]) throw mi::Error(MI_EINVAL, "copy onto itself");
        }
        mi::pair_list_op(*ctx->c, op, level, x, y, npolys, residues, S(stream));
    });
}

int mi_linear_combine(mi_context* ctx, int level, int nout, uint64_t* const* out, int nterms,
                      const uint64_t* const* in, const uint32_t* coef, void* stream) {
    return guard("ckks_aes_walsh", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (nout < 1 || !out || !in || !coef) throw mi::Error(MI_EINVAL, "invalid output list");
        if (nterms < 1 || nterms > 16) throw mi::Error(MI_EINVAL, "1..16 terms per output");
        for (int k = 0; k < nout; ++k) {
            check_ptr(out[k]);
            for (int t = 0; t < nterms; ++t) {
                const size_t i = (size_t)k * nterms + t;
                if (!coef[i]) continue;
                check_ptr(in[i]);
                for (int j = 0; j < nout; ++j)
                    if (in[i] == out[j]) throw mi::Error(MI_EINVAL, "an output aliases an input");
            }
        }
        mi::linear_combine(*ctx->c, level, nout, out, nterms, in, coef, S(stream));
    });
}

int mi_scaled_sum(mi_context* ctx, int level, int npolys, uint64_t* const* out, int nterms,
                  const uint64_t* const* in, const uint64_t* residues, int nconst, const uint64_t* constant,
                  void* stream) {
    return guard("ckks_evalchebyshev", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (npolys < 1 || !out || !in || !residues) throw mi::Error(MI_EINVAL, "invalid poly list");
        if (nterms < 1 || nterms > 64) throw mi::Error(MI_EINVAL, "1..64 terms per scaled sum");
        if (nconst < 0 || nconst > npolys || (nconst && !constant)) throw mi::Error(MI_EINVAL, "invalid constant");
        const size_t W = (size_t)(level + 1) * ctx->c->N;
        auto overlap = [W](const uint64_t* x, const uint64_t* y) { return x < y + W && y < x + W; };
        for (int k = 0; k < npolys; ++k) {
            check_ptr(out[k]);
            for (int j = 0; j < k; ++j)
                if (overlap(out[k], out[j])) throw mi::Error(MI_EINVAL, "output polys overlap");
        }
        for (size_t i = 0; i < (size_t)nterms * npolys; ++i) {
            if (!in[i]) continue;
            for (int j = 0; j < npolys; ++j)
                if (overlap(in[i], out[j])) throw mi::Error(MI_EINVAL, "an output overlaps an input");
        }
        mi::scaled_sum(*ctx->c, level, npolys, out, nterms, in, residues, nconst, constant, S(stream));
    });
}

int mi_walsh_product_sum(mi_context* ctx, int level, const uint64_t* const* low_b, const uint64_t* const* low_a,
                         const uint64_t* const* high_b, const uint64_t* const* high_a, const int16_t* cr,
                         const int16_t* ci, const uint64_t* const_abs, const int32_t* const_neg, const uint64_t* pt,
                         int pt_montgomery, uint64_t* acc_b, uint64_t* acc_a, uint64_t* raw_b, uint64_t* raw_a,
                         uint64_t* raw_c2, void* stream) {
    return guard("walsh_sum_for_output_pair", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (!low_b || !low_a || !high_b || !high_a || !cr || !ci || !const_abs || !const_neg)
            throw mi::Error(MI_EINVAL, "null table");
        check_ptr(pt); check_ptr(acc_b); check_ptr(acc_a); check_ptr(raw_b); check_ptr(raw_a); check_ptr(raw_c2);
        const uint64_t* outs[5] = {acc_b, acc_a, raw_b, raw_a, raw_c2};
        for (int m = 0; m < 15; ++m)
            for (const uint64_t* o : outs)
                if (o == low_b[m] || o == low_a[m] || o == high_b[m] || o == high_a[m] || o == pt)
                    throw mi::Error(MI_EINVAL, "outputs must not alias the inputs");
        mi::walsh_product_sum(*ctx->c, level, low_b, low_a, high_b, high_a, (const short*)cr, (const short*)ci,
                              const_abs, (const int*)const_neg, pt, pt_montgomery != 0, acc_b, acc_a, raw_b,
                              raw_a, raw_c2, S(stream));
    });
}

int mi_walsh_product_sum_batch(mi_context* ctx, int level, int count, const uint64_t* const* low_b,
                               const uint64_t* const* low_a, const uint64_t* const* high_b,
                               const uint64_t* const* high_a, const int16_t* const* cr, const int16_t* const* ci,
                               const uint64_t* const* const_abs, const int32_t* const* const_neg, const uint64_t* pt,
                               int pt_montgomery, uint64_t* const* acc_b, uint64_t* const* acc_a,
                               uint64_t* const* raw_b, uint64_t* const* raw_a, uint64_t* const* raw_c2, void* stream) {
    return guard("walsh_sum_for_output_pair", [&] {
        check_ctx(ctx);
        check_level(*ctx->c, level);
        if (count < 1 || !low_b || !low_a || !high_b || !high_a || !cr || !ci || !const_abs || !const_neg || !acc_b ||
            !acc_a || !raw_b || !raw_a || !raw_c2)
            throw mi::Error(MI_EINVAL, "null table");
        check_ptr(pt);
        const long long W = (long long)(level + 1) * (long long)ctx->c->N;
        std::vector<const uint64_t*> outs, ins;
        for (int i = 0; i < count; ++i) {
            if (!cr[i] || !ci[i] || !const_abs[i] || !const_neg[i]) throw mi::Error(MI_EINVAL, "null table");
            for (const uint64_t* o : {acc_b[i], acc_a[i], raw_b[i], raw_a[i], raw_c2[i]}) {
                check_ptr(o);
                outs.push_back(o);
            }
            for (int m = 0; m < 15; ++m)
                for (const uint64_t* x : {low_b[15 * i + m], low_a[15 * i + m], high_b[15 * i + m],
                                          high_a[15 * i + m]}) {
                    check_ptr(x);
                    ins.push_back(x);
                }
        }
        ins.push_back(pt);
        // members run concurrently: outputs pairwise disjoint, none over an input
        check_disjoint(outs, W, "outputs overlap");
        std::sort(ins.begin(), ins.end());
        ins.erase(std::unique(ins.begin(), ins.end()), ins.end());
        for (const uint64_t* o : outs) {
            auto it = std::upper_bound(ins.begin(), ins.end(), o);
            if (it != ins.end() && *it < o + W) throw mi::Error(MI_EINVAL, "outputs must not alias the inputs");
            if (it != ins.begin() && o < *(it - 1) + W) throw mi::Error(MI_EINVAL, "outputs must not alias the inputs");
        }
        mi::Context& c = *ctx->c;
        mi::over_streams(c, S(stream), count, 0, [&](hipStream_t st, mi::u64*, int i) {
            mi::walsh_product_sum(c, level, low_b + 15 * i, low_a + 15 * i, high_b + 15 * i, high_a + 15 * i,
                                  (const short*)cr[i], (const short*)ci[i], const_abs[i], (const int*)const_neg[i],
                                  pt, pt_montgomery != 0, acc_b[i], acc_a[i], raw_b[i], raw_a[i], raw_c2[i], st);
        });
    });
}

int mi_graph_begin(mi_context* ctx, void* stream) {
    return guard("mi_graph_begin", [&] {
        check_ctx(ctx);
        if (!stream) throw mi::Error(MI_EINVAL, "capture needs a created stream, not the null stream");
        MI_HIP(hipStreamBeginCapture(S(stream), hipStreamCaptureModeRelaxed));
        ctx->c->capture_begin();
    });
}

int mi_graph_end(mi_context* ctx, void* stream, mi_graph** out) {
    return guard("mi_graph_end", [&] {
        check_ctx(ctx);
        if (!stream || !out) throw mi::Error(MI_EINVAL, "null argument");
        *out = nullptr;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        MI_HIP(hipStreamIsCapturing(S(stream), &cap));
        if (cap == hipStreamCaptureStatusNone) throw mi::Error(MI_EINVAL, "stream is not capturing");
        mi_graph* g = new mi_graph;
        hipError_t e = hipStreamEndCapture(S(stream), &g->graph);
        if (e == hipSuccess && g->graph) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess || !g->exec) {
            if (g->graph) (void)hipGraphDestroy(g->graph);
            delete g;
            ctx->c->capture_end(false);
            (void)hipGetLastError();
            throw mi::Error(MI_EHIP, std::string("capture did not produce a graph: ") + hipGetErrorString(e) +
                                         " (a call inside the capture synchronised, allocated a workspace or"
                                         " failed; its own error was returned then)");
        }
        ctx->c->capture_end(true);
        *out = g;
    });
}

int mi_graph_launch(mi_context* ctx, mi_graph* graph, void* stream) {
    return guard("mi_graph_launch", [&] {
        check_ctx(ctx);
        if (!graph || !graph->exec) throw mi::Error(MI_EINVAL, "null graph");
        MI_HIP(hipGraphLaunch(graph->exec, S(stream)));
    });
}

int mi_graph_destroy(mi_context* ctx, mi_graph* graph) {
    return guard("mi_graph_destroy", [&] {
        check_ctx(ctx);
        if (!graph) return;
        (void)hipGraphExecDestroy(graph->exec);
        (void)hipGraphDestroy(graph->graph);
        delete graph;
        ctx->c->graph_destroyed();
    });
}

}  // extern "C"
