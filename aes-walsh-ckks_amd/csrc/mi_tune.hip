// Autotuner for the launch shapes that have more than one configuration:
//   * the NTT passes' limbs per workgroup (1 or 2), per ring dimension and
//     launch size (six buckets of limbs per launch);
//   * the key-switch group size (items per launch sequence: 8, 16 or 32),
//     per ring dimension.
// Each candidate is timed with HIP events on a scratch batch (the timing of
// these kernels does not depend on the residues), the fastest kept.  Results
// are cached per device (GPU architecture and compute-unit count) and ring
// dimension, in memory for the process and in a text file
// ($MICKKS_TUNE_CACHE, else $HOME/.cache/mickks/tune.txt) across processes.
// MICKKS_TUNE=0 skips tuning (the defaults: 1 limb per workgroup, groups of
// 32).  Every configuration computes the same residues: the choice changes
// only speed.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <sys/stat.h>

#include "mi_internal.hpp"
#include "mi_ntt_kernels.hpp"

namespace mi {

namespace {

std::mutex g_mu;
std::map<std::string, TunePlan> g_plans;  // "<arch> <cus> <logN>" -> plan
bool g_loaded = false;

std::string cache_path() {
    if (const char* p = std::getenv("MICKKS_TUNE_CACHE")) return p;
    const char* h = std::getenv("HOME");
    if (!h) return "";
    return std::string(h) + "/.cache/mickks/tune.txt";
}

void load_cache() {
    if (g_loaded) return;
    g_loaded = true;
    const std::string path = cache_path();
    if (path.empty()) return;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string arch;
        int cus, logN;
        TunePlan p;
        if (!(in >> arch >> cus >> logN)) continue;
        bool ok = true;
        for (int b = 0; b < TUNE_BUCKETS; ++b) ok = ok && (in >> p.ntt_variant[b]) && p.ntt_variant[b] >= 1 && p.ntt_variant[b] <= NTT_VARIANTS;
        ok = ok && (in >> p.ks_group) && (p.ks_group == 8 || p.ks_group == 16 || p.ks_group == 32);
        if (ok) {
            p.tuned = true;
            p.from_cache = true;
            g_plans[arch + " " + std::to_string(cus) + " " + std::to_string(logN)] = p;
        }
    }
}

void save_cache() {
    const std::string path = cache_path();
    if (path.empty()) return;
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos) {
        // mkdir -p of the directory (errors surface as a failed open below)
        std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); ++i)
            if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
    }
    std::ofstream f(path, std::ios::trunc);
    if (!f) return;  // read-only home: the plans stay per process
    for (const auto& kv : g_plans) {
        f << kv.first;
        for (int b = 0; b < TUNE_BUCKETS; ++b) f << ' ' << kv.second.ntt_variant[b];
        f << ' ' << kv.second.ks_group << '\n';
    }
}

// launch-size buckets: limbs per launch below 4, 16, 64, 256, 1024, and above
constexpr int kBucketCount[TUNE_BUCKETS] = {2, 8, 32, 128, 512, 2048};

template <int LPB, bool HALF>
void ntt_pair(const Context& c, const PlainJob& j, int count, hipStream_t s) {
    launch_cols<true, LPB, HALF>(c, j, count, s);
    launch_rows<true, LPB, HALF>(c, j, count, s);
    launch_rows<false, LPB, HALF>(c, j, count, s);
    launch_cols<false, LPB, HALF>(c, j, count, s);
}

float time_ms(hipStream_t s, const std::function<void()>& f, int reps) {
    hipEvent_t a, b;
    MI_HIP(hipEventCreate(&a));
    MI_HIP(hipEventCreate(&b));
    f();  // warm-up
    MI_HIP(hipEventRecord(a, s));
    for (int r = 0; r < reps; ++r) f();
    MI_HIP(hipEventRecord(b, s));
    MI_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    MI_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

TunePlan measure(Context& c) {
    TunePlan p;
    p.tuned = true;
    hipStream_t s = nullptr;
    MI_HIP(hipStreamCreate(&s));
    const u64 N = c.N;
    // NTT: limbs of the context's moduli, cycling; at most 256 MiB of scratch
    const int max_count = (int)std::min<u64>(kBucketCount[TUNE_BUCKETS - 1], ((u64)256 << 20) / (8 * N));
    u64* buf = nullptr;
    MI_HIP(hipMalloc(&buf, sizeof(u64) * (size_t)max_count * N));
    MI_HIP(hipMemsetAsync(buf, 0, sizeof(u64) * (size_t)max_count * N, s));
    for (int b = 0; b < TUNE_BUCKETS; ++b) {
        const int count = std::min(kBucketCount[b], max_count);
        const int lpp = std::min(count, c.nb());
        const int npolys = count / lpp, n = npolys * lpp;
        const PlainJob j = plain_job(LimbBatch{buf, (long long)lpp * (long long)N, lpp, 0, n}, N);
        const int reps = n >= 512 ? 3 : 8;
        const float t[NTT_VARIANTS] = {time_ms(s, [&] { ntt_pair<1, false>(c, j, n, s); }, reps),
                                       time_ms(s, [&] { ntt_pair<2, false>(c, j, n, s); }, reps),
                                       time_ms(s, [&] { ntt_pair<1, true>(c, j, n, s); }, reps)};
        // the plain variant unless another is clearly faster
        int best = 0;
        for (int v = 1; v < NTT_VARIANTS; ++v)
            if (t[v] < 0.97f * t[0] && t[v] < t[best]) best = v;
        p.ntt_variant[b] = best + 1;
    }
    MI_HIP(hipStreamSynchronize(s));
    MI_HIP(hipFree(buf));
    // key-switch group size: 32 switches at the middle level in groups of 8, 16, 32
    const int level = c.nq / 2, L = level + 1;
    const size_t kw = ks_item_words(c, level);
    const size_t key_words = (size_t)c.parts(level) * 2 * c.nb() * N;
    int items = 32;
    while (items > 8 && sizeof(u64) * ((size_t)items * (3 * L * N + kw) + key_words) > ((size_t)512 << 20)) items /= 2;
    const size_t words = key_words + (size_t)items * (3 * (size_t)L * N + kw);
    u64* ks = nullptr;
    MI_HIP(hipMalloc(&ks, sizeof(u64) * words));
    MI_HIP(hipMemsetAsync(ks, 0, sizeof(u64) * words, s));
    const u64* key = ks;
    u64* io = ks + key_words;
    u64* ws = io + (size_t)items * 3 * L * N;
    std::vector<KsItem> it(items);
    for (int i = 0; i < items; ++i) {
        u64* base = io + (size_t)i * 3 * L * N;
        it[i] = KsItem{base, key, base + (size_t)L * N, base + 2 * (size_t)L * N, level, 0, 1};
    }
    // Two interleaved trials per candidate (min taken), so clock drift over
    // the measurement does not favour the candidate timed last; a smaller
    // group replaces 32 only on a clear (10%) win, since a group of 32 keeps
    // fewer, larger launches in the longer batches of a bootstrap or S-box.
    const int cand[3] = {32, 16, 8};
    float t[3] = {0.0f, 0.0f, 0.0f};
    for (int trial = 0; trial < 2; ++trial)
        for (int ci = 0; ci < 3; ++ci) {
            const int g = cand[ci];
            if (g > items) continue;
            const float ms = time_ms(s, [&] {
                for (int i0 = 0; i0 < items; i0 += g)
                    keyswitch_items_with(c, it.data() + i0, std::min(g, items - i0), true, ws, s);
            }, 3);
            t[ci] = trial ? std::min(t[ci], ms) : ms;
        }
    p.ks_group = 32;
    float best = t[0];
    for (int ci = 1; ci < 3; ++ci)
        if (cand[ci] <= items && t[ci] > 0.0f && t[ci] < 0.9f * t[0] && t[ci] < best) {
            best = t[ci];
            p.ks_group = cand[ci];
        }
    MI_HIP(hipStreamSynchronize(s));
    MI_HIP(hipFree(ks));
    MI_HIP(hipStreamDestroy(s));
    return p;
}

}  // namespace

void tune_context(Context& c) {
    if (const char* e = std::getenv("MICKKS_TUNE"))
        if (std::string(e) == "0") return;  // defaults
    hipDeviceProp_t prop;
    MI_HIP(hipGetDeviceProperties(&prop, c.device));
    const std::string key = std::string(prop.gcnArchName).substr(0, std::string(prop.gcnArchName).find(':')) + " " +
                            std::to_string(prop.multiProcessorCount) + " " + std::to_string(c.logN);
    std::lock_guard<std::mutex> g(g_mu);
    load_cache();
    auto f = g_plans.find(key);
    if (f == g_plans.end()) {
        TunePlan p = measure(c);
        f = g_plans.emplace(key, p).first;
        save_cache();
    }
    c.tune = f->second;
    c.tune_key = key;
}

int tune_bucket(int count) {
    int b = 0;
    for (int lim = 4; b < TUNE_BUCKETS - 1 && count >= lim; lim *= 4) ++b;
    return b;
}

}  // namespace mi
