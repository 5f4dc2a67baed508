// The hybrid key switch of mi_keyswitch with its scratch in a fixed device
// array instead of the stream workspace (which grows with hipMalloc), for
// callers that cannot allocate device memory: code under graph capture,
// serving loops and harnesses that forbid allocation.
//
// `scratch` is one file-scope device array; each GPU that loads this library
// holds its own copy until the process ends.  Its address on a device comes
// from hipGetSymbolAddress, taken once per device when the first context is
// created there (the device is current then); host code uses only that
// address, never the array's name.  Every call on a device shares the array:
// calls that overlap on one device (concurrent streams, host threads) race.
// The key switch writes every scratch word it reads, so nothing needs
// clearing between calls.
#include <map>
#include <mutex>

#include "../../include/mickks.h"
#include "mi_internal.hpp"

namespace mi {

// 2^22 words = 32 MiB.  One key switch at level L - 1 needs
// (L + parts (L + K) + 2K + 2L) N words: every level of the walsh15 chain
// (N = 2^15, L <= 13, K = 4, dnum 3: at most 98 N words) fits.
constexpr size_t NOALLOC_SCRATCH_WORDS = size_t{1} << 22;

static __device__ __align__(256) u64 scratch[NOALLOC_SCRATCH_WORDS];

namespace {
std::mutex g_mu;
std::map<int, u64*> g_scratch;  // device ordinal -> that device's copy of scratch
}  // namespace

void noalloc_setup(int device) {
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch.count(device)) return;
    void* p = nullptr;
    MI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(MI_EHIP, "no device address for the key-switch scratch array");
    g_scratch[device] = (u64*)p;
}

size_t noalloc_scratch_words() { return NOALLOC_SCRATCH_WORDS; }

void keyswitch_noalloc(Context& c, const KsItem& it, bool montgomery, hipStream_t s) {
    if (it.level < 0 || it.level >= c.nq) throw Error(MI_EINVAL, "invalid ciphertext level");
    const size_t need = ks_item_words(c, it.level);
    if (need > NOALLOC_SCRATCH_WORDS)
        throw Error(MI_EINVAL, "key switch at level " + std::to_string(it.level) + " needs " + std::to_string(need) +
                            " scratch words; the fixed scratch holds " + std::to_string(NOALLOC_SCRATCH_WORDS));
    u64* ptr = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        auto f = g_scratch.find(c.device);
        if (f == g_scratch.end()) throw Error(MI_ERUNTIME, "key-switch scratch not set up on this device");
        ptr = f->second;
    }
    keyswitch_items_with(c, &it, 1, montgomery, ptr, s);
}

}  // namespace mi
