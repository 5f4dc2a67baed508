// Served walsh15 AES-CTR transciphering with continuous batching.
//
// A request is a run of AES-CTR blocks (public counter and ciphertext blocks)
// to be decrypted homomorphically under the server's encrypted AES key.  One
// evaluation (ckks_aes_ctr_walsh_sbox_complex_blocks) holds 2 x 512 blocks:
// logical block j rides in complex slot j mod 512, in the real lane for
// j < 512 and the imaginary lane above (as aes_bench.cpp:156-196 reads them).
//
// The key owner context holds the evaluation keys and the encrypted round
// keys once; each of `lanes` lanes (CKKSContext::makeLane: own stream, pool
// and workspaces, the owner's keys read in place) runs on its own host
// thread.  A free lane takes the waiting requests first come first served,
// whole requests only, up to 1024 blocks, evaluates them (the 16 S-boxes of
// a round as one batch), decrypts the 128 output bits and completes them;
// requests that arrive meanwhile wait for the next free lane, and the lanes'
// evaluations overlap on the device.
//
// Admission is by device memory: a lane opens only if its measured footprint
// (the peak of its memory pool over a warm-up evaluation) fits the budget
// beside the lanes already open; a request is charged its share of that
// footprint (blocks / 1024) and is refused, with its own error and without
// touching the others, if it exceeds one evaluation or if the requests
// queued and in flight would exceed the open lanes' memory, each lane
// counted once in flight and `max_queued_evaluations` times waiting.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <thread>

#include "ckks.hpp"

namespace {

constexpr size_t kBlocksPerEvaluation = 1024;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int64_t pool_peak(const CKKSContext& c) {
    int64_t st[4];
    mi_check(mi_memory_stats(c.handle(), st), "ckks_aes_serve");
    return st[1];
}

}  // namespace

std::vector<size_t> ckks_aes_serve_take(const std::vector<size_t>& queued, size_t capacity) {
    std::vector<size_t> offsets;
    size_t used = 0;
    for (size_t n : queued) {
        if (used + n > capacity) break;
        offsets.push_back(used);
        used += n;
    }
    return offsets;
}

struct CKKSAesCtrServer::Impl {
    struct Pending {
        uint64_t id;
        std::vector<std::array<uint8_t, 16>> counters, ciphertexts;
        int64_t bytes;
        double arrival;
    };

    CKKSContext& owner;
    CKKSAesServerOptions options;
    MyVector<CKKSCiphertext> round_keys;
    std::vector<std::unique_ptr<CKKSContext>> lanes;
    int64_t evaluation_bytes = 0, budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> queue;
    std::map<uint64_t, CKKSAesRequestResult> done;
    int64_t charged = 0;  // bytes of the requests queued and in flight
    uint64_t next_id = 1;
    size_t evaluations = 0;
    bool stopping = false;
    std::vector<std::thread> threads;

    Impl(CKKSContext& o, const std::array<uint8_t, 16>& aes_key, const CKKSAesServerOptions& opt)
        : owner(o), options(opt) {
        const auto& P = owner.getParams();
        if (!P.usesAesWalsh() || 2 * P.getAesWalshBlockCount() != kBlocksPerEvaluation)
            throw std::invalid_argument("ckks_aes_serve: the server expects the AESWalsh15 packing (2 x 512 blocks)");
        if (owner.isLane()) throw std::invalid_argument("ckks_aes_serve: the server needs the key-owning context");
        round_keys = ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(
            owner, aes128_expand_key(aes_key), P.getAesWalshBlockCount());
        mi_check(mi_stream_synchronize(owner.handle(), owner.stream()), "ckks_aes_serve");
        int64_t free_bytes = 0, total = 0;
        mi_check(mi_device_memory_info(owner.handle(), &free_bytes, &total), "ckks_aes_serve");
        budget = options.memory_budget_bytes > 0 ? options.memory_budget_bytes : (int64_t)(0.9 * (double)free_bytes);
        const std::vector<std::array<uint8_t, 16>> zeros(kBlocksPerEvaluation, std::array<uint8_t, 16>{});
        int64_t used = 0;
        for (size_t l = 0; l < std::max<size_t>(1, options.lanes); ++l) {
            auto lane = CKKSContext::makeLane(owner);
            // warm-up: sizes the lane's pool and workspaces, measures its footprint
            (void)ckks_aes_ctr_walsh_sbox_complex_blocks(*lane, round_keys, zeros, zeros);
            mi_check(mi_stream_synchronize(lane->handle(), lane->stream()), "ckks_aes_serve");
            const int64_t need = pool_peak(*lane);
            evaluation_bytes = std::max(evaluation_bytes, need);
            if (!lanes.empty() && used + need > budget) break;
            used += need;
            lanes.push_back(std::move(lane));
        }
        for (size_t l = 0; l < lanes.size(); ++l) threads.emplace_back([this, l] { run(l); });
    }

    ~Impl() {
        {
            std::lock_guard<std::mutex> g(mu);
            stopping = true;
        }
        cv_work.notify_all();
        for (auto& t : threads) t.join();
    }

    // The device memory of the open lanes, each charged once in flight and
    // `max_queued_evaluations` times waiting.
    int64_t admission_limit() const {
        return (int64_t)lanes.size() * evaluation_bytes * (1 + (int64_t)options.max_queued_evaluations);
    }

    uint64_t submit(std::vector<std::array<uint8_t, 16>> counters, std::vector<std::array<uint8_t, 16>> cts) {
        if (counters.size() != cts.size())
            throw std::invalid_argument("ckks_aes_serve: counter and ciphertext block counts differ");
        const double arrival = now_s();
        std::lock_guard<std::mutex> g(mu);
        if (stopping) throw std::runtime_error("ckks_aes_serve: the server is stopping");
        const uint64_t id = next_id++;
        const size_t n = counters.size();
        const int64_t bytes =
            (int64_t)std::ceil((double)evaluation_bytes * (double)n / (double)kBlocksPerEvaluation);
        std::string refusal;
        if (n == 0 || n > std::min(options.max_request_blocks, kBlocksPerEvaluation))
            refusal = "ckks_aes_serve: refused: a request holds 1.." +
                      std::to_string(std::min(options.max_request_blocks, kBlocksPerEvaluation)) + " blocks";
        else if (charged + bytes > admission_limit())
            refusal = "ckks_aes_serve: refused: the requests queued and in flight would exceed the lanes' device "
                      "memory (" + std::to_string(charged + bytes) + " > " + std::to_string(admission_limit()) +
                      " bytes)";
        if (!refusal.empty()) {
            CKKSAesRequestResult r;
            r.id = id;
            r.refused = true;
            r.error = refusal;
            r.arrival = r.finish = arrival;
            done.emplace(id, std::move(r));
            cv_done.notify_all();
            return id;
        }
        charged += bytes;
        queue.push_back(Pending{id, std::move(counters), std::move(cts), bytes, arrival});
        cv_work.notify_one();
        return id;
    }

    CKKSAesRequestResult wait(uint64_t id) {
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return done.count(id) != 0; });
        CKKSAesRequestResult r = std::move(done.at(id));
        done.erase(id);
        return r;
    }

    void run(size_t l) {
        CKKSContext& lane = *lanes[l];
        const size_t half = kBlocksPerEvaluation / 2;
        for (;;) {
            std::vector<Pending> batch;
            std::vector<size_t> offset;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || !queue.empty(); });
                if (queue.empty()) return;  // stopping with nothing left
                std::vector<size_t> sizes;
                for (const auto& r : queue) sizes.push_back(r.counters.size());
                offset = ckks_aes_serve_take(sizes, kBlocksPerEvaluation);
                for (size_t k = 0; k < offset.size(); ++k) {
                    batch.push_back(std::move(queue.front()));
                    queue.pop_front();
                }
            }
            const double start = now_s();
            std::vector<std::array<uint8_t, 16>> counters(kBlocksPerEvaluation), cts(kBlocksPerEvaluation);
            for (size_t k = 0; k < batch.size(); ++k) {
                std::copy(batch[k].counters.begin(), batch[k].counters.end(), counters.begin() + offset[k]);
                std::copy(batch[k].ciphertexts.begin(), batch[k].ciphertexts.end(), cts.begin() + offset[k]);
            }
            std::vector<CKKSAesRequestResult> results(batch.size());
            try {
                auto bits = ckks_aes_ctr_walsh_sbox_complex_blocks(lane, round_keys, counters, cts);
                std::vector<const CKKSCiphertext*> ptrs;
                for (const auto& b : bits) ptrs.push_back(&b);
                auto slots = ckks_decrypt_batch(lane, ptrs);
                const double finish = now_s();
                for (size_t k = 0; k < batch.size(); ++k) {
                    auto& res = results[k];
                    const size_t n = batch[k].counters.size();
                    res.values.assign(n, std::array<double, 128>{});
                    res.plaintexts.assign(n, std::array<uint8_t, 16>{});
                    for (size_t j = 0; j < n; ++j) {
                        const size_t b = offset[k] + j, slot = b % half;
                        for (size_t bit = 0; bit < 128; ++bit) {
                            const double v = b < half ? slots[bit][slot].real() : slots[bit][slot].imag();
                            res.values[j][bit] = v;
                            if (v >= 0.5) res.plaintexts[j][bit / 8] |= (uint8_t)(1u << (bit % 8));
                        }
                    }
                    res.finish = finish;
                }
            } catch (const std::exception& e) {
                for (auto& res : results) {
                    res.error = e.what();
                    res.finish = now_s();
                }
            }
            std::lock_guard<std::mutex> g(mu);
            for (size_t k = 0; k < batch.size(); ++k) {
                results[k].id = batch[k].id;
                results[k].arrival = batch[k].arrival;
                results[k].start = start;
                results[k].lane = l;
                charged -= batch[k].bytes;
                done.emplace(batch[k].id, std::move(results[k]));
            }
            ++evaluations;
            cv_done.notify_all();
        }
    }
};

CKKSAesCtrServer::CKKSAesCtrServer(CKKSContext& owner, const std::array<uint8_t, 16>& aes_key,
                                   const CKKSAesServerOptions& options)
    : impl_(std::make_unique<Impl>(owner, aes_key, options)) {}
CKKSAesCtrServer::~CKKSAesCtrServer() = default;
uint64_t CKKSAesCtrServer::submit(std::vector<std::array<uint8_t, 16>> counter_blocks,
                                  std::vector<std::array<uint8_t, 16>> ciphertext_blocks) {
    return impl_->submit(std::move(counter_blocks), std::move(ciphertext_blocks));
}
CKKSAesRequestResult CKKSAesCtrServer::wait(uint64_t id) { return impl_->wait(id); }
size_t CKKSAesCtrServer::laneCount() const { return impl_->lanes.size(); }
int64_t CKKSAesCtrServer::evaluationBytes() const { return impl_->evaluation_bytes; }
size_t CKKSAesCtrServer::evaluations() const {
    std::lock_guard<std::mutex> g(impl_->mu);
    return impl_->evaluations;
}
