// Seeded samplers of the host layer: uniform residues, the discrete Gaussian
// error and Hamming-weight ternary secrets.  The reference draws these from
// getrandom / std::random_device (secret_key_generation.cpp:67-137,
// ckks_encoding.cpp:211-240, discrete_gaussian_sampler.hpp); here every draw
// comes from an explicit seed so that a run can be replayed.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace ckks_host {

inline uint64_t splitmix64(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded through splitmix64.
class Rng {
public:
    explicit Rng(uint64_t seed) {
        uint64_t st = seed;
        for (auto& w : s) w = splitmix64(st);
    }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in [0, q) by rejection
    uint64_t below(uint64_t q) {
        const uint64_t lim = UINT64_MAX - (UINT64_MAX % q + 1) % q;
        uint64_t x;
        do {
            x = next();
        } while (x > lim);
        return x % q;
    }
    double uniform01() { return (double)(next() >> 11) * 0x1.0p-53; }

private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t s[4];
};

// Discrete Gaussian over the integers by inversion of the cumulative table on
// [-B, B], B = ceil(12.0061 * sigma) (the reference's table sampler,
// discrete_gaussian_sampler.hpp:46-83).
class GaussianSampler {
public:
    explicit GaussianSampler(double sigma) {
        bound = (int64_t)std::ceil(12.00610553538285 * sigma);
        const double var2 = 2.0 * sigma * sigma;
        double total = 0.0;
        for (int64_t x = -bound; x <= bound; ++x) {
            total += std::exp(-(double)(x * x) / var2);
            cdf.push_back(total);
        }
        for (double& v : cdf) v /= total;
    }
    int64_t sample(Rng& rng) const {
        const double u = rng.uniform01();
        size_t lo = 0, hi = cdf.size() - 1;
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (cdf[mid] < u) lo = mid + 1;
            else hi = mid;
        }
        return (int64_t)lo - bound;
    }

private:
    int64_t bound;
    std::vector<double> cdf;
};

// Exactly h nonzero coefficients in {-1, +1} at uniformly chosen positions
// (generate_hamming_weight_ternary_secret, secret_key_generation.cpp:106-137).
inline std::vector<int8_t> hamming_weight_ternary(size_t N, size_t h, Rng& rng) {
    std::vector<size_t> idx(N);
    for (size_t i = 0; i < N; ++i) idx[i] = i;
    for (size_t i = N - 1; i > 0; --i) {
        const size_t j = (size_t)rng.below(i + 1);
        std::swap(idx[i], idx[j]);
    }
    std::vector<int8_t> s(N, 0);
    for (size_t i = 0; i < h && i < N; ++i) s[idx[i]] = (rng.next() & 1) ? int8_t(1) : int8_t(-1);
    return s;
}

}  // namespace ckks_host
