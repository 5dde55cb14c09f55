// spg_rng.h - seeded generator for the synthetic inputs: splitmix64 seeding
// of xoshiro256** (Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", 2018).
#pragma once
#include <cstdint>

namespace spg {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    inline uint64_t next() {
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
    // uniform in [0, 1) with 53 random bits
    inline double uniform01() { return (double)(next() >> 11) * 0x1.0p-53; }
    // uniform in [-1, 1): 24-bit grid, every value exact in FP32
    inline float uniform_pm1() {
        return (float)(int32_t)(next() >> 40) * 0x1.0p-23f - 1.0f;
    }
    // uniform integer in [0, n) (Lemire's multiply-shift, rejection-free
    // bias below 2^-40 for the sizes used here)
    inline uint64_t below(uint64_t n) {
        return (uint64_t)(((unsigned __int128)next() * n) >> 64);
    }
};

}  // namespace spg
