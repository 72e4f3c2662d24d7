// Deterministic PRNG used by the synthetic workload generator.
// xoshiro256** (Blackman & Vigna, 2018) seeded through splitmix64, as published
// at https://prng.di.unimi.it/ . Same seed -> same stream on every platform.
#pragma once
#include <cstdint>

namespace lsim {

inline uint64_t splitmix64(uint64_t &state) {
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

class Xoshiro256ss {
 public:
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t sm = seed;
    for (auto &w : s_) w = splitmix64(sm);
  }
  uint64_t next() {
    const uint64_t result = rotl(s_[1] * 5, 7) * 9;
    const uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return result;
  }
  // Uniform double in [0, 1) from the top 53 bits.
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  // Uniform integer in [0, n) (n > 0), Lemire's multiply-shift with rejection.
  uint64_t below(uint64_t n) {
    unsigned __int128 m = (unsigned __int128)next() * n;
    uint64_t lo = (uint64_t)m;
    if (lo < n) {
      const uint64_t thresh = (0 - n) % n;
      while (lo < thresh) {
        m = (unsigned __int128)next() * n;
        lo = (uint64_t)m;
      }
    }
    return (uint64_t)(m >> 64);
  }

 private:
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t s_[4];
};

}  // namespace lsim
