// Producers following fsim/generate_fault_patterns.py:
//  * faults (:7-24, :26-30): distinct (wire, value) pairs drawn uniformly,
//    PO ids redrawn, F <= 2 (N - O); written as "F" then "wire value" lines;
//  * patterns (:34-48, :56-79): distinct P-bit patterns drawn uniformly,
//    n <= 2^P; written by the "new" writer: the count, then per round of
//    `bits_per_round` patterns one decimal word per PI made of that PI's bits
//    over the round's patterns, first pattern most significant, then a blank
//    line and the same chunks in binary.
// The reference draws from Python's random module and iterates Python sets;
// here the draws come from the library's seeded xoshiro256** stream and the
// lists keep the order drawn, so files are reproducible from the seed.
#include "producers.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "rng.hpp"
#include "textio.hpp"

namespace lsim {

std::string produce_faults(uint64_t seed, int64_t N, int64_t F, int64_t P, int64_t O) {
  if (N < 0 || F < 0 || P < 0 || O < 0 || P + O > N) throw std::invalid_argument("producer: bad node counts");
  const int64_t slots = 2 * (N - O);
  if (F > slots)
    throw std::invalid_argument(
        "num_of_faults must be less than or equal to (2*(num_of_PIs_gates_POs-num_of_POs)).");
  Xoshiro256ss rng(splitmix64(seed) ^ 0x66617573ull);
  // Slot s < 2 (N - O): wire = s / 2 mapped around the PO ids, value = s % 2.
  auto wire_of = [&](int64_t s) {
    const int64_t w = s / 2;
    return w < P ? w : w + O;
  };
  std::vector<int64_t> picked;
  picked.reserve(size_t(F));
  if (2 * F <= slots) {
    std::unordered_set<int64_t> seen;
    seen.reserve(size_t(F) * 2);
    while (int64_t(picked.size()) < F) {
      const int64_t s = int64_t(rng.below(uint64_t(slots)));
      if (seen.insert(s).second) picked.push_back(s);
    }
  } else {
    // Dense request: a partial Fisher-Yates shuffle of every slot.
    std::vector<int64_t> all(static_cast<size_t>(slots));
    for (int64_t s = 0; s < slots; ++s) all[size_t(s)] = s;
    for (int64_t i = 0; i < F; ++i) {
      const int64_t j = i + int64_t(rng.below(uint64_t(slots - i)));
      std::swap(all[size_t(i)], all[size_t(j)]);
      picked.push_back(all[size_t(i)]);
    }
  }
  std::string out;
  TextOut t(out);
  t.u(uint64_t(F)).c('\n');
  for (int64_t s : picked) t.u(uint64_t(wire_of(s))).c(' ').u(uint64_t(s % 2)).c('\n');
  return out;
}

std::string pack_patterns(int64_t P, int64_t n, int bits_per_round, const uint8_t *bits) {
  if (P <= 0 || n <= 0) throw std::invalid_argument("Both num_of_PIs and num_of_output_patterns should be positive integers.");
  if (bits_per_round < 1 || bits_per_round > 64) throw std::invalid_argument("producer: bits per round must be 1..64");
  const int64_t B = bits_per_round;
  const int64_t team = n > B ? (n + B - 1) / B : 1;
  std::string out;
  TextOut t(out);
  t.u(uint64_t(n)).c('\n');
  for (int64_t r = 0; r < team; ++r) {
    const int64_t k0 = r * B, k1 = std::min(n, k0 + B);
    for (int64_t i = 0; i < P; ++i) {
      uint64_t w = 0;
      for (int64_t k = k0; k < k1; ++k) w = (w << 1) | uint64_t(bits[size_t(k * P + i)] & 1);
      t.u(w).c(' ');
    }
    t.c('\n');
  }
  t.c('\n');
  for (int64_t r = 0; r < team; ++r) {
    const int64_t k0 = r * B, k1 = std::min(n, k0 + B);
    for (int64_t i = 0; i < P; ++i) {
      for (int64_t k = k0; k < k1; ++k) t.c(char('0' + (bits[size_t(k * P + i)] & 1)));
      t.c(' ');
    }
    t.c('\n');
  }
  t.c('\n');
  return out;
}

std::string produce_patterns(uint64_t seed, int64_t P, int64_t n, int bits_per_round) {
  if (P <= 0 || n <= 0) throw std::invalid_argument("Both num_of_PIs and num_of_output_patterns should be positive integers.");
  if (P < 63 && n > (int64_t(1) << P)) throw std::invalid_argument("num_of_output_patterns is too large.");
  Xoshiro256ss rng(splitmix64(seed) ^ 0x7074726eull);
  std::vector<uint8_t> bits(size_t(n * P));
  std::unordered_set<std::string> seen;
  seen.reserve(size_t(n) * 2);
  std::string key(size_t(P), '0');
  for (int64_t k = 0; k < n;) {
    uint64_t word = 0;
    for (int64_t i = 0; i < P; ++i) {
      if (i % 64 == 0) word = rng.next();
      key[size_t(i)] = char('0' + ((word >> (i % 64)) & 1));
    }
    if (!seen.insert(key).second) continue;
    for (int64_t i = 0; i < P; ++i) bits[size_t(k * P + i)] = uint8_t(key[size_t(i)] - '0');
    ++k;
  }
  return pack_patterns(P, n, bits_per_round, bits.data());
}

}  // namespace lsim
