// Fault-list and pattern producers with the conventions of the reference's
// fsim/generate_fault_patterns.py.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lsim {

// .flst text: F distinct (wire id, stuck value) pairs, ids uniform over
// [0, N) less the PO ids [P, P+O), values uniform over {0, 1}, listed in the
// order drawn. Throws std::invalid_argument when F > 2 (N - O).
std::string produce_faults(uint64_t seed, int64_t num_nodes, int64_t num_faults, int64_t num_pis, int64_t num_pos);

// .ptn text of explicit patterns: bits[k * P + i] is PI i in pattern k
// (0 or 1). Rounds of `bits_per_round` patterns; within a round PI i's word
// is the binary number read from its bits in pattern order, the first
// pattern being the most significant. Decimal words (P per line, each
// followed by a space), a blank line, the same words in binary, a blank line.
std::string pack_patterns(int64_t num_pis, int64_t num_patterns, int bits_per_round, const uint8_t *bits);

// .ptn text of n distinct random patterns (uniform bits) packed as above.
// Throws std::invalid_argument when P or n is not positive or n > 2^P.
std::string produce_patterns(uint64_t seed, int64_t num_pis, int64_t num_patterns, int bits_per_round);

}  // namespace lsim
