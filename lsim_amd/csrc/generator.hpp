// Seeded synthetic workloads in the reference's .ckt/.flst/.ptn text formats.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lsim {

struct GenSpec {
  uint64_t seed = 1;
  int64_t num_pis = 64;
  int64_t num_pos = 32;
  int64_t num_gates = 5000;
  int64_t num_levels = 40;
  int32_t fanout_cap = 0;     // 0: every source has weight 1 and no fanout cap
  double power_exp = 2.2;     // exponent of the discrete power-law source weights
  double locality = 0.7;      // probability the second fanin comes from the last `window` levels
  int32_t window = 3;
  double unary_frac = 0.1;    // share of INV/CLKBUF gates
  int64_t num_patterns = 1024;
};

struct GenText {
  std::string ckt, flst, ptn;
};

// Builds the circuit, the full SA0/SA1 fault list over inner gates and the
// packed random patterns. Throws std::invalid_argument on an impossible spec.
GenText generate(const GenSpec &spec);

// Circuit only (no faults, no patterns), as .ckt text.
std::string generate_ckt(const GenSpec &spec);

// Random patterns as .ptn text: ceil(n/64) rounds x P decimal words; in a
// partial last round only the low n%64 bits may be set.
std::string generate_ptn(uint64_t seed, int64_t num_pis, int64_t num_patterns);

// .flst text with a seeded subset (fraction in (0,1]) of the full SA0/SA1 list
// over inner gate ids [P+O, N), kept in ascending (id, value) order.
std::string generate_flst(uint64_t seed, int64_t num_pis, int64_t num_pos, int64_t num_gates,
                          double fraction);

uint64_t derive_seed(uint64_t seed, uint64_t stream);

}  // namespace lsim
