// Host-side circuit, fault list and pattern set in the reference's topological
// order, plus the derived structure the device kernels consume.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace lsim {

// Gate type codes, identical to the .ckt type ids (fsim/base_ppg.hpp:19-32).
enum GateCode : uint8_t {
  G_INV = 0,
  G_AND = 1,
  G_OR = 2,
  G_XOR = 3,
  G_NAND = 4,
  G_NOR = 5,
  G_XNOR = 6,
  G_MUX = 7,
  G_CLKBUF = 8,
  G_PI = 9,
  G_PO = 10,
};

// Input rejected by the reader (reference behaviour would be undefined).
struct InputError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Netlist {
  // Header counts (fsim/simulation.cpp:40).
  int64_t P = 0, O = 0, G = 0, W = 0, N = 0;

  // Topological order. Position p holds file id order[p]; order_idx is the
  // reference's _gate_idx_to_tpg_order_idx (fsim/simulation.cpp:781-784).
  // Positions: PIs [0,P), inner gates [P,P+G), POs [P+G,N).
  std::vector<int64_t> order, order_idx;
  std::vector<uint8_t> type;       // per position
  std::vector<int32_t> fin_off;    // N+1, fanin CSR over positions, pin order
  std::vector<int32_t> fin;
  std::vector<int32_t> level;      // per position (fsim/simulation.cpp:882-901)
  std::vector<int64_t> gates_per_level;  // reference _num_gates_per_level
  int32_t max_level = 0;           // POs' level
  bool has_mux = false;

  // Faults (fsim/simulation.cpp:943-959): site is a position, sa in {0,1}.
  std::vector<int32_t> fault_site;
  std::vector<uint8_t> fault_sa;

  // Patterns (fsim/simulation.cpp:961-976): R rounds of P words.
  int64_t num_patterns = 0, R = 0;
  std::vector<uint64_t> patterns;  // [R][P]

  int64_t E() const { return int64_t(fin.size()); }
  // Live bits in round r: 64 except a partial last round (fsim/simulation.cpp:119-122).
  int round_bits(int64_t r) const {
    return (num_patterns / (64 * (r + 1))) ? 64 : int(num_patterns % 64);
  }
};

// Parses the three texts and builds the netlist exactly as the reference
// reader does (fsim/simulation.cpp:37-63,664-976). Throws InputError.
void read_netlist(Netlist &nl, const char *ckt, size_t ckt_len, const char *flst, size_t flst_len,
                  const char *ptn, size_t ptn_len);
// The three parts separately: the circuit (fsim/simulation.cpp:40,664-941),
// a fault list against it (:943-959) and a pattern set against it (:961-976).
void read_circuit(Netlist &nl, const char *ckt, size_t ckt_len);
void read_faults(const Netlist &nl, const char *flst, size_t flst_len, std::vector<int32_t> &site,
                 std::vector<uint8_t> &sa);
void read_patterns(const Netlist &nl, const char *ptn, size_t ptn_len, int64_t &num_patterns, int64_t &R,
                   std::vector<uint64_t> &words);

// Derived structure for fault simulation.
struct FaultStructure {
  std::vector<int32_t> fo_off, fo;   // fanout CSR over distinct fanout nodes
  std::vector<uint8_t> reach_po;     // 1 if a PO is reachable from the node (PO itself included)
  // Stem analysis: every node x that is not a stem has exactly one distinct
  // fanout node, its "parent" in the fanout-free region; stems (>=2 distinct
  // fanout nodes) and POs end a region. Nodes with no fanout and no PO are dead.
  std::vector<int32_t> parent;       // -1 for stems / POs / dead nodes
  std::vector<int32_t> stems;        // stem positions, ascending
  std::vector<int32_t> stem_of;      // region head (stem or PO) for each node, -1 if dead
};

void build_fault_structure(const Netlist &nl, FaultStructure &fs);

}  // namespace lsim
