// Header-only C++ facade over the C ABI (include/lsim.h) with the shape of the
// reference's fsim::Simulator (fsim/simulation.hpp:39-188): read(paths),
// read(istreams), run(Mode, num_threads) and the same accessors returning
// const references that stay valid until the next read/run. Errors from the
// library are thrown as std::runtime_error, as the reference's read(paths)
// does (fsim/simulation.cpp:24-32).
#pragma once
#include <cstdint>
#include <iostream>
#include <istream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "../lsim.h"

namespace lsim {

// Gate type codes of the .ckt format (fsim/base_ppg.hpp:19-32).
enum class GateType { INV = 0, AND, OR, XOR, NAND, NOR, XNOR, MUX, CLKBUF, PI, PO, MAX_GATE_TYPE };

class Element {
 public:
  Element(size_t idx, size_t level) : idx_(idx), level_(level) {}
  size_t idx() const { return idx_; }
  size_t level() const { return level_; }

 private:
  size_t idx_, level_;
};

class Fault {
 public:
  Fault(size_t site, size_t val) : site_(site), val_(val) {}
  size_t gate_with_fault() const { return site_; }
  size_t gate_SA_fault_val() const { return val_; }

 private:
  size_t site_, val_;
};

class Pattern {
 public:
  explicit Pattern(std::vector<size_t> v) : v_(std::move(v)) {}
  std::vector<size_t> output_value() const { return v_; }

 private:
  std::vector<size_t> v_;
};

class Simulator {
 public:
  // Reference mode names (fsim/simulation.hpp:45); all three run the
  // persistent GPU engine.
  enum class Mode { SEQUENTIAL = 0, PARALLEL_TF, PARALLEL_OMP };

  explicit Simulator(int device = 0) {
    if (lsim_create(device, &h_) != LSIM_OK) throw std::runtime_error(lsim_last_error(nullptr));
  }
  ~Simulator() { lsim_destroy(h_); }
  Simulator(const Simulator &) = delete;
  Simulator &operator=(const Simulator &) = delete;

  void read(const std::string &ckt, const std::string &flst, const std::string &ptn) {
    check(lsim_read_files(h_, ckt.c_str(), flst.c_str(), ptn.c_str()));
    refresh_structure();
  }
  void read(std::istream &ckt, std::istream &flst, std::istream &ptn) {
    const std::string c = slurp(ckt), f = slurp(flst), p = slurp(ptn);
    check(lsim_read_buffers(h_, c.data(), c.size(), f.data(), f.size(), p.data(), p.size()));
    refresh_structure();
  }
  void set_option(const std::string &key, int64_t value) { check(lsim_set_option(h_, key.c_str(), value)); }
  void run(Mode mode, const size_t num_threads) {
    check(lsim_run(h_, static_cast<int>(mode), num_threads));
    const size_t P = num_PIs(), G = num_gates(), O = num_POs();
    g_pi_.assign(P, 0);
    g_gate_.assign(G, 0);
    g_po_.assign(O, 0);
    check(lsim_good_results(h_, ptr(g_pi_), P, ptr(g_gate_), G, ptr(g_po_), O));
    table_.assign(2 * faults_.size(), 0);
    const int rc = lsim_found_fault_to_pattern(h_, ptr(table_), table_.size());
    if (rc != LSIM_OK && rc != LSIM_E_STATE) check(rc);
  }

  // No reference counterpart: timing, traffic and the engine of the last run.
  lsim_run_stats last_run_stats() const {
    lsim_run_stats st{};
    check(lsim_last_run_stats(h_, &st));
    return st;
  }

  const std::vector<size_t> &g_pi_results() const { return g_pi_; }
  const std::vector<size_t> &g_gate_results() const { return g_gate_; }
  const std::vector<size_t> &g_po_results() const { return g_po_; }
  const Element &pi(size_t i) const { return pis_[i]; }
  const Element &gate(size_t i) const { return gates_[i]; }
  const Element &po(size_t i) const { return pos_[i]; }
  const std::vector<Element> &gates() const { return gates_; }
  const std::vector<Fault> &faults() const { return faults_; }
  const std::vector<Pattern> &patterns() const { return patterns_; }
  size_t num_PIs() const { return size_t(lsim_num_pis(h_)); }
  size_t num_gates() const { return size_t(lsim_num_gates(h_)); }
  size_t num_POs() const { return size_t(lsim_num_pos(h_)); }
  size_t num_wires() const { return size_t(lsim_num_wires(h_)); }
  size_t num_pos() const { return num_POs(); }
  size_t num_rounds() const { return size_t(lsim_num_rounds(h_)); }
  const std::vector<size_t> &gate_idx_to_tpg_order_idx() const { return order_; }
  const std::vector<size_t> &num_gates_per_level() const { return per_level_; }
  const std::vector<size_t> &found_fault_to_pattern() const { return table_; }
  // fsim/simulation.cpp:1126-1154.
  std::string gateTypeToString(GateType type) {
    static const char *kNames[] = {"INV", "AND", "OR", "XOR", "NAND", "NOR", "XNOR", "MUX", "CLKBUF", "PI", "PO"};
    const int t = static_cast<int>(type);
    return t >= 0 && t < 11 ? kNames[t] : "UNKNOWN";
  }
  // fsim/simulation.cpp:1157-1225: on a file that cannot be opened the
  // reference prints the message and returns; so does this.
  void visualization_graph(const std::string &outputFile) {
    const int rc = lsim_visualization_graph(h_, outputFile.c_str());
    if (rc == LSIM_E_IO) {
      std::cerr << lsim_last_error(h_) << std::endl;
      return;
    }
    check(rc);
  }

  // No reference counterpart: first detecting round per fault (-1: never).
  std::vector<int64_t> first_detect_round() const {
    std::vector<int64_t> out(faults_.size());
    check(lsim_first_detect_round(h_, out.data(), out.size()));
    return out;
  }
  lsim_sim *handle() const { return h_; }

 private:
  static std::string slurp(std::istream &in) {
    return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  }
  static uint64_t *ptr(std::vector<size_t> &v) {
    static_assert(sizeof(size_t) == sizeof(uint64_t), "size_t must be 64-bit");
    return reinterpret_cast<uint64_t *>(v.data());
  }
  void check(int rc) const {
    if (rc != LSIM_OK) throw std::runtime_error(lsim_last_error(h_));
  }
  void refresh_structure() {
    const size_t P = num_PIs(), G = num_gates(), O = num_POs(), N = P + G + O;
    order_.assign(N, 0);
    check(lsim_gate_idx_to_tpg_order_idx(h_, ptr(order_), N));
    per_level_.assign(size_t(lsim_num_levels(h_)), 0);
    check(lsim_num_gates_per_level(h_, ptr(per_level_), per_level_.size()));
    std::vector<size_t> lv(N);
    check(lsim_levels(h_, ptr(lv), N));
    pis_.clear();
    gates_.clear();
    pos_.clear();
    for (size_t i = 0; i < P; ++i) pis_.emplace_back(i, lv[i]);
    for (size_t i = 0; i < G; ++i) gates_.emplace_back(P + i, lv[P + i]);
    for (size_t i = 0; i < O; ++i) pos_.emplace_back(P + G + i, lv[P + G + i]);
    const size_t F = size_t(lsim_num_faults(h_));
    std::vector<size_t> site(F), val(F);
    check(lsim_faults(h_, ptr(site), ptr(val), F));
    faults_.clear();
    for (size_t i = 0; i < F; ++i) faults_.emplace_back(site[i], val[i]);
    const size_t R = num_rounds();
    std::vector<size_t> words(R * P);
    check(lsim_patterns(h_, ptr(words), words.size()));
    patterns_.clear();
    for (size_t r = 0; r < R; ++r)
      patterns_.emplace_back(std::vector<size_t>(words.begin() + r * P, words.begin() + (r + 1) * P));
    g_pi_.clear();
    g_gate_.clear();
    g_po_.clear();
    table_.clear();
  }

  lsim_sim *h_ = nullptr;
  std::vector<size_t> order_, per_level_, g_pi_, g_gate_, g_po_, table_;
  std::vector<Element> pis_, gates_, pos_;
  std::vector<Fault> faults_;
  std::vector<Pattern> patterns_;
};

}  // namespace lsim
