// Drives the C++ facade (include/lsim/simulator.hpp) the way the reference's
// callers drive fsim::Simulator (main/simple.cpp:21-57): read, optionally run,
// then print every accessor as "name v0 v1 ..." lines for the Python test to
// compare against the golden vectors.
//
// usage: facade_check <device> <ckt> <flst> <ptn> <stream|path> <run 0|1> [graph.dot]
// (with graph.dot, also Simulator::visualization_graph as main/simple.cpp:62-65)
#include <fstream>
#include <iostream>
#include <string>

#include "lsim/simulator.hpp"

template <typename V>
static void line(const char *name, const V &v) {
  std::cout << name;
  for (auto x : v) std::cout << ' ' << x;
  std::cout << '\n';
}

int main(int argc, char **argv) {
  if (argc != 7 && argc != 8) {
    std::cerr << "usage: facade_check device ckt flst ptn stream|path run [graph.dot]\n";
    return 2;
  }
  try {
    lsim::Simulator sim(std::stoi(argv[1]));
    if (std::string(argv[5]) == "stream") {
      std::ifstream c(argv[2]), f(argv[3]), p(argv[4]);
      sim.read(c, f, p);
    } else {
      sim.read(std::string(argv[2]), std::string(argv[3]), std::string(argv[4]));
    }
    std::cout << "num_vars " << sim.num_PIs() << ' ' << sim.num_POs() << ' ' << sim.num_gates() << ' '
              << sim.num_wires() << ' ' << sim.num_rounds() << '\n';
    line("tpg", sim.gate_idx_to_tpg_order_idx());
    line("per_level", sim.num_gates_per_level());
    std::cout << "levels";
    for (size_t i = 0; i < sim.num_PIs(); ++i) std::cout << ' ' << sim.pi(i).level();
    for (size_t i = 0; i < sim.num_gates(); ++i) std::cout << ' ' << sim.gate(i).level();
    for (size_t i = 0; i < sim.num_POs(); ++i) std::cout << ' ' << sim.po(i).level();
    std::cout << "\nfaults";
    for (const auto &f : sim.faults()) std::cout << ' ' << f.gate_with_fault() << ' ' << (f.gate_SA_fault_val() ? 1 : 0);
    std::cout << '\n';
    if (!sim.patterns().empty()) line("ptn0", sim.patterns()[0].output_value());
    if (argc == 8) sim.visualization_graph(argv[7]);
    std::cout << "type_names";
    for (int t = 0; t <= static_cast<int>(lsim::GateType::MAX_GATE_TYPE); ++t)
      std::cout << ' ' << sim.gateTypeToString(static_cast<lsim::GateType>(t));
    std::cout << '\n';
    if (std::stoi(argv[6])) {
      sim.run(lsim::Simulator::Mode::SEQUENTIAL, 1);
      line("pi", sim.g_pi_results());
      line("gate", sim.g_gate_results());
      line("po", sim.g_po_results());
      line("table", sim.found_fault_to_pattern());
      std::cout << "engine " << sim.last_run_stats().engine << '\n';
    }
  } catch (const std::exception &e) {
    std::cout << "error " << e.what() << '\n';
    return 1;
  }
  return 0;
}
