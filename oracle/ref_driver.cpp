// Driver for the reference simulator built from its own sources (see
// oracle/Makefile, target `ref`). Calls fsim::Simulator exactly as the
// reference CLI does (main/simple.cpp:21-57): read(paths) then run(mode, n),
// and dumps the observable results as text for comparison.
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline).
//
// usage: fsim_ref <ckt> <flst> <ptn> <mode 0|1|2> <threads> <out.txt|-> [graph.dot]
// prints "read_s <t> run_s <t>" on stdout; with graph.dot, also writes the
// reference's Simulator::visualization_graph there.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "fsim/simulation.hpp"

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s ckt flst ptn mode threads out\n", argv[0]);
    return 2;
  }
  fsim::Simulator sim;
  auto t0 = std::chrono::steady_clock::now();
  sim.read(std::string(argv[1]), std::string(argv[2]), std::string(argv[3]));
  auto t1 = std::chrono::steady_clock::now();
  const int mode = std::atoi(argv[4]);
  const size_t threads = size_t(std::atol(argv[5]));
  sim.run(fsim::Simulator::Mode(mode), threads);
  auto t2 = std::chrono::steady_clock::now();
  std::printf("read_s %.6f run_s %.6f\n", std::chrono::duration<double>(t1 - t0).count(),
              std::chrono::duration<double>(t2 - t1).count());
  if (argc > 7 && std::string(argv[7]) != "-") sim.visualization_graph(std::string(argv[7]));
  const std::string out = argv[6];
  if (out == "-") return 0;
  std::ofstream o(out);
  o << sim.num_PIs() << ' ' << sim.num_POs() << ' ' << sim.num_gates() << ' ' << sim.num_wires() << ' '
    << sim.num_rounds() << ' ' << sim.faults().size() << '\n';
  for (auto v : sim.gate_idx_to_tpg_order_idx()) o << v << ' ';
  o << '\n';
  for (auto v : sim.num_gates_per_level()) o << v << ' ';
  o << '\n';
  for (size_t i = 0; i < sim.num_PIs(); ++i) o << sim.pi(i).level() << ' ';
  for (size_t i = 0; i < sim.num_gates(); ++i) o << sim.gate(i).level() << ' ';
  for (size_t i = 0; i < sim.num_POs(); ++i) o << sim.po(i).level() << ' ';
  o << '\n';
  for (auto v : sim.found_fault_to_pattern()) o << v << ' ';
  o << '\n';
  for (auto v : sim.g_pi_results()) o << v << ' ';
  o << '\n';
  for (auto v : sim.g_gate_results()) o << v << ' ';
  o << '\n';
  for (auto v : sim.g_po_results()) o << v << ' ';
  o << '\n';
  return 0;
}
