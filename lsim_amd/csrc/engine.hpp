// GPU fault-simulation engine (host interface; the HIP code lives in engine.hip).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lsim.h"
#include "netlist.hpp"

namespace lsim {

struct DeviceError : std::runtime_error {
  int code;
  DeviceError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct RunOptions {
  int mode = LSIM_MODE_PLAIN;
  bool drop = false;
  bool report_last_round = true;
  int64_t scratch_mb = 0;
  bool direct = false;
  int block_waves = 8;
};

struct RunResults {
  // Good-machine values of the last round, masked, per topological position.
  std::vector<uint64_t> good_last;
  // First detecting round per fault, -1 if never (all modes).
  std::vector<int32_t> first_round;
  // d(f, r) bit matrix, F rows of RB words (only without dropping).
  std::vector<uint64_t> det_bits;
  int64_t RB = 0;
  bool has_matrix = false;
  // d(f, R-1) per fault, when available.
  std::vector<uint8_t> last_round;
  bool has_last = false;
  lsim_run_stats stats{};
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;
  // Makes the circuit, faults and patterns resident on the device.
  void load(const Netlist &nl, const FaultStructure &fs);
  void run(const RunOptions &opt, RunResults &res);
  int device() const;

 private:
  struct Impl;
  Impl *impl_;
};

// Number of visible HIP devices (0 when the runtime has none).
int device_count();

}  // namespace lsim
