// Random-access probe of the device's memory system (measurement only: built
// into liblsim_probe.so, loaded by the benchmark, not part of liblsim.so).
// C entry points: int lsim_bench_probe_random_reads(int device, int64_t bytes,
// double *reads_per_s, double *latency_ns) and int lsim_bench_copy(int device,
// int64_t bytes, int reps, double *ms), 0 on success.
#pragma once
#include <cstdint>
#include <stdexcept>

namespace lsim {

struct ProbeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct ProbeResult {
  double reads_per_s = 0;  // independent random 8-byte reads served per second
  double latency_ns = 0;   // one read of a dependent chain
  double reads = 0;        // reads timed
  double ms = 0;           // their time
};

// Allocates `bytes` of device memory for the duration of the probe.
ProbeResult probe_random_reads(int device, int64_t bytes);
// `reps` plain copies of `bytes` between two device buffers; ms per copy.
double probe_copy(int device, int64_t bytes, int reps);

}  // namespace lsim
