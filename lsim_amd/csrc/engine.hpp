// GPU fault-simulation engine (host interface; the HIP code lives in engine.hip).
//
// A circuit is made resident once (structure in HBM). A job is one fault
// simulation on a resident circuit: a fault list, a pattern block and options.
// The engine advances any set of jobs together: each iteration moves every
// active job forward by one word step and sweeps the tiles of all of them in
// one persistent launch, so jobs of different circuits and sizes share the
// device and new jobs can join between iterations.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "lsim.h"
#include "netlist.hpp"

namespace lsim {

struct DeviceError : std::runtime_error {
  int code;
  DeviceError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Allocator whose value-less construct leaves elements uninitialised: a
// resize() of a large result array touches no page, so the threads that fill
// it are the first to touch it.
template <typename T>
struct UninitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = UninitAlloc<U>;
  };
  UninitAlloc() = default;
  template <typename U>
  UninitAlloc(const UninitAlloc<U> &) noexcept {}
  template <typename U>
  void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void *>(p)) U;
  }
  template <typename U, typename... A>
  void construct(U *p, A &&...a) {
    ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
  }
};
template <typename T>
using result_vector = std::vector<T, UninitAlloc<T>>;

// Per-job options.
struct JobOptions {
  bool drop = false;              // fault dropping
  bool report_last_round = true;  // with dropping: evaluate the last round for every fault
  bool direct = false;            // inject every fault directly (forced for circuits with MUX)
  bool keep_good_last = false;    // gather the last round's good values (reference accessors)
  bool mux_safe = false;          // no fault's cone reaches a MUX select input: stem tracing stays
                                  // valid in a circuit with MUX (the caller has checked)
  int first_words = 1;            // with dropping: words of the first step
  int word_growth = 2;            // with dropping: factor of words per step after each step (up to 64)
};

// Sweep engines.
enum EngineKind : int {
  ENGINE_PLAIN = 0,       // per-level launches (k_sweep_level)
  ENGINE_PERSISTENT = 1,  // persistent sweeps (k_sweep_pairs / k_sweep_events)
};

// Engine-wide options of a run / iteration.
struct EngineOptions {
  int mode = ENGINE_PERSISTENT;      // ENGINE_PLAIN or ENGINE_PERSISTENT
  int64_t scratch_mb = 0;            // sweep scratch budget, 0 = automatic
  int block_waves = 8;               // waves per workgroup of the node-wave persistent sweep (1..8)
  int sweep_kernel = 0;              // persistent sweep: 0 = flattened (node, lane) pairs, 1 = one node per wave
  bool good_graph = true;            // a lone starting job replays its circuit's captured good-machine graph
  int good_form = 1;                 // good machine of a lone job: 0 one launch per level, 1 level windows
  int win_levels = 4;                // level windows: most levels per window (1..8)
  int win_cap = 0;                   // level windows: nodes per block (0 = from the device's LDS)
  int pair_threads = 0;              // workgroup size of the pair sweep: 512, 256, or 0 = by step density
  int64_t pool_cap = 0;              // pair sweep: at most this many pool words per slot (0 = as memory allows)
  int pair_dense = 24;               // pair sweep: passes averaging this many active lanes per node go one node per wave (0: never)
  int pair_nopend = 1;               // pair sweep without pending flags: 0 never, 1 dense steps, 2 every step
  int pair_nopend_pct = 65;          // with pair_nopend 1: also after a step that staged more than this % of walked nodes
  bool pair_fo_pre = true;           // pair sweep with pending flags: load a staged node's first fanout in the scan
  int max_step_words = 128;          // with dropping: most words of one step (the growth stops there)
  double slot_share = 1.0;           // pair sweep: use at most this share of the resident workgroup slots
  bool autotune = false;             // run(): pick sweep kernel and word schedule by timed probes (cached)
  bool device_steps = true;          // persistent pair sweep: plan each word step on the device (no host round trip)
  bool step_graphs = true;           // device-planned steps: capture recurring steps' launches in graphs and replay them
};

struct RunResults {
  // Good-machine values of the last round, masked, per topological position
  // (only with keep_good_last).
  std::vector<uint64_t> good_last;
  // First detecting round per fault, -1 if never.
  result_vector<int32_t> first_round;
  // d(f, r) bit matrix, F rows of RB words (only without dropping).
  std::vector<uint64_t> det_bits;
  int64_t RB = 0;
  bool has_matrix = false;
  // d(f, R-1) per fault, when available.
  result_vector<uint8_t> last_round;
  bool has_last = false;
  lsim_run_stats stats{};
};

struct CircuitRes;  // resident circuit (engine.hip)
struct JobState;    // job (engine.hip)

class Engine {
 public:
  explicit Engine(int device);
  // An engine that allocates nothing: every device buffer it needs is carved
  // from `scratch` (`bytes` long, on host's device), host staging is pageable,
  // and it borrows host's stream and events (so it must not run while host
  // does). Running past `bytes` fails with DeviceError(LSIM_E_NOMEM).
  Engine(const Engine &host, void *scratch, size_t bytes);
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;
  int device() const;

  // Makes a circuit resident. The engine owns it until remove_circuit.
  CircuitRes *add_circuit(const Netlist &nl, const FaultStructure &fs);
  // The same circuit in this engine, reading src's device arrays (another
  // engine on the same device, which must keep src resident while this one
  // uses it): no second upload.
  CircuitRes *share_circuit(const CircuitRes *src);
  void remove_circuit(CircuitRes *c);
  int64_t circuit_bytes(const CircuitRes *c) const;

  // Creates a job on a resident circuit (faults as positions, patterns as
  // R x P words) and uploads its inputs. The engine owns it until release_job.
  JobState *add_job(CircuitRes *c, std::vector<int32_t> fault_site, std::vector<uint8_t> fault_sa,
                    int64_t num_patterns, int64_t R, const std::vector<uint64_t> &patterns, const JobOptions &opt);
  void release_job(JobState *j);
  // Makes a finished (or never run) job runnable again with the same inputs
  // still resident, under `opt`; false when `opt` needs other device arrays
  // (dropping or last-round reporting changed), in which case the caller
  // releases the job and adds a new one.
  bool reset_job(JobState *j, const JobOptions &opt);
  // Device bytes a job on circuit c would hold (admission by footprint).
  static int64_t job_bytes(const CircuitRes *c, int64_t faults, int64_t R, bool drop);
  // Least device memory a lone job on c needs in an engine that must not
  // allocate: the circuit's copy, the job's arrays, and (twice, as the
  // sweeps take half of what is left) one tile slot with its pool.
  static int64_t noalloc_min_bytes(const CircuitRes *c, int64_t faults, int64_t R, bool with_last, bool matrix);
  int64_t free_bytes() const;

  // Good machine for jobs not started yet (one batched launch per level).
  void start(const std::vector<JobState *> &jobs, const EngineOptions &eo);
  // One word step for every started, unfinished job. Jobs that finish are
  // finalised (results copied to the host). Returns the number still running.
  int64_t iterate(const std::vector<JobState *> &jobs, const EngineOptions &eo);
  // start + iterate until every job is done.
  void run(const std::vector<JobState *> &jobs, const EngineOptions &eo);

  bool done(const JobState *j) const;
  RunResults &results(JobState *j);

 private:
  struct Impl;
  Impl *impl_;
};

// Number of visible HIP devices (0 when the runtime has none).
int device_count();

// Device and page-locked host allocations the library has made so far, on
// all threads.
int64_t device_allocations();

// The library's fixed no-allocation scratch on `device` (a static device
// array of kNoAllocScratchBytes, one copy per device): its device address,
// looked up once per device with that device current.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
void *noalloc_scratch(int device);

}  // namespace lsim
