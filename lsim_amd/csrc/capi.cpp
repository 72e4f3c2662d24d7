// C ABI of the simulator (include/lsim.h). Every entry point catches C++
// exceptions and turns them into LSIM_E_* codes plus a message.
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <unordered_set>
#include <memory>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <thread>

#include "engine.hpp"
#include "generator.hpp"
#include "ingest.hpp"
#include "loops.hpp"
#include "producers.hpp"
#include "lsim.h"
#include "netlist.hpp"

namespace {

thread_local std::string g_create_error;

template <typename Fn>
int guarded(std::string *err, Fn &&fn) {
  try {
    fn();
    if (err) err->clear();
    return LSIM_OK;
  } catch (const lsim::InputError &e) {
    if (err) *err = e.what();
    return LSIM_E_INPUT;
  } catch (const lsim::DeviceError &e) {
    if (err) *err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    if (err) *err = "host out of memory";
    return LSIM_E_NOMEM;
  } catch (const std::exception &e) {
    if (err) *err = e.what();
    return LSIM_E_ARG;
  }
}

int fail(std::string *err, int code, const std::string &msg) {
  if (err) *err = msg;
  return code;
}

bool read_file(const char *path, std::string &out) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  std::ostringstream ss;
  ss << in.rdbuf();
  out = ss.str();
  return true;
}

int check_n(std::string *err, const void *p, size_t n, size_t want) {
  if (!p && want) return fail(err, LSIM_E_ARG, "null output buffer");
  if (n != want)
    return fail(err, LSIM_E_ARG,
                "output buffer holds " + std::to_string(n) + " elements, expected " + std::to_string(want));
  return LSIM_OK;
}

// Report helpers shared by the three front ends.
void fill_table(const lsim::RunResults &r, int64_t F, int64_t R, uint64_t *out) {
  // {found in the last round, last round} (fsim/simulation.cpp:142-144); no
  // rounds leaves the table zero-filled (fsim/simulation.cpp:116).
  for (int64_t f = 0; f < F; ++f) {
    out[2 * f] = R > 0 ? r.last_round[f] : 0;
    out[2 * f + 1] = R > 0 ? uint64_t(R - 1) : 0;
  }
}

void fill_matrix(const lsim::RunResults &r, int64_t F, int64_t R, uint8_t *out) {
  for (int64_t f = 0; f < F; ++f)
    for (int64_t k = 0; k < R; ++k) out[f * R + k] = uint8_t((r.det_bits[f * r.RB + (k >> 6)] >> (k & 63)) & 1);
}

struct ResidentCircuit {
  lsim::Netlist nl;
  lsim::FaultStructure fs;
  lsim::CircuitRes *res = nullptr;
  lsim::CircuitRes *res2 = nullptr;  // the server's second lane
};

}  // namespace

// ------------------------------------------------------------ single simulator

struct lsim_sim {
  int device = -1;  // -1: host-only handle (read and accessors, no run)
  std::unique_ptr<lsim::Engine> engine;
  void *scratch = nullptr;  // the device's no-allocation scratch (lsim_run_noalloc)
  lsim::Netlist nl;
  lsim::FaultStructure fs;
  lsim::CircuitRes *circuit = nullptr;
  lsim::JobState *job = nullptr;  // the inputs' job, kept resident from run to run
  // Circuits with MUX: the faults whose cone reaches a MUX select input are
  // injected directly (job2); the others keep stem tracing (job).
  lsim::JobState *job2 = nullptr;
  bool job_split = false;
  std::vector<int32_t> split_stem, split_direct;  // fault indices of job and job2
  bool loaded = false;
  bool ran = false;
  lsim::JobOptions jopt;
  lsim::EngineOptions eopt;
  lsim::RunResults res;
  std::string err;
};

namespace {

// The reference's MUX reads its select as a whole word (`!s`,
// fsim/simulation.cpp:1056), which couples the 64 patterns of a word. A fault
// whose cone reaches no MUX select propagates pattern by pattern (a MUX's data
// inputs do), so stem tracing stays exact for it; the others are injected
// directly. Splits the fault indices accordingly (empty lists without MUX).
void mux_split(const lsim::Netlist &nl, const lsim::FaultStructure &fs, std::vector<int32_t> &stem,
               std::vector<int32_t> &direct) {
  stem.clear();
  direct.clear();
  if (!nl.has_mux) return;
  const int64_t N = nl.N;
  std::vector<uint8_t> reach(size_t(N), 0);  // reaches a MUX select input
  for (int64_t p = 0; p < N; ++p)
    if (nl.type[size_t(p)] == lsim::G_MUX && nl.fin_off[p + 1] - nl.fin_off[p] >= 3)
      reach[size_t(nl.fin[size_t(nl.fin_off[p] + 2)])] = 1;
  for (int64_t p = N - 1; p >= 0; --p)  // fanouts sit at higher positions
    for (int32_t e = fs.fo_off[p]; e < fs.fo_off[p + 1] && !reach[size_t(p)]; ++e)
      reach[size_t(p)] = reach[size_t(fs.fo[e])];
  for (size_t f = 0; f < nl.fault_site.size(); ++f)
    (reach[size_t(nl.fault_site[f])] ? direct : stem).push_back(int32_t(f));
}

void release_jobs(lsim_sim *s) {
  for (lsim::JobState **j : {&s->job, &s->job2})
    if (s->engine && *j) {
      s->engine->release_job(*j);
      *j = nullptr;
    }
  s->job_split = false;
}

// Results of the split jobs in the fault order of the inputs.
lsim::RunResults merge_split(lsim::RunResults &a, const std::vector<int32_t> &ia, lsim::RunResults &b,
                             const std::vector<int32_t> &ib, int64_t F) {
  lsim::RunResults out;
  out.good_last = std::move(a.good_last);
  out.RB = a.RB;
  out.has_matrix = a.has_matrix;
  out.has_last = a.has_last;
  out.first_round.resize(size_t(F));
  out.last_round.resize(size_t(F));
  if (out.has_matrix) out.det_bits.assign(size_t(F * out.RB), 0);
  for (int part = 0; part < 2; ++part) {
    const lsim::RunResults &r = part ? b : a;
    const std::vector<int32_t> &idx = part ? ib : ia;
    for (size_t k = 0; k < idx.size(); ++k) {
      const size_t f = size_t(idx[k]);
      out.first_round[f] = r.first_round[k];
      out.last_round[f] = r.has_last ? r.last_round[k] : uint8_t(0);
      if (out.has_matrix)
        std::memcpy(out.det_bits.data() + f * size_t(out.RB), r.det_bits.data() + k * size_t(r.RB),
                    size_t(out.RB) * sizeof(uint64_t));
    }
  }
  out.stats = a.stats;  // the two ran together: shared device times
  out.stats.live_faults += b.stats.live_faults;
  out.stats.stems += b.stats.stems;
  out.stats.sweep_bytes = std::max(a.stats.sweep_bytes, b.stats.sweep_bytes);
  return out;
}

int load_texts(lsim_sim *s, const char *ckt, size_t ckt_len, const char *flst, size_t flst_len, const char *ptn,
               size_t ptn_len) {
  return guarded(&s->err, [&] {
    s->loaded = false;
    s->ran = false;
    s->res = lsim::RunResults();
    release_jobs(s);
    if (s->engine && s->circuit) {
      s->engine->remove_circuit(s->circuit);
      s->circuit = nullptr;
    }
    lsim::read_netlist(s->nl, ckt, ckt_len, flst, flst_len, ptn, ptn_len);
    lsim::build_fault_structure(s->nl, s->fs);
    mux_split(s->nl, s->fs, s->split_stem, s->split_direct);
    if (s->engine) s->circuit = s->engine->add_circuit(s->nl, s->fs);
    s->loaded = true;
  });
}

int check_loaded(const lsim_sim *s) {
  if (!s) return LSIM_E_ARG;
  if (!s->loaded) return fail(&const_cast<lsim_sim *>(s)->err, LSIM_E_STATE, "no circuit has been read");
  return LSIM_OK;
}

int check_ran(const lsim_sim *s) {
  int rc = check_loaded(s);
  if (rc) return rc;
  if (!s->ran) return fail(&const_cast<lsim_sim *>(s)->err, LSIM_E_STATE, "run() has not completed");
  return LSIM_OK;
}

int set_engine_option(lsim::EngineOptions &eo, std::string *err, const std::string &k, int64_t value, bool &known) {
  known = true;
  if (k == "scratch_mb") {
    eo.scratch_mb = value < 0 ? 0 : value;
  } else if (k == "block_waves") {
    if (value < 1 || value > 8) return fail(err, LSIM_E_ARG, "block_waves must be 1..8");
    eo.block_waves = int(value);
  } else if (k == "pair_threads") {
    if (value != 0 && value != 256 && value != 512) return fail(err, LSIM_E_ARG, "pair_threads must be 0, 256 or 512");
    eo.pair_threads = int(value);
  } else if (k == "pool_cap") {
    eo.pool_cap = value < 0 ? 0 : value;
  } else if (k == "pair_nopend") {
    if (value < 0 || value > 2) return fail(err, LSIM_E_ARG, "pair_nopend must be 0, 1 or 2");
    eo.pair_nopend = int(value);
  } else if (k == "pair_nopend_pct") {
    if (value < 0 || value > 100) return fail(err, LSIM_E_ARG, "pair_nopend_pct must be 0..100");
    eo.pair_nopend_pct = int(value);
  } else if (k == "pair_fo_pre") {
    if (value < 0 || value > 1) return fail(err, LSIM_E_ARG, "pair_fo_pre must be 0 or 1");
    eo.pair_fo_pre = value != 0;
  } else if (k == "max_step_words") {
    if (value < 1 || value > 4096) return fail(err, LSIM_E_ARG, "max_step_words must be 1..4096");
    eo.max_step_words = int(value);
  } else if (k == "pair_dense") {
    if (value < 0 || value > 64) return fail(err, LSIM_E_ARG, "pair_dense must be 0..64");
    eo.pair_dense = int(value);
  } else if (k == "autotune") {
    eo.autotune = value != 0;
  } else if (k == "step_graphs") {
    eo.step_graphs = value != 0;
  } else if (k == "device_steps") {
    if (value < 0 || value > 1) return fail(err, LSIM_E_ARG, "device_steps must be 0 or 1");
    eo.device_steps = value != 0;
  } else if (k == "good_form") {
    if (value < 0 || value > 1) return fail(err, LSIM_E_ARG, "good_form must be 0 or 1");
    eo.good_form = int(value);
  } else if (k == "win_levels") {
    if (value < 1 || value > 8) return fail(err, LSIM_E_ARG, "win_levels must be 1..8");
    eo.win_levels = int(value);
  } else if (k == "win_cap") {
    if (value < 0 || value > 4096) return fail(err, LSIM_E_ARG, "win_cap must be 0..4096");
    eo.win_cap = int(value);
  } else if (k == "good_graph") {
    eo.good_graph = value != 0;
  } else if (k == "sweep_kernel") {
    if (value < 0 || value > 1) return fail(err, LSIM_E_ARG, "sweep_kernel must be 0 or 1");
    eo.sweep_kernel = int(value);
  } else if (k == "slot_pct") {
    if (value < 1 || value > 100) return fail(err, LSIM_E_ARG, "slot_pct must be 1..100");
    eo.slot_share = double(value) / 100.0;
  } else {
    known = false;
  }
  return LSIM_OK;
}

}  // namespace

extern "C" {

int lsim_create(int device, lsim_sim **out) {
  if (!out) return LSIM_E_ARG;
  *out = nullptr;
  std::unique_ptr<lsim_sim> s(new (std::nothrow) lsim_sim);
  if (!s) return LSIM_E_NOMEM;
  s->device = device;
  if (device >= 0) {
    // Set-up of the device's no-allocation scratch (its address, looked up
    // once per device) happens here, with the engine.
    int rc = guarded(&s->err, [&] {
      s->engine.reset(new lsim::Engine(device));
      s->scratch = lsim::noalloc_scratch(device);
    });
    if (rc) {
      g_create_error = s->err;
      return rc;
    }
  }
  *out = s.release();
  return LSIM_OK;
}

void lsim_destroy(lsim_sim *s) {
  if (!s) return;
  try {
    s->engine.reset();
  } catch (...) {
  }
  delete s;
}

const char *lsim_last_error(const lsim_sim *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int lsim_read_files(lsim_sim *s, const char *ckt_path, const char *flst_path, const char *ptn_path) {
  if (!s || !ckt_path || !flst_path || !ptn_path) return LSIM_E_ARG;
  std::string c, f, p;
  // Messages as in fsim/simulation.cpp:24-32.
  if (!read_file(ckt_path, c)) return fail(&s->err, LSIM_E_IO, std::string("cannot open circut file ") + ckt_path);
  if (!read_file(flst_path, f)) return fail(&s->err, LSIM_E_IO, std::string("cannot open fault file ") + flst_path);
  if (!read_file(ptn_path, p)) return fail(&s->err, LSIM_E_IO, std::string("cannot open pattern file ") + ptn_path);
  return load_texts(s, c.data(), c.size(), f.data(), f.size(), p.data(), p.size());
}

int lsim_read_buffers(lsim_sim *s, const char *ckt, size_t ckt_len, const char *flst, size_t flst_len, const char *ptn,
                      size_t ptn_len) {
  if (!s || (!ckt && ckt_len) || (!flst && flst_len) || (!ptn && ptn_len)) return LSIM_E_ARG;
  return load_texts(s, ckt, ckt_len, flst, flst_len, ptn, ptn_len);
}

int lsim_set_option(lsim_sim *s, const char *key, int64_t value) {
  if (!s || !key) return LSIM_E_ARG;
  const std::string k(key);
  bool known = false;
  int rc = set_engine_option(s->eopt, &s->err, k, value, known);
  if (rc || known) return rc;
  if (k == "drop") s->jopt.drop = value != 0;
  else if (k == "report_last_round") s->jopt.report_last_round = value != 0;
  else if (k == "direct") s->jopt.direct = value != 0;
  else if (k == "first_words" || k == "word_growth") {
    if (value < 1 || value > 64) return fail(&s->err, LSIM_E_ARG, k + " must be 1..64");
    (k == "first_words" ? s->jopt.first_words : s->jopt.word_growth) = int(value);
  } else return fail(&s->err, LSIM_E_ARG, "unknown option " + k);
  return LSIM_OK;
}

static int engine_of_mode(lsim_sim *s, int mode, int *m) {
  switch (mode) {
    case LSIM_MODE_REF_SEQUENTIAL:
    case LSIM_MODE_REF_PARALLEL_TF:
    case LSIM_MODE_REF_PARALLEL_OMP:
    case LSIM_MODE_PARTITIONED: *m = lsim::ENGINE_PERSISTENT; return LSIM_OK;
    case LSIM_MODE_PLAIN: *m = lsim::ENGINE_PLAIN; return LSIM_OK;
    default: return fail(&s->err, LSIM_E_ARG, "unknown mode " + std::to_string(mode));
  }
}

int lsim_run(lsim_sim *s, int mode, size_t num_threads) {
  (void)num_threads;
  int rc = check_loaded(s);
  if (rc) return rc;
  if (!s->engine) return fail(&s->err, LSIM_E_DEVICE, "handle was created without a device (device -1)");
  int m;
  if ((rc = engine_of_mode(s, mode, &m))) return rc;
  const int64_t allocs0 = lsim::device_allocations();
  return guarded(&s->err, [&] {
    s->ran = false;
    lsim::EngineOptions eo = s->eopt;
    eo.mode = m;
    lsim::JobOptions jo = s->jopt;
    jo.keep_good_last = true;
    // In a circuit with MUX the faults whose cone reaches no MUX select keep
    // stem tracing; the others are injected directly (mux_split).
    const bool split = !jo.direct && !s->split_stem.empty() && !s->split_direct.empty();
    if (split != s->job_split) release_jobs(s);
    if (s->nl.has_mux && s->split_direct.empty()) jo.mux_safe = true;
    lsim::JobOptions jd = jo;
    if (split) {
      jo.mux_safe = true;
      jd.direct = true;
    }
    // The fault lists and patterns stay resident between runs of the same
    // inputs; only a change of dropping or last-round reporting rebuilds them.
    for (lsim::JobState **j : {&s->job, &s->job2})
      if (*j && !s->engine->reset_job(*j, j == &s->job ? jo : jd)) {
        s->engine->release_job(*j);
        *j = nullptr;
      }
    auto subset = [&](const std::vector<int32_t> &idx, std::vector<int32_t> &site, std::vector<uint8_t> &sa) {
      site.resize(idx.size());
      sa.resize(idx.size());
      for (size_t k = 0; k < idx.size(); ++k) {
        site[k] = s->nl.fault_site[size_t(idx[k])];
        sa[k] = s->nl.fault_sa[size_t(idx[k])];
      }
    };
    if (!s->job) {
      if (split) {
        std::vector<int32_t> site;
        std::vector<uint8_t> sa;
        subset(s->split_stem, site, sa);
        s->job = s->engine->add_job(s->circuit, std::move(site), std::move(sa), s->nl.num_patterns, s->nl.R,
                                    s->nl.patterns, jo);
      } else {
        s->job = s->engine->add_job(s->circuit, s->nl.fault_site, s->nl.fault_sa, s->nl.num_patterns, s->nl.R,
                                    s->nl.patterns, jo);
      }
    }
    if (split && !s->job2) {
      std::vector<int32_t> site;
      std::vector<uint8_t> sa;
      subset(s->split_direct, site, sa);
      s->job2 = s->engine->add_job(s->circuit, std::move(site), std::move(sa), s->nl.num_patterns, s->nl.R,
                                   s->nl.patterns, jd);
    }
    s->job_split = split;
    try {
      if (split) {
        s->engine->run({s->job, s->job2}, eo);
        s->res = merge_split(s->engine->results(s->job), s->split_stem, s->engine->results(s->job2),
                             s->split_direct, int64_t(s->nl.fault_site.size()));
      } else {
        s->engine->run({s->job}, eo);
        s->res = std::move(s->engine->results(s->job));
      }
    } catch (...) {
      release_jobs(s);
      throw;
    }
    s->res.stats.device_allocs = lsim::device_allocations() - allocs0;
    s->ran = true;
  });
}

int lsim_run_noalloc(lsim_sim *s, int mode, size_t num_threads) {
  (void)num_threads;
  int rc = check_loaded(s);
  if (rc) return rc;
  if (!s->engine || !s->circuit)
    return fail(&s->err, LSIM_E_DEVICE, "handle was created without a device (device -1)");
  int m;
  if ((rc = engine_of_mode(s, mode, &m))) return rc;
  if (!s->scratch) return fail(&s->err, LSIM_E_STATE, "no-allocation scratch was not set up for this device");
  const int64_t need = lsim::Engine::noalloc_min_bytes(s->circuit, int64_t(s->nl.fault_site.size()), s->nl.R,
                                                       s->jopt.drop && s->jopt.report_last_round, !s->jopt.drop);
  if (need > int64_t(lsim::kNoAllocScratchBytes))
    return fail(&s->err, LSIM_E_NOMEM,
                "this run needs at least " + std::to_string((need + (1 << 20) - 1) >> 20) +
                    " MiB of device memory; the no-allocation scratch holds " +
                    std::to_string(lsim::kNoAllocScratchBytes >> 20) + " MiB");
  const int64_t allocs0 = lsim::device_allocations();
  return guarded(&s->err, [&] {
    s->ran = false;
    lsim::EngineOptions eo = s->eopt;
    eo.mode = m;
    eo.good_graph = false;  // a graph's instantiation allocates
    eo.autotune = false;
    lsim::JobOptions jo = s->jopt;
    jo.keep_good_last = true;
    // A fresh engine per call, carving every buffer from the scratch from
    // its start; it borrows the handle engine's stream and events.
    lsim::Engine e(*s->engine, s->scratch, lsim::kNoAllocScratchBytes);
    lsim::CircuitRes *c = e.add_circuit(s->nl, s->fs);
    lsim::JobState *j =
        e.add_job(c, s->nl.fault_site, s->nl.fault_sa, s->nl.num_patterns, s->nl.R, s->nl.patterns, jo);
    e.run({j}, eo);
    s->res = std::move(e.results(j));
    s->res.stats.device_allocs = lsim::device_allocations() - allocs0;
    s->ran = true;
  });
}

int64_t lsim_num_pis(const lsim_sim *s) { return s && s->loaded ? s->nl.P : -1; }
int64_t lsim_num_pos(const lsim_sim *s) { return s && s->loaded ? s->nl.O : -1; }
int64_t lsim_num_gates(const lsim_sim *s) { return s && s->loaded ? s->nl.G : -1; }
int64_t lsim_num_wires(const lsim_sim *s) { return s && s->loaded ? s->nl.W : -1; }
int64_t lsim_num_rounds(const lsim_sim *s) { return s && s->loaded ? s->nl.R : -1; }
int64_t lsim_num_faults(const lsim_sim *s) { return s && s->loaded ? int64_t(s->nl.fault_site.size()) : -1; }
int64_t lsim_num_levels(const lsim_sim *s) { return s && s->loaded ? int64_t(s->nl.gates_per_level.size()) : -1; }
int64_t lsim_num_patterns(const lsim_sim *s) { return s && s->loaded ? s->nl.num_patterns : -1; }

int lsim_gate_idx_to_tpg_order_idx(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(&const_cast<lsim_sim *>(s)->err, out, n, size_t(s->nl.N)))) return rc;
  for (size_t i = 0; i < n; ++i) out[i] = uint64_t(s->nl.order_idx[i]);
  return LSIM_OK;
}

int lsim_num_gates_per_level(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(&const_cast<lsim_sim *>(s)->err, out, n, s->nl.gates_per_level.size()))) return rc;
  for (size_t i = 0; i < n; ++i) out[i] = uint64_t(s->nl.gates_per_level[i]);
  return LSIM_OK;
}

int lsim_levels(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(&const_cast<lsim_sim *>(s)->err, out, n, size_t(s->nl.N)))) return rc;
  for (size_t i = 0; i < n; ++i) out[i] = uint64_t(s->nl.level[i]);
  return LSIM_OK;
}

int lsim_faults(const lsim_sim *s, uint64_t *site, uint64_t *value, size_t n) {
  int rc = check_loaded(s);
  std::string *e = s ? &const_cast<lsim_sim *>(s)->err : nullptr;
  if (rc || (rc = check_n(e, site, n, s->nl.fault_site.size())) || (rc = check_n(e, value, n, n))) return rc;
  for (size_t i = 0; i < n; ++i) {
    site[i] = uint64_t(s->nl.fault_site[i]);
    value[i] = s->nl.fault_sa[i] ? ~uint64_t(0) : 0;
  }
  return LSIM_OK;
}

int lsim_patterns(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(&const_cast<lsim_sim *>(s)->err, out, n, s->nl.patterns.size()))) return rc;
  if (n) std::memcpy(out, s->nl.patterns.data(), n * sizeof(uint64_t));
  return LSIM_OK;
}

int lsim_good_results(const lsim_sim *s, uint64_t *pi, size_t n_pi, uint64_t *gate, size_t n_gate, uint64_t *po,
                      size_t n_po) {
  int rc = check_ran(s);
  std::string *e = s ? &const_cast<lsim_sim *>(s)->err : nullptr;
  if (rc || (rc = check_n(e, pi, n_pi, size_t(s->nl.P))) || (rc = check_n(e, gate, n_gate, size_t(s->nl.G))) ||
      (rc = check_n(e, po, n_po, size_t(s->nl.O))))
    return rc;
  const auto &g = s->res.good_last;
  for (size_t i = 0; i < n_pi; ++i) pi[i] = g[i];
  for (size_t i = 0; i < n_gate; ++i) gate[i] = g[s->nl.P + i];
  for (size_t i = 0; i < n_po; ++i) po[i] = g[s->nl.P + s->nl.G + i];
  return LSIM_OK;
}

int lsim_found_fault_to_pattern(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_ran(s);
  const size_t F = s ? s->nl.fault_site.size() : 0;
  std::string *e = s ? &const_cast<lsim_sim *>(s)->err : nullptr;
  if (rc || (rc = check_n(e, out, n, 2 * F))) return rc;
  if (!s->res.has_last)
    return fail(e, LSIM_E_STATE, "the last round was not evaluated for every fault (dropping on, report_last_round off)");
  fill_table(s->res, int64_t(F), s->nl.R, out);
  return LSIM_OK;
}

int lsim_detected(const lsim_sim *s, uint8_t *out, size_t n) {
  int rc = check_ran(s);
  if (rc || (rc = check_n(&const_cast<lsim_sim *>(s)->err, out, n, s->nl.fault_site.size()))) return rc;
  for (size_t f = 0; f < n; ++f) out[f] = s->res.first_round[f] >= 0;
  return LSIM_OK;
}

int lsim_first_detect_round(const lsim_sim *s, int64_t *out, size_t n) {
  int rc = check_ran(s);
  if (rc || (rc = check_n(&const_cast<lsim_sim *>(s)->err, out, n, s->nl.fault_site.size()))) return rc;
  for (size_t f = 0; f < n; ++f) out[f] = s->res.first_round[f];
  return LSIM_OK;
}

int lsim_detect_matrix(const lsim_sim *s, uint8_t *out, size_t n) {
  int rc = check_ran(s);
  const size_t F = s ? s->nl.fault_site.size() : 0;
  const int64_t R = s ? s->nl.R : 0;
  std::string *e = s ? &const_cast<lsim_sim *>(s)->err : nullptr;
  if (rc || (rc = check_n(e, out, n, F * size_t(R)))) return rc;
  if (!s->res.has_matrix) return fail(e, LSIM_E_STATE, "the run used fault dropping; no per-round matrix");
  fill_matrix(s->res, int64_t(F), R, out);
  return LSIM_OK;
}

int lsim_last_run_stats(const lsim_sim *s, lsim_run_stats *out) {
  int rc = check_ran(s);
  if (rc) return rc;
  if (!out) return LSIM_E_ARG;
  *out = s->res.stats;
  return LSIM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------- batching

struct BatchJob {
  int64_t circuit = -1;
  int64_t F = 0, R = 0;
  lsim::JobState *js = nullptr;
  bool finished = false;
  lsim::RunResults res;
};

struct lsim_batch {
  std::unique_ptr<lsim::Engine> engine;
  lsim::EngineOptions eopt;
  std::vector<std::unique_ptr<ResidentCircuit>> circuits;
  std::map<int64_t, BatchJob> jobs;
  int64_t next_job = 0;
  lsim_run_stats stats{};
  std::string err;
};

extern "C" {

int lsim_batch_create(int device, lsim_batch **out) {
  if (!out) return LSIM_E_ARG;
  *out = nullptr;
  std::unique_ptr<lsim_batch> b(new (std::nothrow) lsim_batch);
  if (!b) return LSIM_E_NOMEM;
  int rc = guarded(&b->err, [&] { b->engine.reset(new lsim::Engine(device)); });
  if (rc) {
    g_create_error = b->err;
    return rc;
  }
  *out = b.release();
  return LSIM_OK;
}

void lsim_batch_destroy(lsim_batch *b) {
  if (!b) return;
  try {
    b->engine.reset();
  } catch (...) {
  }
  delete b;
}

const char *lsim_batch_last_error(const lsim_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int lsim_batch_set_option(lsim_batch *b, const char *key, int64_t value) {
  if (!b || !key) return LSIM_E_ARG;
  const std::string k(key);
  bool known = false;
  int rc = set_engine_option(b->eopt, &b->err, k, value, known);
  if (rc || known) return rc;
  if (k == "mode") {
    if (value != 0 && value != 1) return fail(&b->err, LSIM_E_ARG, "mode must be 0 or 1");
    b->eopt.mode = value ? lsim::ENGINE_PERSISTENT : lsim::ENGINE_PLAIN;
    return LSIM_OK;
  }
  return fail(&b->err, LSIM_E_ARG, "unknown option " + k);
}

int lsim_batch_add_circuit(lsim_batch *b, const char *ckt, size_t ckt_len, int64_t *circuit_id) {
  if (!b || (!ckt && ckt_len) || !circuit_id) return LSIM_E_ARG;
  return guarded(&b->err, [&] {
    auto rc = std::make_unique<ResidentCircuit>();
    lsim::read_circuit(rc->nl, ckt, ckt_len);
    lsim::build_fault_structure(rc->nl, rc->fs);
    rc->res = b->engine->add_circuit(rc->nl, rc->fs);
    *circuit_id = int64_t(b->circuits.size());
    b->circuits.push_back(std::move(rc));
  });
}

int lsim_batch_add_job(lsim_batch *b, int64_t circuit_id, const char *flst, size_t flst_len, const char *ptn,
                       size_t ptn_len, int drop, int64_t *job_id) {
  if (!b || !job_id || (!flst && flst_len) || (!ptn && ptn_len)) return LSIM_E_ARG;
  if (circuit_id < 0 || circuit_id >= int64_t(b->circuits.size())) return fail(&b->err, LSIM_E_ARG, "unknown circuit");
  return guarded(&b->err, [&] {
    const ResidentCircuit &c = *b->circuits[size_t(circuit_id)];
    std::vector<int32_t> site;
    std::vector<uint8_t> sa;
    int64_t n = 0, R = 0;
    std::vector<uint64_t> words;
    lsim::read_faults(c.nl, flst, flst_len, site, sa);
    lsim::read_patterns(c.nl, ptn, ptn_len, n, R, words);
    lsim::JobOptions jo;
    jo.drop = drop != 0;
    jo.report_last_round = true;
    BatchJob bj;
    bj.circuit = circuit_id;
    bj.F = int64_t(site.size());
    bj.R = R;
    bj.js = b->engine->add_job(c.res, std::move(site), std::move(sa), n, R, words, jo);
    *job_id = b->next_job++;
    b->jobs[*job_id] = std::move(bj);
  });
}

int lsim_batch_run(lsim_batch *b) {
  if (!b) return LSIM_E_ARG;
  return guarded(&b->err, [&] {
    std::vector<lsim::JobState *> js;
    std::vector<BatchJob *> owners;
    for (auto &kv : b->jobs)
      if (!kv.second.finished) {
        js.push_back(kv.second.js);
        owners.push_back(&kv.second);
      }
    const auto t0 = std::chrono::steady_clock::now();
    b->engine->run(js, b->eopt);
    b->stats = lsim_run_stats{};
    b->stats.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (BatchJob *o : owners) {
      o->res = std::move(b->engine->results(o->js));
      b->engine->release_job(o->js);
      o->js = nullptr;
      o->finished = true;
      const lsim_run_stats &s = o->res.stats;
      // Device phases are shared by the batch: keep the largest.
      b->stats.t_good_ms = std::max(b->stats.t_good_ms, s.t_good_ms);
      b->stats.t_sweep_ms = std::max(b->stats.t_sweep_ms, s.t_sweep_ms);
      b->stats.t_detect_ms = std::max(b->stats.t_detect_ms, s.t_detect_ms);
      b->stats.live_faults += s.live_faults;
      b->stats.stems += s.stems;
      b->stats.steps = std::max(b->stats.steps, s.steps);
      b->stats.sweep_launches = std::max(b->stats.sweep_launches, s.sweep_launches);
      b->stats.sweep_tiles = std::max(b->stats.sweep_tiles, s.sweep_tiles);
      b->stats.sweep_node_steps = std::max(b->stats.sweep_node_steps, s.sweep_node_steps);
      b->stats.sweep_bytes = std::max(b->stats.sweep_bytes, s.sweep_bytes);
      b->stats.tune_trials = s.tune_trials;
      b->stats.tune_candidates = s.tune_candidates;
      b->stats.good_replayed = s.good_replayed;
      b->stats.steps_replayed = std::max(b->stats.steps_replayed, s.steps_replayed);
      b->stats.sweep_reruns = std::max(b->stats.sweep_reruns, s.sweep_reruns);
      b->stats.tuned = s.tuned;
      b->stats.engine = s.engine;
      b->stats.slots = s.slots;
      b->stats.good_bytes += s.good_bytes;
    }
  });
}

static const BatchJob *batch_job(const lsim_batch *b, int64_t id, bool need_finished) {
  if (!b) return nullptr;
  auto it = b->jobs.find(id);
  if (it == b->jobs.end()) return nullptr;
  if (need_finished && !it->second.finished) return nullptr;
  return &it->second;
}

int64_t lsim_batch_job_faults(const lsim_batch *b, int64_t job_id) {
  const BatchJob *j = batch_job(b, job_id, false);
  return j ? j->F : -1;
}

int64_t lsim_batch_job_rounds(const lsim_batch *b, int64_t job_id) {
  const BatchJob *j = batch_job(b, job_id, false);
  return j ? j->R : -1;
}

int lsim_batch_first_detect_round(const lsim_batch *b, int64_t job_id, int64_t *out, size_t n) {
  const BatchJob *j = batch_job(b, job_id, true);
  if (!j) return LSIM_E_STATE;
  int rc = check_n(&const_cast<lsim_batch *>(b)->err, out, n, size_t(j->F));
  if (rc) return rc;
  for (size_t f = 0; f < n; ++f) out[f] = j->res.first_round[f];
  return LSIM_OK;
}

int lsim_batch_detect_matrix(const lsim_batch *b, int64_t job_id, uint8_t *out, size_t n) {
  const BatchJob *j = batch_job(b, job_id, true);
  if (!j) return LSIM_E_STATE;
  std::string *e = &const_cast<lsim_batch *>(b)->err;
  int rc = check_n(e, out, n, size_t(j->F * j->R));
  if (rc) return rc;
  if (!j->res.has_matrix) return fail(e, LSIM_E_STATE, "the job used fault dropping; no per-round matrix");
  fill_matrix(j->res, j->F, j->R, out);
  return LSIM_OK;
}

int lsim_batch_found_fault_to_pattern(const lsim_batch *b, int64_t job_id, uint64_t *out, size_t n) {
  const BatchJob *j = batch_job(b, job_id, true);
  if (!j) return LSIM_E_STATE;
  std::string *e = &const_cast<lsim_batch *>(b)->err;
  int rc = check_n(e, out, n, size_t(2 * j->F));
  if (rc) return rc;
  if (!j->res.has_last) return fail(e, LSIM_E_STATE, "the last round was not evaluated for every fault");
  fill_table(j->res, j->F, j->R, out);
  return LSIM_OK;
}

int lsim_batch_last_run_stats(const lsim_batch *b, lsim_run_stats *out) {
  if (!b || !out) return LSIM_E_ARG;
  *out = b->stats;
  return LSIM_OK;
}

int lsim_batch_release_job(lsim_batch *b, int64_t job_id) {
  if (!b) return LSIM_E_ARG;
  auto it = b->jobs.find(job_id);
  if (it == b->jobs.end()) return fail(&b->err, LSIM_E_ARG, "unknown job");
  return guarded(&b->err, [&] {
    if (it->second.js) b->engine->release_job(it->second.js);
    b->jobs.erase(it);
  });
}

}  // extern "C"

// -------------------------------------------------------------------- serving

struct ServerJob {
  int64_t id = 0;
  int64_t circuit = -1;
  int lane = 0;
  std::vector<int32_t> site;
  std::vector<uint8_t> sa;
  int64_t n = 0, R = 0;
  std::vector<uint64_t> words;
  lsim::JobOptions opt;
  int64_t bytes = 0;
  lsim::JobState *js = nullptr;
  double t_submit = 0, t_admit = 0, t_finish = 0;
  int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
  int code = 0;
  std::string error;
  lsim::RunResults res;
};

// A lane: an engine on its own HIP stream with its own worker thread. Lane 0
// batches small jobs continuously; lane 1 runs large jobs one at a time in
// arrival order. The lanes run concurrently on the device, each sweep taking
// a share of the workgroup slots while the other lane has work.
struct ServerLane {
  std::unique_ptr<lsim::Engine> engine;
  lsim::EngineOptions eopt;
  std::mutex engine_mu;  // every call on this lane's engine
  std::condition_variable cv_work;
  std::deque<ServerJob *> queue;
  std::vector<ServerJob *> running;
  int max_running = 0;  // 0: no limit
  double busy_share = 1.0;  // slot share while the other lane has work
  std::thread worker;
};

struct lsim_server {
  ServerLane lane[2];
  int64_t budget = 0, used = 0;  // HBM budget; held by resident circuits and admitted jobs
  int64_t circuit_bytes = 0;     // held by resident circuits (one device copy for both lanes)
  std::mutex mu;  // queues, job table, counters, budget
  std::condition_variable cv_done;
  std::vector<std::unique_ptr<ResidentCircuit>> circuits;
  std::map<int64_t, std::unique_ptr<ServerJob>> jobs;
  int64_t next_id = 0;
  int64_t iterations = 0, finished = 0, failed = 0, peak = 0;
  int first_words = 1, word_growth = 4;
  double big_evals = 4e9;       // fault x pattern evaluations from which a job is large (<= 0: one lane)
  double big_node_words = 5e7;  // circuit nodes x rounds from which a job is large
  bool stop = false;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::string err;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  bool lane_has_work(int l) const { return !lane[l].queue.empty() || !lane[l].running.empty(); }
  void loop(int l);
};

void lsim_server::loop(int l) {
  ServerLane &L = lane[l];
  for (;;) {
    std::vector<ServerJob *> admit;
    lsim::EngineOptions eo;
    {
      std::unique_lock<std::mutex> lk(mu);
      L.cv_work.wait(lk, [&] {
        if (stop || !L.running.empty()) return true;
        return !L.queue.empty() && used + L.queue.front()->bytes <= budget;
      });
      if (stop) return;
      // Admission in arrival order while the footprint fits the budget (a
      // job larger than the whole budget was failed at submit).
      while (!L.queue.empty() && (L.max_running == 0 || int(L.running.size() + admit.size()) < L.max_running)) {
        ServerJob *j = L.queue.front();
        if (used + j->bytes > budget) break;
        used += j->bytes;
        admit.push_back(j);
        L.queue.pop_front();
      }
      eo = L.eopt;
      // The lane's share of the slots while the other lane has work, within
      // the engine option's cap (slot_pct).
      eo.slot_share = L.eopt.slot_share * (lane_has_work(1 - l) ? L.busy_share : 1.0);
    }
    std::vector<ServerJob *> done_now;
    std::vector<ServerJob *> failed_now;
    {
      std::lock_guard<std::mutex> elk(L.engine_mu);
      std::vector<ServerJob *> fresh;
      for (ServerJob *j : admit) {
        const ResidentCircuit &c = *circuits[size_t(j->circuit)];
        int rc = guarded(&j->error, [&] {
          j->js = L.engine->add_job(l == 0 ? c.res : c.res2, std::move(j->site), std::move(j->sa), j->n, j->R,
                                    j->words, j->opt);
        });
        j->words.clear();
        j->words.shrink_to_fit();
        if (rc) {
          j->code = rc;
          failed_now.push_back(j);
        } else {
          j->t_admit = now();
          fresh.push_back(j);
        }
      }
      std::vector<lsim::JobState *> js;
      std::vector<ServerJob *> live;
      for (ServerJob *j : L.running) live.push_back(j);
      for (ServerJob *j : fresh) live.push_back(j);
      for (ServerJob *j : live) js.push_back(j->js);
      std::string msg;
      int rc = guarded(&msg, [&] {
        std::vector<lsim::JobState *> fs;
        for (ServerJob *j : fresh) fs.push_back(j->js);
        L.engine->start(fs, eo);
        L.engine->iterate(js, eo);
      });
      std::vector<ServerJob *> still;
      for (ServerJob *j : live) {
        if (rc) {
          // The iteration failed. Each of its jobs runs again alone from its
          // start, so that one job's failure does not fail the others; a job
          // that fails alone too reports its own error.
          std::string msg1;
          const int rc1 = guarded(&msg1, [&] {
            if (!L.engine->reset_job(j->js, j->opt)) throw lsim::DeviceError(LSIM_E_STATE, msg);
            lsim::EngineOptions alone = eo;
            alone.autotune = false;
            L.engine->run({j->js}, alone);
          });
          if (rc1) {
            j->code = rc1;
            j->error = msg1;
            L.engine->release_job(j->js);
            j->js = nullptr;
            failed_now.push_back(j);
          } else {
            j->res = std::move(L.engine->results(j->js));
            L.engine->release_job(j->js);
            j->js = nullptr;
            done_now.push_back(j);
          }
        } else if (L.engine->done(j->js)) {
          j->res = std::move(L.engine->results(j->js));
          L.engine->release_job(j->js);
          j->js = nullptr;
          done_now.push_back(j);
        } else {
          still.push_back(j);
        }
      }
      std::lock_guard<std::mutex> lk(mu);
      L.running.swap(still);
      peak = std::max<int64_t>(peak, int64_t(lane[0].running.size() + lane[1].running.size() + done_now.size()));
      ++iterations;
      const double t = now();
      for (ServerJob *j : done_now) {
        j->state = 2;
        j->t_finish = t;
        used -= j->bytes;
        ++finished;
      }
      for (ServerJob *j : failed_now) {
        j->state = 3;
        j->t_finish = t;
        used -= j->bytes;
        ++failed;
      }
    }
    if (!done_now.empty() || !failed_now.empty()) {
      cv_done.notify_all();
      lane[1 - l].cv_work.notify_all();  // budget freed for the other lane
    }
  }
}

extern "C" {

int lsim_server_create(int device, int64_t job_budget_mb, int64_t scratch_mb, lsim_server **out) {
  if (!out || job_budget_mb <= 0) return LSIM_E_ARG;
  *out = nullptr;
  std::unique_ptr<lsim_server> s(new (std::nothrow) lsim_server);
  if (!s) return LSIM_E_NOMEM;
  int rc = guarded(&s->err, [&] {
    for (ServerLane &L : s->lane) L.engine.reset(new lsim::Engine(device));
    s->budget = job_budget_mb << 20;
    int64_t scratch = scratch_mb;
    if (scratch <= 0) {
      // Free HBM less the job budget and a 4 GiB margin for resident circuits.
      const int64_t free_mb = s->lane[0].engine->free_bytes() >> 20;
      scratch = std::max<int64_t>(2048, (free_mb - job_budget_mb - 4096) / 10 * 9);
    }
    // Small jobs: 30 % of the scratch and a quarter of the slots while large
    // jobs run; large jobs: 70 % and the other three quarters.
    const int64_t floor_mb = scratch_mb > 0 ? 32 : 512;  // an explicit scratch size is taken as given
    s->lane[0].eopt.scratch_mb = std::max<int64_t>(floor_mb, scratch * 3 / 10);
    s->lane[1].eopt.scratch_mb = std::max<int64_t>(floor_mb, scratch - s->lane[0].eopt.scratch_mb);
    s->lane[0].busy_share = 0.25;
    s->lane[1].busy_share = 0.75;
    s->lane[1].max_running = 1;
    for (ServerLane &L : s->lane) {
      L.eopt.mode = lsim::ENGINE_PERSISTENT;
      // Serving jobs rarely repeat a circuit's round count, so a captured
      // good-machine graph would seldom be replayed.
      L.eopt.good_graph = false;
    }
    for (int l = 0; l < 2; ++l) s->lane[l].worker = std::thread([p = s.get(), l] { p->loop(l); });
  });
  if (rc) {
    g_create_error = s->err;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    for (ServerLane &L : s->lane) {
      L.cv_work.notify_all();
      if (L.worker.joinable()) L.worker.join();
    }
    return rc;
  }
  *out = s.release();
  return LSIM_OK;
}

void lsim_server_destroy(lsim_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  for (ServerLane &L : s->lane) {
    L.cv_work.notify_all();
    if (L.worker.joinable()) L.worker.join();
  }
  try {
    for (int l = 0; l < 2; ++l) {
      std::lock_guard<std::mutex> elk(s->lane[l].engine_mu);
      for (auto &kv : s->jobs)
        if (kv.second->js && kv.second->lane == l) s->lane[l].engine->release_job(kv.second->js);
      s->lane[l].engine.reset();
    }
  } catch (...) {
  }
  delete s;
}

const char *lsim_server_last_error(const lsim_server *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int lsim_server_add_circuit(lsim_server *s, const char *ckt, size_t ckt_len, int64_t *circuit_id) {
  if (!s || (!ckt && ckt_len) || !circuit_id) return LSIM_E_ARG;
  auto rc_ptr = std::make_unique<ResidentCircuit>();
  std::string msg;
  int rc = guarded(&msg, [&] {
    lsim::read_circuit(rc_ptr->nl, ckt, ckt_len);
    lsim::build_fault_structure(rc_ptr->nl, rc_ptr->fs);
    {
      std::lock_guard<std::mutex> elk(s->lane[0].engine_mu);
      rc_ptr->res = s->lane[0].engine->add_circuit(rc_ptr->nl, rc_ptr->fs);
    }
    // One device copy serves both lanes (lane 0 owns it).
    std::lock_guard<std::mutex> elk(s->lane[1].engine_mu);
    rc_ptr->res2 = s->lane[1].engine->share_circuit(rc_ptr->res);
  });
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc) {
    s->err = msg;
    return rc;
  }
  // The resident circuit counts against the HBM budget that admission uses.
  const int64_t cb = s->lane[0].engine->circuit_bytes(rc_ptr->res);
  s->used += cb;
  s->circuit_bytes += cb;
  *circuit_id = int64_t(s->circuits.size());
  s->circuits.push_back(std::move(rc_ptr));
  return LSIM_OK;
}

int lsim_server_submit(lsim_server *s, int64_t circuit_id, const char *flst, size_t flst_len, const char *ptn,
                       size_t ptn_len, int drop, int64_t *job_id) {
  if (!s || !job_id || (!flst && flst_len) || (!ptn && ptn_len)) return LSIM_E_ARG;
  const ResidentCircuit *c;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (circuit_id < 0 || circuit_id >= int64_t(s->circuits.size())) return fail(&s->err, LSIM_E_ARG, "unknown circuit");
    c = s->circuits[size_t(circuit_id)].get();
  }
  auto j = std::make_unique<ServerJob>();
  std::string msg;
  int rc = guarded(&msg, [&] {
    lsim::read_faults(c->nl, flst, flst_len, j->site, j->sa);
    lsim::read_patterns(c->nl, ptn, ptn_len, j->n, j->R, j->words);
  });
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc) {
    s->err = msg;
    return rc;
  }
  j->circuit = circuit_id;
  j->opt.drop = drop != 0;
  j->opt.report_last_round = false;
  // Fewer, larger word steps than a lone run: every iteration a job waits
  // for is shared with the other admitted jobs.
  j->opt.first_words = s->first_words;
  j->opt.word_growth = s->word_growth;
  j->bytes = lsim::Engine::job_bytes(c->res, int64_t(j->site.size()), j->R, j->opt.drop);
  j->t_submit = s->now();
  j->id = s->next_id++;
  *job_id = j->id;
  // Large: many fault x pattern evaluations, or a large good machine (nodes
  // x rounds), which also sets how far every sweep tile walks.
  const double evals = double(j->site.size()) * double(j->n);
  const double node_words = double(c->nl.N) * double(j->R);
  j->lane = (s->big_evals > 0 && (evals >= s->big_evals || node_words >= s->big_node_words)) ? 1 : 0;
  ServerJob *jp = j.get();
  s->jobs[j->id] = std::move(j);
  if (jp->bytes > s->budget - s->circuit_bytes) {
    // Larger than what the budget leaves beside the resident circuits: it
    // could never be admitted.
    jp->state = 3;
    jp->code = LSIM_E_NOMEM;
    jp->error = "job needs " + std::to_string(jp->bytes) + " bytes, above the server's job budget";
    jp->t_finish = s->now();
    ++s->failed;
    s->cv_done.notify_all();
    return LSIM_OK;
  }
  s->lane[jp->lane].queue.push_back(jp);
  s->lane[jp->lane].cv_work.notify_all();
  s->lane[1 - jp->lane].cv_work.notify_all();  // the other lane's slot share changes
  return LSIM_OK;
}

int lsim_server_wait(lsim_server *s, int64_t job_id, double timeout_s) {
  if (!s) return LSIM_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(&s->err, LSIM_E_ARG, "unknown job");
  ServerJob *j = it->second.get();
  auto ready = [&] { return j->state >= 2; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    return fail(&s->err, LSIM_E_STATE, "job not finished within the timeout");
  }
  if (j->state == 3) return fail(&s->err, j->code ? j->code : LSIM_E_DEVICE, j->error);
  return LSIM_OK;
}

int lsim_server_first_detect_round(lsim_server *s, int64_t job_id, int64_t *out, size_t n) {
  if (!s) return LSIM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end() || it->second->state != 2) return fail(&s->err, LSIM_E_STATE, "job not finished");
  const ServerJob &j = *it->second;
  int rc = check_n(&s->err, out, n, j.res.first_round.size());
  if (rc) return rc;
  for (size_t f = 0; f < n; ++f) out[f] = j.res.first_round[f];
  return LSIM_OK;
}

int64_t lsim_server_job_faults(lsim_server *s, int64_t job_id) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return -1;
  return it->second->state == 2 ? int64_t(it->second->res.first_round.size()) : int64_t(it->second->site.size());
}

int lsim_server_job_times(lsim_server *s, int64_t job_id, double *times) {
  if (!s || !times) return LSIM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(&s->err, LSIM_E_ARG, "unknown job");
  times[0] = it->second->t_submit;
  times[1] = it->second->t_admit;
  times[2] = it->second->t_finish;
  return LSIM_OK;
}

int lsim_server_release_job(lsim_server *s, int64_t job_id) {
  if (!s) return LSIM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(&s->err, LSIM_E_ARG, "unknown job");
  if (it->second->state < 2) return fail(&s->err, LSIM_E_STATE, "job still queued or running");
  s->jobs.erase(it);
  return LSIM_OK;
}

int lsim_server_set_option(lsim_server *s, const char *key, int64_t value) {
  if (!s || !key) return LSIM_E_ARG;
  const std::string k(key);
  std::lock_guard<std::mutex> elk0(s->lane[0].engine_mu);  // the workers read eopt under these locks
  std::lock_guard<std::mutex> elk1(s->lane[1].engine_mu);
  std::lock_guard<std::mutex> lk(s->mu);
  if (k == "first_words" || k == "word_growth") {
    if (value < 1 || value > 64) return fail(&s->err, LSIM_E_ARG, k + " must be 1..64");
    (k == "first_words" ? s->first_words : s->word_growth) = int(value);
    return LSIM_OK;
  }
  if (k == "big_evals") {
    s->big_evals = double(value);
    return LSIM_OK;
  }
  if (k == "big_node_words") {
    s->big_node_words = double(value);
    return LSIM_OK;
  }
  if (k == "small_share_pct") {
    if (value < 1 || value > 99) return fail(&s->err, LSIM_E_ARG, "small_share_pct must be 1..99");
    s->lane[0].busy_share = double(value) / 100.0;
    s->lane[1].busy_share = 1.0 - s->lane[0].busy_share;
    return LSIM_OK;
  }
  bool known = false;
  int rc = 0;
  for (ServerLane &L : s->lane) {
    const int64_t keep = L.eopt.scratch_mb;
    rc = set_engine_option(L.eopt, &s->err, k, value, known);
    if (k == "scratch_mb") L.eopt.scratch_mb = keep;  // the server splits its scratch at creation
    if (rc) return rc;
  }
  if (known) return LSIM_OK;
  return fail(&s->err, LSIM_E_ARG, "unknown option " + k);
}

int lsim_server_counters(lsim_server *s, int64_t *counters) {
  if (!s || !counters) return LSIM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  counters[0] = s->iterations;
  counters[1] = s->finished;
  counters[2] = s->failed;
  counters[3] = s->peak;
  return LSIM_OK;
}

// ------------------------------------------------------------------ generator

int lsim_generate(const lsim_gen_spec *spec, char **ckt, size_t *ckt_len, char **flst, size_t *flst_len, char **ptn,
                  size_t *ptn_len) {
  if (!spec || !ckt || !ckt_len || !flst || !flst_len || !ptn || !ptn_len) return LSIM_E_ARG;
  *ckt = *flst = *ptn = nullptr;
  std::string msg;
  int rc = guarded(&msg, [&] {
    lsim::GenSpec g;
    g.seed = spec->seed;
    g.num_pis = spec->num_pis;
    g.num_pos = spec->num_pos;
    g.num_gates = spec->num_gates;
    g.num_levels = spec->num_levels;
    g.fanout_cap = spec->fanout_cap;
    g.power_exp = spec->power_exp;
    g.locality = spec->locality;
    g.window = spec->window;
    g.unary_frac = spec->unary_frac;
    g.num_patterns = spec->num_patterns;
    std::string c = lsim::generate_ckt(g);
    std::string f = lsim::generate_flst(g.seed, g.num_pis, g.num_pos, g.num_gates,
                                        spec->fault_fraction > 0 ? spec->fault_fraction : 1.0);
    std::string p = lsim::generate_ptn(g.seed, g.num_pis, g.num_patterns);
    auto dup = [](const std::string &x, char **o, size_t *n) {
      *o = static_cast<char *>(std::malloc(x.size() + 1));
      if (!*o) throw std::bad_alloc();
      std::memcpy(*o, x.data(), x.size());
      (*o)[x.size()] = 0;
      *n = x.size();
    };
    dup(c, ckt, ckt_len);
    dup(f, flst, flst_len);
    dup(p, ptn, ptn_len);
  });
  if (rc) {
    g_create_error = msg;
    std::free(*ckt);
    std::free(*flst);
    std::free(*ptn);
    *ckt = *flst = *ptn = nullptr;
  }
  return rc;
}

void lsim_free_text(char *p) { std::free(p); }

static int emit_text(const std::string &x, char **o, size_t *n) {
  *o = static_cast<char *>(std::malloc(x.size() + 1));
  if (!*o) return LSIM_E_NOMEM;
  std::memcpy(*o, x.data(), x.size());
  (*o)[x.size()] = 0;
  *n = x.size();
  return LSIM_OK;
}

// Graphviz text of the circuit in the reference's layout
// (fsim/simulation.cpp:1157-1225): gates, then POs, one edge per fanin in pin
// order; PIs are named by position, gates by position among the gates and
// their type, POs by position among the POs.
static std::string dot_text(const lsim::Netlist &nl) {
  static const char *kNames[] = {"INV", "AND", "OR", "XOR", "NAND", "NOR", "XNOR", "MUX", "CLKBUF", "PI", "PO"};
  auto tname = [](uint8_t t) { return t < 11 ? kNames[t] : "UNKNOWN"; };
  const int64_t P = nl.P, G = nl.G;
  std::string o = "digraph G {\n\tsubgraph cluster_0 {\n\t\tcolor=black;\n\t\tnode [style=filled,color=lightgrey];\n";
  auto source = [&](int32_t q) {
    if (nl.type[size_t(q)] == lsim::G_PI) return "PI_" + std::to_string(q) + " -> ";
    return "Gate_" + std::to_string(int64_t(q) - P) + "_" + tname(nl.type[size_t(q)]) + " -> ";
  };
  for (int64_t p = P; p < P + G; ++p) {
    const std::string now = "Gate_" + std::to_string(p - P) + "_" + tname(nl.type[size_t(p)]) + ";";
    for (int32_t e = nl.fin_off[size_t(p)]; e < nl.fin_off[size_t(p) + 1]; ++e)
      o += "\t\t" + source(nl.fin[size_t(e)]) + now + "\n";
  }
  for (int64_t p = P + G; p < nl.N; ++p) {
    const std::string now = "PO_" + std::to_string(p - P - G) + ";";
    for (int32_t e = nl.fin_off[size_t(p)]; e < nl.fin_off[size_t(p) + 1]; ++e)
      o += "\t\t" + source(nl.fin[size_t(e)]) + now + "\n";
  }
  o += "\t}\n\n";
  for (int64_t i = 0; i < P; ++i) o += "\tPI_" + std::to_string(i) + "  [shape=Msquare];\n";
  for (int64_t i = 0; i < nl.O; ++i) o += "\tPO_" + std::to_string(i) + "  [shape=Msquare];\n";
  o += "}\n\n";
  return o;
}

int lsim_level_diagnostics(const lsim_sim *s, int wave, int64_t *nodes, double *avg_dist, int64_t *max_span,
                           double *avg_range, double *avg_unique, size_t n_levels) {
  int rc = check_loaded(s);
  if (rc) return rc;
  std::string *err = &const_cast<lsim_sim *>(s)->err;
  const lsim::Netlist &nl = s->nl;
  const size_t L = size_t(nl.max_level) + 1;
  if (wave < 1) return fail(err, LSIM_E_ARG, "wave must be positive");
  if (!nodes || !avg_dist || !max_span || !avg_range || !avg_unique || n_levels != L)
    return fail(err, LSIM_E_ARG, "level arrays must hold " + std::to_string(L) + " entries");
  std::vector<std::vector<int32_t>> by_level(L);
  for (int64_t p = 0; p < nl.N; ++p) by_level[size_t(nl.level[size_t(p)])].push_back(int32_t(p));
  std::unordered_set<int32_t> uniq;
  for (size_t l = 0; l < L; ++l) {
    const std::vector<int32_t> &v = by_level[l];
    nodes[l] = int64_t(v.size());
    double dsum = 0;
    int64_t span = 0;
    for (int32_t p : v) {
      const int32_t a = nl.fin_off[size_t(p)], b = nl.fin_off[size_t(p) + 1];
      if (a == b) continue;
      double d = 0;
      for (int32_t e = a; e < b; ++e) {
        const int64_t dist = int64_t(l) - nl.level[size_t(nl.fin[size_t(e)])];
        d += double(dist);
        span = std::max(span, dist);
      }
      dsum += d / double(b - a);
    }
    avg_dist[l] = v.empty() ? 0.0 : dsum / double(v.size());
    max_span[l] = span;
    double rsum = 0, usum = 0;
    int64_t groups = 0;
    for (size_t i = 0; i < v.size(); i += size_t(wave)) {
      int32_t lo = INT32_MAX, hi = 0;
      uniq.clear();
      for (size_t j = i; j < std::min(v.size(), i + size_t(wave)); ++j)
        for (int32_t e = nl.fin_off[size_t(v[j])]; e < nl.fin_off[size_t(v[j]) + 1]; ++e) {
          const int32_t q = nl.fin[size_t(e)];
          lo = std::min(lo, q);
          hi = std::max(hi, q);
          uniq.insert(q);
        }
      if (lo != INT32_MAX) rsum += double(hi - lo);
      usum += double(uniq.size());
      ++groups;
    }
    avg_range[l] = groups ? rsum / double(groups) : 0.0;
    avg_unique[l] = groups ? usum / double(groups) : 0.0;
  }
  return LSIM_OK;
}

int lsim_visualization_text(const lsim_sim *s, char **text, size_t *len) {
  if (!text || !len) return LSIM_E_ARG;
  *text = nullptr;
  int rc = check_loaded(s);
  if (rc) return rc;
  return emit_text(dot_text(s->nl), text, len);
}

int lsim_visualization_graph(const lsim_sim *s, const char *path) {
  int rc = check_loaded(s);
  if (rc) return rc;
  if (!path) return LSIM_E_ARG;
  std::string *err = &const_cast<lsim_sim *>(s)->err;
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(err, LSIM_E_IO, std::string("Error opening output file: ") + path);
  const std::string t = dot_text(s->nl);
  const bool ok = std::fwrite(t.data(), 1, t.size(), f) == t.size();
  if (std::fclose(f) != 0 || !ok) return fail(err, LSIM_E_IO, std::string("cannot write ") + path);
  return LSIM_OK;
}

int lsim_generate_faults(uint64_t seed, int64_t num_pis, int64_t num_pos, int64_t num_gates, double fraction,
                         char **flst, size_t *flst_len) {
  if (!flst || !flst_len) return LSIM_E_ARG;
  *flst = nullptr;
  std::string text;
  int rc = guarded(&g_create_error, [&] { text = lsim::generate_flst(seed, num_pis, num_pos, num_gates, fraction); });
  return rc ? rc : emit_text(text, flst, flst_len);
}

int lsim_break_loops(const char *ckt, size_t ckt_len, int mode, char **out, size_t *out_len, int64_t *removed) {
  if (!out || !out_len || (!ckt && ckt_len) || (mode != LSIM_ORDER_BFS && mode != LSIM_ORDER_DFS)) return LSIM_E_ARG;
  *out = nullptr;
  std::string text;
  lsim::LoopStats st;
  int rc = guarded(&g_create_error, [&] { text = lsim::break_loops(ckt, ckt_len, mode == LSIM_ORDER_DFS, &st); });
  if (rc) return rc;
  if (removed) *removed = st.removed;
  return emit_text(text, out, out_len);
}

int lsim_convert_timing_dump(const char *dump, size_t dump_len, int mode, char **ckt, size_t *ckt_len,
                             lsim_ingest_stats *stats) {
  if (!ckt || !ckt_len || (!dump && dump_len) || (mode != LSIM_ORDER_BFS && mode != LSIM_ORDER_DFS))
    return LSIM_E_ARG;
  *ckt = nullptr;
  std::string text;
  lsim::IngestStats st;
  int rc = guarded(&g_create_error,
                   [&] { text = lsim::convert_timing_dump(dump, dump_len, mode == LSIM_ORDER_DFS, &st); });
  if (rc) return rc;
  if (stats) {
    stats->pis = st.pis;
    stats->pos = st.pos;
    stats->cells = st.cells;
    stats->wires = st.wires;
    stats->connections = st.connections;
    stats->records = st.records;
    stats->removed = st.removed;
    stats->unknown_cells = st.unknown_cells;
    stats->unknown_pins = st.unknown_pins;
    stats->unreached = st.unreached;
  }
  return emit_text(text, ckt, ckt_len);
}

int lsim_produce_faults(uint64_t seed, int64_t num_nodes, int64_t num_faults, int64_t num_pis, int64_t num_pos,
                        char **flst, size_t *flst_len) {
  if (!flst || !flst_len) return LSIM_E_ARG;
  *flst = nullptr;
  std::string text;
  int rc = guarded(&g_create_error,
                   [&] { text = lsim::produce_faults(seed, num_nodes, num_faults, num_pis, num_pos); });
  return rc ? rc : emit_text(text, flst, flst_len);
}

int lsim_pack_patterns(int64_t num_pis, int64_t num_patterns, int bits_per_round, const uint8_t *bits, char **ptn,
                       size_t *ptn_len) {
  if (!ptn || !ptn_len || (!bits && num_pis > 0 && num_patterns > 0)) return LSIM_E_ARG;
  *ptn = nullptr;
  std::string text;
  int rc = guarded(&g_create_error,
                   [&] { text = lsim::pack_patterns(num_pis, num_patterns, bits_per_round, bits); });
  return rc ? rc : emit_text(text, ptn, ptn_len);
}

int lsim_produce_patterns(uint64_t seed, int64_t num_pis, int64_t num_patterns, int bits_per_round, char **ptn,
                          size_t *ptn_len) {
  if (!ptn || !ptn_len) return LSIM_E_ARG;
  *ptn = nullptr;
  std::string text;
  int rc = guarded(&g_create_error,
                   [&] { text = lsim::produce_patterns(seed, num_pis, num_patterns, bits_per_round); });
  return rc ? rc : emit_text(text, ptn, ptn_len);
}

int lsim_generate_patterns(uint64_t seed, int64_t num_pis, int64_t num_patterns, char **ptn, size_t *ptn_len) {
  if (!ptn || !ptn_len) return LSIM_E_ARG;
  *ptn = nullptr;
  std::string text;
  int rc = guarded(&g_create_error, [&] { text = lsim::generate_ptn(seed, num_pis, num_patterns); });
  return rc ? rc : emit_text(text, ptn, ptn_len);
}

}  // extern "C"
