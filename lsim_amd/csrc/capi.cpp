// C ABI of the simulator (include/lsim.h). Every entry point catches C++
// exceptions and turns them into LSIM_E_* codes plus a message.
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <sstream>
#include <string>

#include "engine.hpp"
#include "generator.hpp"
#include "lsim.h"
#include "netlist.hpp"

struct lsim_sim {
  int device = -1;  // -1: host-only handle (read and accessors, no run)
  std::unique_ptr<lsim::Engine> engine;
  lsim::Netlist nl;
  lsim::FaultStructure fs;
  bool loaded = false;
  bool ran = false;
  lsim::RunOptions opt;
  lsim::RunResults res;
  std::string err;
};

namespace {

thread_local std::string g_create_error;

template <typename Fn>
int guarded(lsim_sim *s, Fn &&fn) {
  try {
    fn();
    if (s) s->err.clear();
    return LSIM_OK;
  } catch (const lsim::InputError &e) {
    if (s) s->err = e.what();
    return LSIM_E_INPUT;
  } catch (const lsim::DeviceError &e) {
    if (s) s->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    if (s) s->err = "host out of memory";
    return LSIM_E_NOMEM;
  } catch (const std::exception &e) {
    if (s) s->err = e.what();
    return LSIM_E_ARG;
  }
}

int fail(lsim_sim *s, int code, const std::string &msg) {
  if (s) s->err = msg;
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

int load_texts(lsim_sim *s, const char *ckt, size_t ckt_len, const char *flst, size_t flst_len, const char *ptn,
               size_t ptn_len) {
  return guarded(s, [&] {
    s->loaded = false;
    s->ran = false;
    s->res = lsim::RunResults();
    lsim::read_netlist(s->nl, ckt, ckt_len, flst, flst_len, ptn, ptn_len);
    lsim::build_fault_structure(s->nl, s->fs);
    if (s->engine) s->engine->load(s->nl, s->fs);
    s->loaded = true;
  });
}

int check_loaded(const lsim_sim *s) {
  if (!s) return LSIM_E_ARG;
  if (!s->loaded) {
    const_cast<lsim_sim *>(s)->err = "no circuit has been read";
    return LSIM_E_STATE;
  }
  return LSIM_OK;
}

int check_n(const lsim_sim *s, const void *p, size_t n, size_t want) {
  if (!p && want) return fail(const_cast<lsim_sim *>(s), LSIM_E_ARG, "null output buffer");
  if (n != want)
    return fail(const_cast<lsim_sim *>(s), LSIM_E_ARG,
                "output buffer holds " + std::to_string(n) + " elements, expected " + std::to_string(want));
  return LSIM_OK;
}

int check_ran(const lsim_sim *s) {
  int rc = check_loaded(s);
  if (rc) return rc;
  if (!s->ran) return fail(const_cast<lsim_sim *>(s), LSIM_E_STATE, "run() has not completed");
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
    int rc = guarded(s.get(), [&] { s->engine.reset(new lsim::Engine(device)); });
    if (rc) {
      g_create_error = s->err;
      return rc;
    }
  }
  *out = s.release();
  return LSIM_OK;
}

void lsim_destroy(lsim_sim *s) { delete s; }

const char *lsim_last_error(const lsim_sim *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int lsim_read_files(lsim_sim *s, const char *ckt_path, const char *flst_path, const char *ptn_path) {
  if (!s || !ckt_path || !flst_path || !ptn_path) return LSIM_E_ARG;
  std::string c, f, p;
  // Messages as in fsim/simulation.cpp:24-32.
  if (!read_file(ckt_path, c)) return fail(s, LSIM_E_IO, std::string("cannot open circut file ") + ckt_path);
  if (!read_file(flst_path, f)) return fail(s, LSIM_E_IO, std::string("cannot open fault file ") + flst_path);
  if (!read_file(ptn_path, p)) return fail(s, LSIM_E_IO, std::string("cannot open pattern file ") + ptn_path);
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
  if (k == "drop") s->opt.drop = value != 0;
  else if (k == "report_last_round") s->opt.report_last_round = value != 0;
  else if (k == "scratch_mb") s->opt.scratch_mb = value < 0 ? 0 : value;
  else if (k == "direct") s->opt.direct = value != 0;
  else if (k == "block_waves") {
    if (value < 1 || value > 16) return fail(s, LSIM_E_ARG, "block_waves must be 1..16");
    s->opt.block_waves = int(value);
  } else
    return fail(s, LSIM_E_ARG, "unknown option " + k);
  return LSIM_OK;
}

int lsim_run(lsim_sim *s, int mode, size_t num_threads) {
  (void)num_threads;
  int rc = check_loaded(s);
  if (rc) return rc;
  if (!s->engine) return fail(s, LSIM_E_DEVICE, "handle was created without a device (device -1)");
  int m;
  switch (mode) {
    case LSIM_MODE_REF_SEQUENTIAL:
    case LSIM_MODE_REF_PARALLEL_OMP: m = LSIM_MODE_PLAIN; break;
    case LSIM_MODE_REF_PARALLEL_TF: m = LSIM_MODE_PARTITIONED; break;
    default: return fail(s, LSIM_E_ARG, "unknown mode " + std::to_string(mode));
  }
  return guarded(s, [&] {
    s->ran = false;
    lsim::RunOptions o = s->opt;
    o.mode = m;
    s->engine->run(o, s->res);
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
  if (rc || (rc = check_n(s, out, n, size_t(s->nl.N)))) return rc;
  for (size_t i = 0; i < n; ++i) out[i] = uint64_t(s->nl.order_idx[i]);
  return LSIM_OK;
}

int lsim_num_gates_per_level(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(s, out, n, s->nl.gates_per_level.size()))) return rc;
  for (size_t i = 0; i < n; ++i) out[i] = uint64_t(s->nl.gates_per_level[i]);
  return LSIM_OK;
}

int lsim_levels(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(s, out, n, size_t(s->nl.N)))) return rc;
  for (size_t i = 0; i < n; ++i) out[i] = uint64_t(s->nl.level[i]);
  return LSIM_OK;
}

int lsim_faults(const lsim_sim *s, uint64_t *site, uint64_t *value, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(s, site, n, s->nl.fault_site.size())) || (rc = check_n(s, value, n, n))) return rc;
  for (size_t i = 0; i < n; ++i) {
    site[i] = uint64_t(s->nl.fault_site[i]);
    value[i] = s->nl.fault_sa[i] ? ~uint64_t(0) : 0;
  }
  return LSIM_OK;
}

int lsim_patterns(const lsim_sim *s, uint64_t *out, size_t n) {
  int rc = check_loaded(s);
  if (rc || (rc = check_n(s, out, n, s->nl.patterns.size()))) return rc;
  std::memcpy(out, s->nl.patterns.data(), n * sizeof(uint64_t));
  return LSIM_OK;
}

int lsim_good_results(const lsim_sim *s, uint64_t *pi, size_t n_pi, uint64_t *gate, size_t n_gate, uint64_t *po,
                      size_t n_po) {
  int rc = check_ran(s);
  if (rc || (rc = check_n(s, pi, n_pi, size_t(s->nl.P))) || (rc = check_n(s, gate, n_gate, size_t(s->nl.G))) ||
      (rc = check_n(s, po, n_po, size_t(s->nl.O))))
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
  if (rc || (rc = check_n(s, out, n, 2 * F))) return rc;
  if (!s->res.has_last)
    return fail(const_cast<lsim_sim *>(s), LSIM_E_STATE,
                "the last round was not evaluated for every fault (dropping on, report_last_round off)");
  // {found in the last round, last round} (fsim/simulation.cpp:142-144);
  // no rounds leaves the table zero-filled (fsim/simulation.cpp:116).
  const int64_t R = s->nl.R;
  for (size_t f = 0; f < F; ++f) {
    out[2 * f] = R > 0 ? s->res.last_round[f] : 0;
    out[2 * f + 1] = R > 0 ? uint64_t(R - 1) : 0;
  }
  return LSIM_OK;
}

int lsim_detected(const lsim_sim *s, uint8_t *out, size_t n) {
  int rc = check_ran(s);
  if (rc || (rc = check_n(s, out, n, s->nl.fault_site.size()))) return rc;
  for (size_t f = 0; f < n; ++f) out[f] = s->res.first_round[f] >= 0;
  return LSIM_OK;
}

int lsim_first_detect_round(const lsim_sim *s, int64_t *out, size_t n) {
  int rc = check_ran(s);
  if (rc || (rc = check_n(s, out, n, s->nl.fault_site.size()))) return rc;
  for (size_t f = 0; f < n; ++f) out[f] = s->res.first_round[f];
  return LSIM_OK;
}

int lsim_detect_matrix(const lsim_sim *s, uint8_t *out, size_t n) {
  int rc = check_ran(s);
  const size_t F = s ? s->nl.fault_site.size() : 0;
  const int64_t R = s ? s->nl.R : 0;
  if (rc || (rc = check_n(s, out, n, F * size_t(R)))) return rc;
  if (!s->res.has_matrix)
    return fail(const_cast<lsim_sim *>(s), LSIM_E_STATE, "the run used fault dropping; no per-round matrix");
  const int64_t RB = s->res.RB;
  for (size_t f = 0; f < F; ++f)
    for (int64_t r = 0; r < R; ++r) out[f * R + r] = uint8_t((s->res.det_bits[f * RB + (r >> 6)] >> (r & 63)) & 1);
  return LSIM_OK;
}

int lsim_last_run_stats(const lsim_sim *s, lsim_run_stats *out) {
  int rc = check_ran(s);
  if (rc) return rc;
  if (!out) return LSIM_E_ARG;
  *out = s->res.stats;
  return LSIM_OK;
}

int lsim_generate(const lsim_gen_spec *spec, char **ckt, size_t *ckt_len, char **flst, size_t *flst_len, char **ptn,
                  size_t *ptn_len) {
  if (!spec || !ckt || !ckt_len || !flst || !flst_len || !ptn || !ptn_len) return LSIM_E_ARG;
  *ckt = *flst = *ptn = nullptr;
  lsim_sim scratch;  // carries the error message of this call
  int rc = guarded(&scratch, [&] {
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
    g_create_error = scratch.err;
    std::free(*ckt);
    std::free(*flst);
    std::free(*ptn);
    *ckt = *flst = *ptn = nullptr;
  }
  return rc;
}

void lsim_free_text(char *p) { std::free(p); }

}  // extern "C"
