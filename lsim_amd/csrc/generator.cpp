// Seeded synthetic levelized DAG generator.
//
// Format conventions follow the reference's producers: node ids are PIs [0,P),
// POs [P,P+O), inner gates [P+O,N) (fsim/parser.py:213-227); a wire record is
// `src 0 k (dst pin)*k` and the last line lists the G inner-gate type codes
// (fsim/simulation.cpp:684-711,760-772); type codes are the GateType enum
// (fsim/base_ppg.hpp:19-32). Pattern words are decimal uint64, one line per
// 64-pattern round with one word per PI (fsim/simulation.cpp:961-976).
//
// The generator never emits what the reference reader cannot handle: no PI->PO
// edge, no gate without fanin, no unreachable node, at most 2 fanins per gate,
// and rounds <= PIs.
#include "generator.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "rng.hpp"
#include "textio.hpp"

namespace lsim {

uint64_t derive_seed(uint64_t seed, uint64_t stream) {
  uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull);
  return splitmix64(s);
}

namespace {

constexpr int kINV = 0, kCLKBUF = 8;
constexpr int kBinaryTypes[6] = {1 /*AND*/, 2 /*OR*/, 4 /*NAND*/, 5 /*NOR*/, 3 /*XOR*/, 6 /*XNOR*/};

struct Edge {
  int64_t src, dst;
  int pin;
};

struct Circuit {
  int64_t P, O, G, N;
  std::vector<int> type;    // per inner gate (index id-P-O)
  std::vector<Edge> edges;  // creation order
};

Circuit build_circuit(const GenSpec &s) {
  if (s.num_pis < 1 || s.num_pos < 0 || s.num_gates < 1 || s.num_levels < 1)
    throw std::invalid_argument("generator: need P>=1, O>=0, G>=1, L>=1");
  if (s.num_levels > s.num_gates) throw std::invalid_argument("generator: more levels than gates");
  if (s.num_pos > s.num_gates) throw std::invalid_argument("generator: more POs than gates");
  Circuit c;
  c.P = s.num_pis;
  c.O = s.num_pos;
  c.G = s.num_gates;
  c.N = c.P + c.O + c.G;
  const int64_t L = s.num_levels;
  Xoshiro256ss rng(derive_seed(s.seed, 1));

  // Slots: [0,P) are the PIs, slot P+k is inner gate id P+O+k. Gates are dealt
  // to levels 1..L evenly and in id order, so every level is a slot range.
  std::vector<int64_t> lvl_begin(L + 2);
  lvl_begin[0] = 0;
  lvl_begin[1] = c.P;
  for (int64_t l = 1; l <= L; ++l) {
    int64_t cnt = c.G / L + ((l - 1) < (c.G % L) ? 1 : 0);
    lvl_begin[l + 1] = lvl_begin[l] + cnt;
  }
  const int64_t S = c.P + c.G;
  auto slot_to_id = [&](int64_t slot) { return slot < c.P ? slot : slot + c.O; };

  // Source weights: discrete power law P(k) ~ k^-a on [1,cap], or all 1.
  std::vector<uint64_t> cw(S + 1, 0);
  const bool weighted = s.fanout_cap > 0;
  if (weighted) {
    std::vector<double> cdf(s.fanout_cap);
    double acc = 0;
    for (int k = 1; k <= s.fanout_cap; ++k) {
      acc += std::pow(double(k), -s.power_exp);
      cdf[k - 1] = acc;
    }
    for (auto &v : cdf) v /= acc;
    for (int64_t i = 0; i < S; ++i) {
      double u = rng.uniform();
      int64_t k = std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin() + 1;
      if (k > s.fanout_cap) k = s.fanout_cap;
      cw[i + 1] = cw[i] + uint64_t(k);
    }
  } else {
    for (int64_t i = 0; i < S; ++i) cw[i + 1] = cw[i] + 1;
  }
  std::vector<int32_t> fanout(S, 0);
  const int32_t cap = weighted ? s.fanout_cap : INT32_MAX;

  // Sample a slot in [a,b) proportional to weight, rejecting full sources.
  auto pick = [&](int64_t a, int64_t b, int64_t avoid) -> int64_t {
    int64_t got = -1;
    for (int tries = 0; tries < 64; ++tries) {
      uint64_t r = cw[a] + rng.below(cw[b] - cw[a]);
      int64_t i = int64_t(std::upper_bound(cw.begin() + a, cw.begin() + b + 1, r) - cw.begin()) - 1;
      got = i;
      if (fanout[i] >= cap) continue;
      if (i == avoid && b - a > 1) continue;
      break;
    }
    return got;
  };

  c.type.resize(c.G);
  c.edges.reserve(size_t(c.G) * 2 + c.O);
  for (int64_t l = 1; l <= L; ++l) {
    const int64_t prev_a = lvl_begin[l - 1], prev_b = lvl_begin[l];
    const int64_t win_a = lvl_begin[std::max<int64_t>(0, l - s.window)];
    for (int64_t slot = lvl_begin[l]; slot < lvl_begin[l + 1]; ++slot) {
      const int64_t g = slot - c.P;
      const bool unary = rng.uniform() < s.unary_frac;
      int nin;
      if (unary) {
        c.type[g] = rng.below(2) == 0 ? kINV : kCLKBUF;
        nin = 1;
      } else {
        c.type[g] = kBinaryTypes[rng.below(6)];
        nin = 2;
      }
      const int64_t dst = slot_to_id(slot);
      // First fanin pins the level: it always comes from level l-1.
      int64_t f0 = pick(prev_a, prev_b, -1);
      fanout[f0]++;
      c.edges.push_back({slot_to_id(f0), dst, 1});
      if (nin == 2) {
        const bool local = rng.uniform() < s.locality;
        int64_t f1 = local ? pick(win_a, prev_b, f0) : pick(0, prev_b, f0);
        fanout[f1]++;
        c.edges.push_back({slot_to_id(f1), dst, 2});
      }
    }
  }

  // POs: O distinct inner gates, deepest levels first, random within a level.
  std::vector<int64_t> drivers;
  drivers.reserve(c.O);
  for (int64_t l = L; l >= 1 && int64_t(drivers.size()) < c.O; --l) {
    std::vector<int64_t> lv;
    for (int64_t slot = lvl_begin[l]; slot < lvl_begin[l + 1]; ++slot) lv.push_back(slot);
    for (int64_t i = int64_t(lv.size()) - 1; i > 0; --i) std::swap(lv[i], lv[rng.below(uint64_t(i + 1))]);
    for (int64_t slot : lv) {
      if (int64_t(drivers.size()) == c.O) break;
      drivers.push_back(slot);
    }
  }
  for (int64_t i = 0; i < c.O; ++i) {
    fanout[drivers[i]]++;
    c.edges.push_back({slot_to_id(drivers[i]), c.P + i, 1});
  }
  return c;
}

std::string ckt_text(const Circuit &c) {
  // Group edges by source, keeping creation order inside a source.
  std::vector<int64_t> cnt(c.N + 1, 0);
  for (const auto &e : c.edges) cnt[e.src + 1]++;
  for (int64_t i = 0; i < c.N; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> at(cnt.begin(), cnt.end() - 1);
  std::vector<const Edge *> by_src(c.edges.size());
  for (const auto &e : c.edges) by_src[at[e.src]++] = &e;
  int64_t wires = 0;
  for (int64_t v = 0; v < c.N; ++v) wires += (cnt[v + 1] > cnt[v]);

  std::string out;
  out.reserve(size_t(c.edges.size()) * 16 + size_t(c.G) * 2 + 64);
  TextOut t(out);
  t.u(c.P).c(' ').u(c.O).c(' ').u(c.G).c(' ').u(wires).c('\n');
  for (int64_t v = 0; v < c.N; ++v) {
    if (cnt[v + 1] == cnt[v]) continue;
    t.u(v).c(' ').u(0).c(' ').u(cnt[v + 1] - cnt[v]);
    for (int64_t k = cnt[v]; k < cnt[v + 1]; ++k) t.c(' ').u(by_src[k]->dst).c(' ').u(by_src[k]->pin);
    t.c('\n');
  }
  for (int64_t g = 0; g < c.G; ++g) {
    if (g) t.c(' ');
    t.u(c.type[g]);
  }
  t.c('\n');
  return out;
}

}  // namespace

std::string generate_ptn(uint64_t seed, int64_t num_pis, int64_t num_patterns) {
  if (num_patterns < 0) throw std::invalid_argument("generator: negative pattern count");
  const int64_t R = (num_patterns + 63) / 64;
  if (R > num_pis) throw std::invalid_argument("generator: rounds exceed PIs (reader requires R <= P)");
  Xoshiro256ss rng(derive_seed(seed, 2));
  std::string out;
  out.reserve(size_t(R * num_pis) * 21 + 32);
  TextOut t(out);
  t.u(uint64_t(num_patterns)).c('\n');
  const int tail = int(num_patterns % 64);
  for (int64_t r = 0; r < R; ++r) {
    const uint64_t mask = (r == R - 1 && tail) ? ((uint64_t(1) << tail) - 1) : ~uint64_t(0);
    for (int64_t i = 0; i < num_pis; ++i) {
      if (i) t.c(' ');
      t.u(rng.next() & mask);
    }
    t.c('\n');
  }
  return out;
}

std::string generate_flst(uint64_t seed, int64_t num_pis, int64_t num_pos, int64_t num_gates,
                          double fraction) {
  if (!(fraction > 0.0 && fraction <= 1.0)) throw std::invalid_argument("generator: fault fraction must be in (0,1]");
  Xoshiro256ss rng(derive_seed(seed, 3));
  std::string body;
  TextOut t(body);
  int64_t F = 0;
  const int64_t first = num_pis + num_pos;
  for (int64_t id = first; id < first + num_gates; ++id) {
    for (int v = 0; v < 2; ++v) {
      if (fraction < 1.0 && !(rng.uniform() < fraction)) continue;
      t.u(id).c(' ').u(v).c('\n');
      ++F;
    }
  }
  std::string out;
  TextOut h(out);
  h.u(F).c('\n');
  out += body;
  return out;
}

std::string generate_ckt(const GenSpec &spec) { return ckt_text(build_circuit(spec)); }

GenText generate(const GenSpec &spec) {
  GenText g;
  g.ckt = generate_ckt(spec);
  g.flst = generate_flst(spec.seed, spec.num_pis, spec.num_pos, spec.num_gates, 1.0);
  g.ptn = generate_ptn(spec.seed, spec.num_pis, spec.num_patterns);
  return g;
}

}  // namespace lsim
