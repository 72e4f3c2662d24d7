// Reader and levelizer for the reference's .ckt/.flst/.ptn inputs.
//
// Follows fsim::Simulator::read (fsim/simulation.cpp:37-63) and _read_graph /
// _read_fault / _read_pattern (fsim/simulation.cpp:664-976): FIFO Kahn order
// seeded with PIs 0..P-1, children in wire-file order, POs moved to the end in
// id order, fanins in pin order with pin 5 (MUX select) stored as pin 3, levels
// = 1 + max fanin level with every PO one level above the deepest gate.
//
// Pins follow the reference's zero-filled pin table (fsim/simulation.cpp:676-711,
// 851-877): a node with k wires reads pins 1..k, a pin slot no wire filled reads
// node id 0 (PI 0), a second wire into one pin overwrites the first, and wires
// into pins above k are not read (they still order the node in Kahn's queue).
// The levels follow the pins actually read, so they need not be sorted in
// topological order.
//
// Inputs on which the reference has undefined behaviour are rejected with an
// InputError instead of being reproduced: PI->PO edges, a PO reading an unfilled
// pin (the reference then indexes its gate array with a PI), PO fanout, fanin
// into a PI, >4 fanins, gates or POs without fanin, unreachable or cyclic nodes,
// unknown gate types, fault ids >= N, more rounds than PIs, and truncated or
// non-numeric text.
#include "netlist.hpp"

#include <algorithm>
#include <string>

#include "textio.hpp"

namespace lsim {

namespace {

uint64_t need(TextIn &in, const char *what) {
  uint64_t v;
  if (!in.next(v)) throw InputError(std::string(in.bad() ? "malformed number in " : "truncated ") + what);
  return v;
}

}  // namespace

void read_circuit(Netlist &nl, const char *ckt, size_t ckt_len) {
  nl = Netlist();
  TextIn c(ckt, ckt_len);
  nl.P = int64_t(need(c, "circuit header"));
  nl.O = int64_t(need(c, "circuit header"));
  nl.G = int64_t(need(c, "circuit header"));
  nl.W = int64_t(need(c, "circuit header"));
  nl.N = nl.P + nl.O + nl.G;
  if (nl.P < 1) throw InputError("circuit needs at least one PI");
  if (nl.N >= (int64_t(1) << 31) - 1) throw InputError("circuit too large (node count must fit in int32)");
  const int64_t N = nl.N, P = nl.P, O = nl.O;
  auto is_po = [&](int64_t id) { return id >= P && id < P + O; };

  // Wires: adjacency in file order and the per-destination pin table.
  std::vector<int64_t> adj_head(N, -1), adj_next, adj_dst;
  std::vector<int64_t> adj_tail(N, -1);
  adj_next.reserve(size_t(nl.G) * 2 + O);
  adj_dst.reserve(size_t(nl.G) * 2 + O);
  std::vector<int32_t> indeg(N, 0);
  std::vector<int64_t> table(size_t(N) * 4, -1);  // pins 1..4 -> slot 0..3
  std::vector<uint8_t> pin_count(N, 0);
  for (int64_t w = 0; w < nl.W; ++w) {
    const uint64_t src = need(c, "wire record");
    need(c, "wire record");  // output pin id, unused by the reference
    const uint64_t k = need(c, "wire record");
    if (src >= uint64_t(N)) throw InputError("wire source id out of range");
    if (is_po(int64_t(src))) throw InputError("a PO drives other nodes");
    for (uint64_t j = 0; j < k; ++j) {
      const uint64_t dst = need(c, "wire record");
      uint64_t pin = need(c, "wire record");
      if (dst >= uint64_t(N)) throw InputError("wire destination id out of range");
      if (int64_t(dst) < P) throw InputError("a wire drives a PI");
      if (src < uint64_t(P) && is_po(int64_t(dst))) throw InputError("PI drives a PO directly");
      if (pin < 1 || pin > 5) throw InputError("input pin id must be 1..5");
      if (pin == 5) pin = 3;
      table[size_t(dst) * 4 + (pin - 1)] = int64_t(src);
      if (++pin_count[dst] > 4) throw InputError("more than 4 fanins");
      const int64_t e = int64_t(adj_dst.size());
      adj_dst.push_back(int64_t(dst));
      adj_next.push_back(-1);
      if (adj_tail[src] < 0) adj_head[src] = e; else adj_next[adj_tail[src]] = e;
      adj_tail[src] = e;
      indeg[dst]++;
    }
  }

  // FIFO Kahn from the PIs (fsim/simulation.cpp:719-738).
  std::vector<int64_t> order;
  order.reserve(N);
  {
    std::vector<int64_t> queue(N);
    size_t qh = 0, qt = 0;
    for (int64_t i = 0; i < P; ++i) queue[qt++] = i;
    while (qh < qt) {
      const int64_t u = queue[qh++];
      order.push_back(u);
      for (int64_t e = adj_head[u]; e >= 0; e = adj_next[e]) {
        const int64_t v = adj_dst[e];
        if (--indeg[v] == 0) queue[qt++] = v;
      }
    }
  }
  if (int64_t(order.size()) != N) throw InputError("circuit has unreachable nodes or a cycle");
  // Move POs to the end in id order (fsim/simulation.cpp:742-751).
  order.erase(std::remove_if(order.begin(), order.end(), is_po), order.end());
  for (int64_t i = 0; i < O; ++i) order.push_back(P + i);
  nl.order = order;
  nl.order_idx.assign(N, 0);
  for (int64_t p = 0; p < N; ++p) nl.order_idx[order[p]] = p;

  // Types (fsim/simulation.cpp:760-772).
  std::vector<uint8_t> type_by_id(N);
  for (int64_t i = 0; i < P; ++i) type_by_id[i] = G_PI;
  for (int64_t i = P; i < P + O; ++i) type_by_id[i] = G_PO;
  for (int64_t i = P + O; i < N; ++i) {
    const uint64_t t = need(c, "gate type list");
    if (t > G_CLKBUF) throw InputError("inner gate type code must be 0..8");
    type_by_id[i] = uint8_t(t);
  }

  // Per-position type and fanins in pin order (fsim/simulation.cpp:851-877).
  nl.type.resize(N);
  nl.fin_off.assign(N + 1, 0);
  nl.fin.clear();
  nl.fin.reserve(adj_dst.size());
  for (int64_t p = 0; p < N; ++p) {
    const int64_t id = order[p];
    nl.type[p] = type_by_id[id];
    const int cnt = pin_count[id];
    if (p >= P) {
      if (cnt == 0) throw InputError(is_po(id) ? "PO without fanin" : "gate without fanin");
      for (int j = 0; j < cnt; ++j) {
        int64_t src = table[size_t(id) * 4 + j];
        if (src < 0) {
          // Unfilled pin slot: node id 0 (fsim/simulation.cpp:678-680,856-861).
          if (is_po(id)) throw InputError("a PO reads an input pin no wire drives");
          src = 0;
        }
        nl.fin.push_back(int32_t(nl.order_idx[src]));
      }
      if (nl.type[p] == G_MUX) {
        if (cnt < 3) throw InputError("MUX needs 3 inputs");
        nl.has_mux = true;
      }
    }
    nl.fin_off[p + 1] = int32_t(nl.fin.size());
  }

  // Levels (fsim/simulation.cpp:882-916).
  nl.level.assign(N, 0);
  int32_t maxg = 0;
  for (int64_t p = P; p < P + nl.G; ++p) {
    int32_t l = 0;
    for (int32_t e = nl.fin_off[p]; e < nl.fin_off[p + 1]; ++e) l = std::max(l, nl.level[nl.fin[e]] + 1);
    nl.level[p] = l;
    maxg = std::max(maxg, l);
  }
  nl.max_level = maxg + 1;
  for (int64_t p = P + nl.G; p < N; ++p) nl.level[p] = nl.max_level;
  nl.gates_per_level.assign(size_t(nl.max_level) + 1, 0);
  for (int64_t p = 0; p < N; ++p) nl.gates_per_level[nl.level[p]]++;

}

void read_faults(const Netlist &nl, const char *flst, size_t flst_len, std::vector<int32_t> &site,
                 std::vector<uint8_t> &sa) {
  TextIn f(flst, flst_len);
  const uint64_t F = need(f, "fault list header");
  if (F > uint64_t(INT32_MAX)) throw InputError("fault list too long");
  site.resize(F);
  sa.resize(F);
  for (uint64_t i = 0; i < F; ++i) {
    const uint64_t id = need(f, "fault list");
    const uint64_t v = need(f, "fault list");
    if (id >= uint64_t(nl.N)) throw InputError("fault id out of range");
    site[i] = int32_t(nl.order_idx[id]);
    sa[i] = v ? 1 : 0;
  }
}

void read_patterns(const Netlist &nl, const char *ptn, size_t ptn_len, int64_t &num_patterns, int64_t &R,
                   std::vector<uint64_t> &words) {
  // The reference stores round i at _patterns[order_idx[i]], which is round i
  // itself whenever R <= P.
  TextIn t(ptn, ptn_len);
  num_patterns = int64_t(need(t, "pattern header"));
  R = (num_patterns + 63) / 64;
  if (R > nl.P) throw InputError("more 64-pattern rounds than PIs");
  words.resize(size_t(R * nl.P));
  for (int64_t i = 0; i < R * nl.P; ++i) words[i] = need(t, "pattern words");
}

void read_netlist(Netlist &nl, const char *ckt, size_t ckt_len, const char *flst, size_t flst_len,
                  const char *ptn, size_t ptn_len) {
  read_circuit(nl, ckt, ckt_len);
  read_faults(nl, flst, flst_len, nl.fault_site, nl.fault_sa);
  read_patterns(nl, ptn, ptn_len, nl.num_patterns, nl.R, nl.patterns);
}

void build_fault_structure(const Netlist &nl, FaultStructure &fs) {
  const int64_t N = nl.N;
  // Distinct fanout nodes per position (a node feeding two pins of one gate
  // has that gate once).
  std::vector<int32_t> cnt(N + 1, 0);
  for (int64_t p = 0; p < N; ++p)
    for (int32_t e = nl.fin_off[p]; e < nl.fin_off[p + 1]; ++e) {
      bool dup = false;
      for (int32_t e2 = nl.fin_off[p]; e2 < e; ++e2) dup |= nl.fin[e2] == nl.fin[e];
      if (!dup) cnt[nl.fin[e] + 1]++;
    }
  fs.fo_off.assign(N + 1, 0);
  for (int64_t p = 0; p < N; ++p) fs.fo_off[p + 1] = fs.fo_off[p] + cnt[p + 1];
  fs.fo.assign(fs.fo_off[N], 0);
  std::vector<int32_t> at(fs.fo_off.begin(), fs.fo_off.end() - 1);
  for (int64_t p = 0; p < N; ++p)
    for (int32_t e = nl.fin_off[p]; e < nl.fin_off[p + 1]; ++e) {
      bool dup = false;
      for (int32_t e2 = nl.fin_off[p]; e2 < e; ++e2) dup |= nl.fin[e2] == nl.fin[e];
      if (!dup) fs.fo[at[nl.fin[e]]++] = int32_t(p);
    }
  fs.reach_po.assign(N, 0);
  fs.parent.assign(N, -1);
  fs.stem_of.assign(N, -1);
  fs.stems.clear();
  for (int64_t p = N - 1; p >= 0; --p) {
    if (nl.type[p] == G_PO) {
      fs.reach_po[p] = 1;
      fs.stem_of[p] = int32_t(p);
      continue;
    }
    // Only fanouts that can reach a PO matter: a node with a single live
    // fanout belongs to that fanout's region even if it also feeds dead logic.
    int32_t live = 0, last = -1;
    for (int32_t e = fs.fo_off[p]; e < fs.fo_off[p + 1]; ++e)
      if (fs.reach_po[fs.fo[e]]) {
        ++live;
        last = fs.fo[e];
      }
    fs.reach_po[p] = live > 0;
    if (!live) continue;  // dead: no PO can observe it
    if (live == 1) {
      fs.parent[p] = last;
      fs.stem_of[p] = fs.stem_of[last];
    } else {
      fs.stem_of[p] = int32_t(p);
    }
  }
  for (int64_t p = 0; p < N; ++p)
    if (fs.reach_po[p] && nl.type[p] != G_PO && fs.parent[p] < 0) fs.stems.push_back(int32_t(p));
}

}  // namespace lsim
