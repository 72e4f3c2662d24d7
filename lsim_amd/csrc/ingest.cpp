// OpenTimer timing-graph dump -> .ckt, following the reference's converter
// (fsim/parser.py:1-227):
//  * sections are located by their marker lines ("// Start PIs", "// Start POs",
//    "// Start wires", "// Start cells", "digraph TimingGraph {") and sliced as
//    the converter slices them (parser.py:11-25);
//  * a PI/PO/wire line contributes its second token less its last character,
//    a cell line its instance name (second token) and cell type (first);
//    connections are the "a -> b" lines of the digraph with ' ";' stripped;
//  * cell types map to gate codes by the first matching name prefix
//    (parser.py:58-86, unknown -> AND) and pins to pin numbers through the
//    converter's per-cell table (parser.py:36-54, unknown -> 0);
//  * node ids: PIs, then POs, then cells, in file order (parser.py:104);
//  * loops are removed by a BFS or DFS traversal from the PIs (parser.py:163-211,
//    the rule lsim::break_loops applies to a .ckt);
//  * the output lists P, O, G, the record count, one record per source pin with
//    its destinations, then the cells' gate codes (parser.py:213-227).
// The converter keeps each source's destinations in a Python set, whose order
// is not fixed from run to run; here they keep the order of the dump.
#include <algorithm>
#include <deque>
#include <string>
#include <string_view>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "ingest.hpp"
#include "netlist.hpp"
#include "textio.hpp"

namespace lsim {

namespace {

using SV = std::string_view;

std::vector<SV> split_lines(SV t) {
  std::vector<SV> out;
  size_t a = 0;
  for (;;) {
    const size_t b = t.find('\n', a);
    if (b == SV::npos) {
      out.push_back(t.substr(a));
      return out;
    }
    out.push_back(t.substr(a, b - a));
    a = b + 1;
  }
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

std::vector<SV> tokens(SV x) {
  std::vector<SV> out;
  size_t i = 0;
  while (i < x.size()) {
    while (i < x.size() && is_ws(x[i])) ++i;
    const size_t a = i;
    while (i < x.size() && !is_ws(x[i])) ++i;
    if (i > a) out.push_back(x.substr(a, i - a));
  }
  return out;
}

SV strip_set(SV x, SV chars) {
  size_t a = 0, b = x.size();
  while (a < b && chars.find(x[a]) != SV::npos) ++a;
  while (b > a && chars.find(x[b - 1]) != SV::npos) --b;
  return x.substr(a, b - a);
}

SV before_colon(SV x) {
  const size_t c = x.find(':');
  return c == SV::npos ? x : x.substr(0, c);
}

bool starts_with(SV s, SV p) { return s.size() >= p.size() && s.compare(0, p.size(), p) == 0; }

size_t marker(const std::vector<SV> &lines, SV m) {
  for (size_t i = 0; i < lines.size(); ++i)
    if (lines[i] == m) return i;
  throw InputError("timing dump has no line \"" + std::string(m) + "\"");
}

// lines[a:b] with Python slice bounds (b may be negative).
std::vector<SV> slice(const std::vector<SV> &lines, int64_t a, int64_t b) {
  const int64_t n = int64_t(lines.size());
  if (b < 0) b += n;
  a = std::max<int64_t>(0, std::min(a, n));
  b = std::max<int64_t>(0, std::min(b, n));
  std::vector<SV> out;
  for (int64_t i = a; i < b; ++i) out.push_back(lines[size_t(i)]);
  return out;
}

SV second_token_chopped(SV line, const char *what) {
  const std::vector<SV> t = tokens(line);
  if (t.size() < 2) throw InputError(std::string("timing dump: malformed ") + what + " line \"" + std::string(line) + "\"");
  return t[1].substr(0, t[1].empty() ? 0 : t[1].size() - 1);
}

struct PinEntry {
  const char *cell;
  const char *pins[5];
  int ids[5];
};

// Per-cell pin numbers (parser.py:36-54); output pins map to 0.
const PinEntry kPins[] = {
    {"CLKBUF", {"A", "Z"}, {1, 0}},
    {"INV", {"A", "ZN"}, {1, 0}},
    {"AND2", {"A", "B", "Y"}, {1, 2, 0}},
    {"NAND2", {"A1", "A2", "ZN"}, {1, 2, 0}},
    {"NAND3", {"A1", "A2", "A3", "ZN"}, {1, 2, 3, 0}},
    {"NAND4", {"A1", "A2", "A3", "A4", "ZN"}, {1, 2, 3, 4, 0}},
    {"NOR2", {"A1", "A2", "ZN"}, {1, 2, 0}},
    {"NOR3", {"A1", "A2", "A3", "ZN"}, {1, 2, 3, 0}},
    {"NOR4", {"A1", "A2", "A3", "A4", "ZN"}, {1, 2, 3, 4, 0}},
    {"AOI21", {"A", "B1", "B2", "ZN"}, {1, 2, 3, 0}},
    {"AOI22", {"A1", "A2", "B1", "B2", "ZN"}, {1, 2, 3, 4, 0}},
    {"DFF", {"D", "CK", "Q"}, {1, 2, 0}},
    {"OAI21", {"A", "B1", "B2", "ZN"}, {1, 2, 3, 0}},
    {"OAI22", {"A1", "A2", "B1", "B2", "ZN"}, {1, 2, 3, 4, 0}},
    {"AND2X1", {"A", "B", "Y"}, {1, 2, 0}},
    {"INVX1", {"A", "Y"}, {1, 0}},
    {"NAND3X1", {"A", "B", "C", "Y"}, {1, 2, 3, 0}},
};

// Cell type -> gate code by the first matching prefix (parser.py:58-86);
// the complex cells become the NAND/NOR of their inputs.
const std::pair<const char *, int> kTypes[] = {
    {"INV", 0},   {"AND", 1},   {"OR", 2},    {"XOR", 3},   {"NAND", 4},  {"NOR", 5},   {"XNOR", 6},
    {"MUX", 7},   {"CLKBUF", 8}, {"AOI21", 4}, {"AOI22", 4}, {"DFF", 4},   {"OAI21", 5}, {"OAI22", 5},
};

int type_code(SV cell, bool &known) {
  for (const auto &e : kTypes)
    if (starts_with(cell, e.first)) {
      known = true;
      return e.second;
    }
  known = false;
  return 1;
}

// An insertion-ordered set of destination names.
struct Dests {
  std::vector<std::string> v;
  std::unordered_set<std::string> has;
  void add(SV x) {
    std::string s(x);
    if (has.insert(s).second) v.push_back(std::move(s));
  }
};

}  // namespace

std::string convert_timing_dump(const char *text, size_t len, bool dfs, IngestStats *stats) {
  const std::vector<SV> lines = split_lines(SV(text ? text : "", text ? len : 0));
  const int64_t i_pi = int64_t(marker(lines, "// Start PIs")), i_po = int64_t(marker(lines, "// Start POs")),
                i_wire = int64_t(marker(lines, "// Start wires")), i_cell = int64_t(marker(lines, "// Start cells")),
                i_dg = int64_t(marker(lines, "digraph TimingGraph {"));
  IngestStats st;
  std::vector<std::string> pis, pos;
  for (SV x : slice(lines, i_pi + 1, i_po - 1)) pis.emplace_back(second_token_chopped(x, "PI"));
  for (SV x : slice(lines, i_po + 1, i_wire - 1)) pos.emplace_back(second_token_chopped(x, "PO"));
  for (SV x : slice(lines, i_wire + 1, i_cell - 1)) {
    (void)second_token_chopped(x, "wire");
    ++st.wires;
  }
  std::vector<std::pair<std::string, std::string>> cells;  // (instance, cell type)
  for (SV x : slice(lines, i_cell + 1, i_dg - 2)) {
    const std::vector<SV> t = tokens(x);
    if (t.size() < 2) throw InputError("timing dump: malformed cell line \"" + std::string(x) + "\"");
    cells.emplace_back(std::string(t[1]), std::string(t[0]));
  }
  // Connections "src -> dst" of the digraph, grouped by source in order.
  std::vector<std::string> keys;
  std::unordered_map<std::string, Dests> conn;
  for (SV x : slice(lines, i_dg + 1, -2)) {
    if (x.find("->") == SV::npos) continue;
    const size_t arrow = x.find(" -> ");
    if (arrow == SV::npos) throw InputError("timing dump: malformed connection \"" + std::string(x) + "\"");
    const size_t arrow2 = x.find(" -> ", arrow + 4);
    const SV a = strip_set(x.substr(0, arrow), " \";");
    const SV b = strip_set(x.substr(arrow + 4, arrow2 == SV::npos ? SV::npos : arrow2 - arrow - 4), " \";");
    ++st.connections;
    if (before_colon(a) == before_colon(b)) continue;  // within one cell
    auto it = conn.find(std::string(a));
    if (it == conn.end()) {
      keys.emplace_back(a);
      it = conn.emplace(std::string(a), Dests()).first;
    }
    it->second.add(b);
  }
  // Node ids and per-instance pin tables.
  std::unordered_map<std::string, int64_t> id;
  {
    int64_t k = 0;
    for (const auto &n : pis) id[n] = k++;
    for (const auto &n : pos) id[n] = k++;
    for (const auto &c : cells) id[c.first] = k++;
  }
  std::unordered_map<std::string, std::string> pin_key;
  for (const auto &c : cells) {
    const size_t u = c.second.find('_');
    pin_key[c.first] = u == std::string::npos ? c.second : c.second.substr(0, u);
  }
  std::vector<int> types;
  for (const auto &c : cells) {
    bool known = false;
    types.push_back(type_code(c.second, known));
    st.unknown_cells += known ? 0 : 1;
  }
  const std::unordered_set<std::string> pi_set(pis.begin(), pis.end()), po_set(pos.begin(), pos.end());
  auto gate_and_pin = [&](const std::string &s) -> std::pair<int64_t, int> {
    const size_t c = s.find(':');
    if (c != std::string::npos) {
      if (s.find(':', c + 1) != std::string::npos) throw InputError("timing dump: malformed pin \"" + s + "\"");
      const std::string g = s.substr(0, c), p = s.substr(c + 1);
      const auto gi = id.find(g);
      if (gi == id.end()) throw InputError("timing dump: unknown node \"" + g + "\"");
      const auto k = pin_key.find(g);
      if (k != pin_key.end())
        for (const PinEntry &e : kPins)
          if (k->second == e.cell)
            for (int j = 0; j < 5 && e.pins[j]; ++j)
              if (p == e.pins[j]) return {gi->second, e.ids[j]};
      ++st.unknown_pins;
      return {gi->second, 0};
    }
    if (pi_set.count(s)) return {id[s], 0};
    if (po_set.count(s)) return {id[s], 1};
    throw InputError("timing dump: unknown gate:pin \"" + s + "\"");
  };
  // Loop removal (parser.py:163-211).
  {
    static const char *kPost[] = {"", ":Z", ":ZN", ":Q", ":Y", ":QN"};
    std::deque<std::string> queue(pis.begin(), pis.end());
    std::unordered_set<std::string> visited;
    while (!queue.empty()) {
      std::string cur;
      if (dfs) {
        cur = std::move(queue.back());
        queue.pop_back();
      } else {
        cur = std::move(queue.front());
        queue.pop_front();
      }
      if (!visited.insert(cur).second) continue;
      Dests *outs = nullptr;
      for (const char *post : kPost) {
        auto it = conn.find(cur + post);
        if (it != conn.end()) {
          outs = &it->second;
          break;
        }
      }
      if (!outs) {
        if (!po_set.count(cur)) ++st.unreached;
        continue;
      }
      std::vector<std::string> names;
      for (const std::string &d : outs->v) names.emplace_back(before_colon(d));
      for (const std::string &out : names) {
        if (visited.count(out)) {
          Dests kept;
          for (const std::string &d : outs->v)
            if (starts_with(d, out)) ++st.removed; else kept.add(d);
          *outs = std::move(kept);
        } else {
          queue.push_back(out);
        }
      }
    }
  }
  std::string o;
  TextOut w(o);
  w.u(uint64_t(pis.size())).c('\n').u(uint64_t(pos.size())).c('\n').u(uint64_t(cells.size())).c('\n');
  w.u(uint64_t(keys.size())).c('\n');
  for (const std::string &k : keys) {
    const auto src = gate_and_pin(k);
    const Dests &d = conn[k];
    w.u(uint64_t(src.first)).c(' ').u(uint64_t(src.second)).c(' ').u(uint64_t(d.v.size()));
    for (const std::string &x : d.v) {
      const auto dst = gate_and_pin(x);
      w.c(' ').u(uint64_t(dst.first)).c(' ').u(uint64_t(dst.second));
    }
    w.c('\n');
  }
  for (size_t g = 0; g < types.size(); ++g) {
    if (g) w.c(' ');
    w.u(uint64_t(types[g]));
  }
  w.c('\n');
  st.pis = int64_t(pis.size());
  st.pos = int64_t(pos.size());
  st.cells = int64_t(cells.size());
  st.records = int64_t(keys.size());
  if (stats) *stats = st;
  return o;
}

}  // namespace lsim
