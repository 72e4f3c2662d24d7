// Loop removal on a .ckt netlist by traversal from the PIs, the rule the
// reference's netlist producer applies before writing a .ckt
// (fsim/parser.py:163-211): nodes are taken from a queue seeded with the PIs,
// from its front (BFS) or its back (DFS); a node already taken is skipped;
// otherwise every wire from it to a node already taken is dropped and the
// other destinations are appended to the queue. Every kept wire then runs
// from an earlier-taken node to a later-taken one, so the nodes reached from
// the PIs carry no cycle.
//
// The .ckt has node ids where the producer has cell and pin names:
//  * a node's outgoing wires are its first wire record in file order (the
//    producer follows one output pin per cell, the first of '', Z, ZN, Q, Y,
//    QN it finds);
//  * a dropped destination is matched by id, where the producer drops every
//    pin whose name starts with the destination's name.
#include <deque>
#include <string>
#include <vector>

#include "loops.hpp"
#include "netlist.hpp"
#include "textio.hpp"

namespace lsim {

namespace {

uint64_t need(TextIn &in, const char *what) {
  uint64_t v = 0;
  if (!in.next(v)) throw InputError(std::string(in.bad() ? "malformed " : "truncated ") + what);
  return v;
}

struct Record {
  uint64_t src = 0, ypin = 0;
  std::vector<uint64_t> dst, pin;
};

}  // namespace

std::string break_loops(const char *ckt, size_t len, bool dfs, LoopStats *stats) {
  TextIn in(ckt, len);
  const uint64_t P = need(in, "circuit header"), O = need(in, "circuit header"), G = need(in, "circuit header"),
                 W = need(in, "circuit header");
  const uint64_t N = P + O + G;
  if (N >= (uint64_t(1) << 31)) throw InputError("circuit too large (node count must fit in int32)");
  std::vector<Record> rec(W);
  std::vector<int64_t> first(N, -1);  // a node's first wire record
  for (uint64_t w = 0; w < W; ++w) {
    Record &r = rec[w];
    r.src = need(in, "wire record");
    r.ypin = need(in, "wire record");
    const uint64_t k = need(in, "wire record");
    if (r.src >= N) throw InputError("wire source id out of range");
    for (uint64_t j = 0; j < k; ++j) {
      const uint64_t d = need(in, "wire record");
      const uint64_t p = need(in, "wire record");
      if (d >= N) throw InputError("wire destination id out of range");
      r.dst.push_back(d);
      r.pin.push_back(p);
    }
    if (first[r.src] < 0) first[r.src] = int64_t(w);
  }
  std::vector<uint64_t> types(G);
  for (uint64_t g = 0; g < G; ++g) types[g] = need(in, "gate types");

  LoopStats st;
  std::vector<uint8_t> taken(N, 0);
  std::deque<uint64_t> queue;
  for (uint64_t i = 0; i < P; ++i) queue.push_back(i);
  std::vector<uint64_t> outs;
  while (!queue.empty()) {
    uint64_t cur;
    if (dfs) {
      cur = queue.back();
      queue.pop_back();
    } else {
      cur = queue.front();
      queue.pop_front();
    }
    if (taken[cur]) continue;
    taken[cur] = 1;
    ++st.visited;
    if (first[cur] < 0) continue;  // a PO, or a node that drives nothing
    Record &r = rec[size_t(first[cur])];
    outs = r.dst;  // the destinations as they were when the node was taken
    for (uint64_t out : outs) {
      if (taken[out]) {
        size_t keep = 0;
        for (size_t j = 0; j < r.dst.size(); ++j)
          if (r.dst[j] != out) {
            r.dst[keep] = r.dst[j];
            r.pin[keep] = r.pin[j];
            ++keep;
          } else {
            ++st.removed;
          }
        r.dst.resize(keep);
        r.pin.resize(keep);
      } else {
        queue.push_back(out);
      }
    }
  }

  std::string text;
  TextOut o(text);
  o.u(P).c(' ').u(O).c(' ').u(G).c(' ').u(W).c('\n');
  for (const Record &r : rec) {
    o.u(r.src).c(' ').u(r.ypin).c(' ').u(r.dst.size());
    for (size_t j = 0; j < r.dst.size(); ++j) o.c(' ').u(r.dst[j]).c(' ').u(r.pin[j]);
    o.c('\n');
  }
  for (uint64_t g = 0; g < G; ++g) {
    if (g) o.c(' ');
    o.u(types[g]);
  }
  o.c('\n');
  if (stats) *stats = st;
  return text;
}

}  // namespace lsim
