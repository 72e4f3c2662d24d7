// Loop removal on .ckt netlists by BFS/DFS traversal from the PIs.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace lsim {

struct LoopStats {
  int64_t visited = 0;  // nodes taken from the queue
  int64_t removed = 0;  // (wire, destination pin) entries dropped
};

// Returns the .ckt text with the wires into already-taken nodes dropped
// (BFS when dfs is false). Throws InputError on malformed text.
std::string break_loops(const char *ckt, size_t len, bool dfs, LoopStats *stats);

}  // namespace lsim
