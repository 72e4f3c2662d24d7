// Netlist ingestion: an OpenTimer timing-graph dump to a .ckt netlist.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace lsim {

struct IngestStats {
  int64_t pis = 0, pos = 0, cells = 0, wires = 0;
  int64_t connections = 0;    // "a -> b" lines read
  int64_t records = 0;        // wire records written (distinct source pins)
  int64_t removed = 0;        // destinations dropped by the loop removal
  int64_t unknown_cells = 0;  // cell types mapped to AND by default
  int64_t unknown_pins = 0;   // pins mapped to pin 0 by default
  int64_t unreached = 0;      // nodes taken from the queue that drive no connection (POs excluded)
};

// Returns the .ckt text (BFS traversal when dfs is false). Throws InputError
// on text the conversion cannot read.
std::string convert_timing_dump(const char *text, size_t len, bool dfs, IngestStats *stats);

}  // namespace lsim
