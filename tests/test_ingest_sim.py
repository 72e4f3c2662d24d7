"""Real-circuit ingestion end to end: an OpenTimer timing-graph dump with
loops and reconvergent paths -> .ckt (fsim/parser.py:1-227, BFS and DFS loop
removal) -> simulation.

The traversal drops every arc into a node taken before (parser.py:198-204),
so cells of looped or reconvergent dumps keep non-dense pin sets: a pin slot
no kept arc fills, which the reference reads as node id 0 through its
zero-filled pin table (fsim/simulation.cpp:676-711,851-877), and pins above
the kept count, which it does not read. The CPU tests check that the reader
accepts such netlists with the oracle's order, levels and good values; the GPU
test simulates them against the oracle's literal restatement on every fault
and round."""
import random

import numpy as np
import pytest

import lsim_amd
import oracle
from oracle import ingest

# Cells whose pins the converter's table knows (parser.py:36-54), with the
# gate code each maps to (parser.py:58-74).
CELLS = [("NAND2_X1", ["A1", "A2"], "ZN"), ("NOR3_X2", ["A1", "A2", "A3"], "ZN"), ("INV_X1", ["A"], "ZN"),
         ("AOI21_X1", ["A", "B1", "B2"], "ZN"), ("OAI22_X1", ["A1", "A2", "B1", "B2"], "ZN"),
         ("DFF_X1", ["D", "CK"], "Q"), ("CLKBUF_X1", ["A"], "Z"), ("AND2X1", ["A", "B"], "Y"),
         ("NAND4_X1", ["A1", "A2", "A3", "A4"], "ZN"), ("NOR2_X1", ["A1", "A2"], "ZN"),
         ("NAND3X1", ["A", "B", "C"], "Y"), ("INVX1", ["A"], "Y")]


def looped_dump(seed, P, O, C, back=0.12):
    """A dump whose cells read PIs and earlier cells, plus arcs from any cell
    (loops and forward skips); fixed-width names, so no name prefixes
    another. Every cell's first pin reads a PI or an earlier cell, so every
    cell is reached from the PIs."""
    rng = random.Random(seed)
    pis = [f"pi{i:04d}" for i in range(P)]
    pos = [f"po{i:04d}" for i in range(O)]
    kinds = [rng.choice(CELLS) for _ in range(C)]
    names = [f"g{c:05d}" for c in range(C)]
    arcs, cells = [], []
    for c, (t, ins, out) in enumerate(kinds):
        cells.append((names[c], t))
        for j, pin in enumerate(ins):
            if j and rng.random() < back:
                k = rng.randrange(C)
                src = f"{names[k]}:{kinds[k][2]}"
            elif c == 0 or rng.random() < 0.25:
                src = rng.choice(pis)
            else:
                k = rng.randrange(max(0, c - 40), c)
                src = f"{names[k]}:{kinds[k][2]}"
            arcs.append((src, f"{names[c]}:{pin}"))
            arcs.append((f"{names[c]}:{pin}", f"{names[c]}:{out}"))
    for i, p in enumerate(pos):
        k = C - 1 - i * (C // (2 * O))
        arcs.append((f"{names[k]}:{kinds[k][2]}", p))
    rng.shuffle(arcs)
    lines = ["module top;", "// Start PIs"] + [f"input {p};" for p in pis] + ["", "// Start POs"]
    lines += [f"output {p};" for p in pos] + ["", "// Start wires", "wire w0;", "", "// Start cells"]
    lines += [f"{t} {n} ( );" for n, t in cells] + ["endmodule", "", "digraph TimingGraph {"]
    lines += [f'  "{a}" -> "{b}";' for a, b in arcs] + ["}", ""]
    return "\n".join(lines)


def _inputs(ckt: bytes, seed: int, n_patterns: int):
    P, O, G = (int(x) for x in ckt.split()[:3])
    faults = [(P + O + g, v) for g in range(G) for v in (0, 1)] + [(i, 1) for i in range(P)]
    flst = f"{len(faults)}\n" + "".join(f"{i} {v}\n" for i, v in faults)
    ptn = lsim_amd.generate_patterns(seed, P, n_patterns)
    return flst, ptn


CASES = [(11, 12, 6, 300, "bfs"), (12, 12, 6, 300, "dfs"), (13, 24, 8, 2000, "bfs"), (14, 24, 8, 2000, "dfs")]


def _gappy(ckt: bytes) -> int:
    """Gates whose pin set is not dense 1..k."""
    toks = [int(x) for x in ckt.split()]
    P, O, G, W = toks[:4]
    pins = {}
    i = 4
    for _ in range(W):
        k = toks[i + 2]
        for j in range(k):
            pins.setdefault(toks[i + 3 + 2 * j], []).append(toks[i + 4 + 2 * j])
        i += 3 + 2 * k
    return sum(1 for d, ps in pins.items() if d >= P + O and sorted(ps) != list(range(1, len(ps) + 1)))


@pytest.mark.parametrize("seed,P,O,C,mode", CASES)
def test_looped_dump_reads_like_the_oracle(seed, P, O, C, mode):
    dump = looped_dump(seed, P, O, C)
    ckt, st = lsim_amd.convert_timing_dump(dump, mode)
    assert ckt.decode() == ingest.convert(dump, mode)
    assert st["removed"] > 0 and st["unknown_pins"] == 0 and st["unknown_cells"] == 0
    assert _gappy(ckt) > 0
    flst, ptn = _inputs(ckt, seed, 64 * 3 + 9)
    s = lsim_amd.Simulator(device=-1)
    s.read_text(ckt, flst, ptn)
    o = oracle.Oracle(ckt, flst, ptn)
    assert np.array_equal(s.gate_idx_to_tpg_order_idx(), o.order_map())
    assert np.array_equal(s.levels(), o.levels())
    assert np.array_equal(s.num_gates_per_level(), o.gates_per_level())


@pytest.mark.gpu
@pytest.mark.parametrize("seed,P,O,C,mode", CASES)
def test_looped_dump_simulates_like_the_oracle(seed, P, O, C, mode):
    dump = looped_dump(seed, P, O, C)
    ckt, _ = lsim_amd.convert_timing_dump(dump, mode)
    flst, ptn = _inputs(ckt, seed, 64 * 3 + 9)
    ref = oracle.Oracle(ckt, flst, ptn).run_literal(threads=0)
    for direct in (0, 1):
        s = lsim_amd.Simulator(device=0)
        s.set_option("direct", direct)
        s.read_text(ckt, flst, ptn)
        s.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
        assert np.array_equal(s.detect_matrix(), ref["det"]), direct
        assert np.array_equal(s.found_fault_to_pattern(), ref["table"]), direct
        pi, gate, po = s.good_results()
        assert np.array_equal(gate, ref["g_gate"]) and np.array_equal(po, ref["g_po"])
        s.close()
