"""Timing-dump ingestion (fsim/parser.py:1-227): OpenTimer dump -> .ckt.

Pinned by hand-derived answers (a chain of mapped cells; a loop broken by the
BFS traversal), by the plain-Python restatement of parser.py (oracle/ingest.py)
on seeded random dumps with loops, reconvergence, unknown cells and pins,
cell-internal arcs and name-prefix collisions (g1 / g10), in both traversal
orders, and by reading the converted netlists back."""
import random

import numpy as np
import pytest

import lsim_amd
from oracle import ingest


def _dump(pis, pos, wires, cells, arcs):
    lines = ["module top;", "// Start PIs"] + [f"input {p};" for p in pis] + ["", "// Start POs"]
    lines += [f"output {p};" for p in pos] + ["", "// Start wires"] + [f"wire {w};" for w in wires]
    lines += ["", "// Start cells"] + [f"{t} {n} ( );" for n, t in cells] + ["endmodule", "", "digraph TimingGraph {"]
    lines += [f'  "{a}" -> "{b}";' for a, b in arcs] + ["}", ""]
    return "\n".join(lines)


CHAIN = _dump(["a", "b"], ["z"], ["n1", "n2"], [("g1", "NAND2_X1"), ("g2", "INV_X1"), ("g3", "CLKBUF_X1")],
              [("a", "g1:A1"), ("b", "g1:A2"), ("g1:A1", "g1:ZN"), ("g1:A2", "g1:ZN"), ("g1:ZN", "g2:A"),
               ("g2:A", "g2:ZN"), ("g2:ZN", "g3:A"), ("g3:A", "g3:Z"), ("g3:Z", "z")])


def test_chain_known_answer():
    ckt, st = lsim_amd.convert_timing_dump(CHAIN)
    # ids a=0 b=1 z=2 g1=3 g2=4 g3=5; NAND2 A1/A2 -> pins 1/2, INV A -> 1,
    # CLKBUF A -> 1, outputs -> 0, a PO destination -> pin 1.
    assert ckt.decode() == "2\n1\n3\n5\n0 0 1 3 1\n1 0 1 3 2\n3 0 1 4 1\n4 0 1 5 1\n5 0 1 2 1\n4 0 8\n"
    assert st["removed"] == 0 and st["records"] == 5 and st["connections"] == 9
    assert st["unknown_cells"] == 0 and st["unknown_pins"] == 0
    # z = CLKBUF(INV(NAND(a, b))) = a AND b
    s = lsim_amd.Simulator(device=-1)
    s.read_text(ckt, "0", "4 12 10")  # a = 1100, b = 1010 over 4 patterns
    assert list(s.levels()) == [0, 0, 1, 2, 3, 4]


def test_loop_broken_by_bfs_known_answer():
    dump = _dump(["a"], ["z"], [], [("g1", "NAND2_X1"), ("g2", "INV_X1"), ("g3", "CLKBUF_X1")],
                 [("a", "g1:A1"), ("g1:ZN", "g2:A"), ("g2:ZN", "g1:A2"), ("g2:ZN", "g3:A"), ("g3:Z", "z")])
    ckt, st = lsim_amd.convert_timing_dump(dump)
    # BFS takes a, g1, g2: g2's arc back into g1 (taken) is dropped.
    assert ckt.decode() == "1\n1\n3\n4\n0 0 1 2 1\n2 0 1 3 1\n3 0 1 4 1\n4 0 1 1 1\n4 0 8\n"
    assert st["removed"] == 1
    assert ckt.decode() == ingest.convert(dump, "bfs")


CELLS = [("NAND2_X1", ["A1", "A2"], "ZN"), ("NOR3_X2", ["A1", "A2", "A3"], "ZN"), ("INV_X1", ["A"], "ZN"),
         ("AOI21_X1", ["A", "B1", "B2"], "ZN"), ("OAI22_X1", ["A1", "A2", "B1", "B2"], "ZN"),
         ("DFF_X1", ["D", "CK"], "Q"), ("CLKBUF_X1", ["A"], "Z"), ("AND2X1", ["A", "B"], "Y"),
         ("XOR2_X1", ["A", "B"], "Z"), ("FOO_X1", ["A"], "Y"), ("NAND3X1", ["A", "B", "C"], "Y")]


def _random_dump(seed, P, O, C, back=0.05):
    rng = random.Random(seed)
    pis = [f"pi{i}" for i in range(P)]
    pos = [f"po{i}" for i in range(O)]
    cells, arcs = [], []
    kinds = [rng.choice(CELLS) for _ in range(C)]
    for c, (t, ins, out) in enumerate(kinds):
        name = f"g{c}"
        cells.append((name, t))
        for pin in ins:
            if c and rng.random() < back:  # from any cell: loops and forward arcs
                k = rng.randrange(C)
                src = f"g{k}:{kinds[k][2]}"
            elif c == 0 or rng.random() < 0.3:
                src = rng.choice(pis)
            else:
                k = rng.randrange(c)
                src = f"g{k}:{kinds[k][2]}"
            arcs.append((src, f"{name}:{pin}"))
            arcs.append((f"{name}:{pin}", f"{name}:{out}"))
    for i, p in enumerate(pos):
        k = rng.randrange(C)
        arcs.append((f"g{k}:{kinds[k][2]}", p))
    arcs += [arcs[rng.randrange(len(arcs))] for _ in range(len(arcs) // 20)]  # repeated arcs
    rng.shuffle(arcs)
    return _dump(pis, pos, [f"w{i}" for i in range(C // 2)], cells, arcs)


@pytest.mark.parametrize("mode", ["bfs", "dfs"])
@pytest.mark.parametrize("seed,P,O,C", [(1, 4, 2, 12), (2, 8, 4, 60), (3, 16, 8, 400), (4, 3, 3, 1500)])
def test_matches_restatement_on_random_dumps(seed, P, O, C, mode):
    dump = _random_dump(seed, P, O, C)
    ckt, st = lsim_amd.convert_timing_dump(dump, mode)
    assert ckt.decode() == ingest.convert(dump, mode)
    assert st["pis"] == P and st["pos"] == O and st["cells"] == C


def test_converted_netlist_reads_back():
    """A layered dump (every cell's inputs one level up, fixed-width names)
    loses no arc to the BFS traversal, and converts into a netlist the reader
    accepts with the dump's structure: each cell a gate of the mapped type
    whose fanins are the dump's sources in pin order. (The traversal drops an
    arc into any node taken before it, so reconvergent paths of unequal length
    lose arcs, and a name that prefixes another drops that one's pins too:
    parser.py:198-204.)"""
    rng = random.Random(7)
    known = [c for c in CELLS if c[0].split("_")[0] in ingest.PIN_D]
    P, O, layers, width = 8, 4, 6, 10
    pis = [f"pi{i:03d}" for i in range(P)]
    pos = [f"po{i:03d}" for i in range(O)]
    prev = [(p, p) for p in pis]  # (node name, output arc name)
    cells, arcs, fanins = [], [], {}
    for layer in range(layers):
        cur = []
        for j in range(width):
            t, ins, out = rng.choice(known)
            name = f"c{layer:02d}{j:03d}"
            cells.append((name, t))
            fanins[name] = []
            for pin in ins:
                node, arc = rng.choice(prev)
                arcs.append((arc, f"{name}:{pin}"))
                fanins[name].append(node)
            cur.append((name, f"{name}:{out}"))
        prev = cur
    for i, p in enumerate(pos):
        arcs.append((prev[i][1], p))
    dump = _dump(pis, pos, [], cells, arcs)
    ckt, st = lsim_amd.convert_timing_dump(dump, "bfs")
    assert st["removed"] == 0 and st["unknown_pins"] == 0 and st["unknown_cells"] == 0
    assert ckt.decode() == ingest.convert(dump, "bfs")
    s = lsim_amd.Simulator(device=-1)
    words = " ".join(str(rng.getrandbits(64)) for _ in range(P))
    s.read_text(ckt, "0", f"64 {words}")
    assert s.num_gates() == len(cells) and s.num_PIs() == P and s.num_POs() == O
    assert list(s.levels()[P:P + len(cells)]) == [1 + k // width for k in range(len(cells))]
    # the DOT export lists every cell's fanins in pin order
    ids = {n: i for i, n in enumerate(pis + pos + [c[0] for c in cells])}
    order = s.gate_idx_to_tpg_order_idx()
    names = {int(order[ids[n]]): n for n in ids}
    edges = [line.strip().rstrip(";").split(" -> ") for line in s.visualization_text().splitlines()
             if "-> Gate_" in line]
    got = {}
    for a_, b_ in edges:
        dst = names[P + int(b_.split("_")[1])]
        src = names[int(a_.split("_")[1])] if a_.startswith("PI_") else names[P + int(a_.split("_")[1])]
        got.setdefault(dst, []).append(src)
    assert got == fanins


def test_rejects_dump_without_sections():
    with pytest.raises(lsim_amd.LsimError) as e:
        lsim_amd.convert_timing_dump("module top;\nendmodule\n")
    assert e.value.code == lsim_amd.LSIM_E_INPUT
    assert "// Start PIs" in e.value.message
