"""Simulator::visualization_graph (fsim/simulation.hpp:93,
fsim/simulation.cpp:1157-1225), called by the reference CLI
(main/simple.cpp:62-65): the Graphviz text of the circuit.

Pinned by two hand-derived answers, by the text the compiled reference writes
for the reference's own unit-test circuits and reference-made cases, and on a
generated circuit by an edge list rebuilt from the .ckt wires and the oracle's
order map."""
import os
import re

import numpy as np
import pytest

import lsim_amd
import oracle
from lsim_amd import workloads


def _host(ckt, flst, ptn):
    s = lsim_amd.Simulator(device=-1)
    s.read_text(ckt, flst, ptn)
    return s


HEAD = "digraph G {\n\tsubgraph cluster_0 {\n\t\tcolor=black;\n\t\tnode [style=filled,color=lightgrey];\n"


def test_one_gate_known_answer():
    # PI 0, 1; PO 2; gate 3 = AND(pin 1: node 1, pin 2: node 0) -> PO 2.
    s = _host("2 1 1 3 1 0 1 3 1 0 0 1 3 2 3 0 1 2 1 1", "3 1 1 2 0 0 0", "4 6 10")
    want = (HEAD + "\t\tPI_1 -> Gate_0_AND;\n\t\tPI_0 -> Gate_0_AND;\n\t\tGate_0_AND -> PO_0;\n\t}\n\n"
            "\tPI_0  [shape=Msquare];\n\tPI_1  [shape=Msquare];\n\tPO_0  [shape=Msquare];\n}\n\n")
    assert s.visualization_text() == want


def test_two_gates_known_answer():
    # Gate 3 = NAND(0, 1); gate 4 = XOR(3, 1) -> PO 2. FIFO Kahn from the PIs
    # takes 0, 1, 3, 4, so gate 3 is Gate_0 and gate 4 is Gate_1.
    ckt = "2 1 2 4\n0 0 1 3 1\n1 0 2 3 2 4 2\n3 0 1 4 1\n4 0 1 2 1\n4 3\n"
    s = _host(ckt, "0", "4 6 10")
    want = (HEAD + "\t\tPI_0 -> Gate_0_NAND;\n\t\tPI_1 -> Gate_0_NAND;\n"
            "\t\tGate_0_NAND -> Gate_1_XOR;\n\t\tPI_1 -> Gate_1_XOR;\n\t\tGate_1_XOR -> PO_0;\n\t}\n\n"
            "\tPI_0  [shape=Msquare];\n\tPI_1  [shape=Msquare];\n\tPO_0  [shape=Msquare];\n}\n\n")
    assert s.visualization_text() == want


def test_writes_file_and_reports_unopenable_path(tmp_path):
    s = _host("2 1 1 3 1 0 1 3 1 0 0 1 3 2 3 0 1 2 1 1", "3 1 1 2 0 0 0", "4 6 10")
    path = tmp_path / "g.dot"
    s.visualization_graph(str(path))
    assert path.read_text() == s.visualization_text()
    with pytest.raises(lsim_amd.LsimError) as e:
        s.visualization_graph(str(tmp_path / "no_such_dir" / "g.dot"))
    assert e.value.code == lsim_amd.LSIM_E_IO
    assert "Error opening output file" in e.value.message


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")
def test_matches_compiled_reference(unit_vectors, ref_vectors):
    for c in list(unit_vectors) + list(ref_vectors):
        s = _host(c["ckt"], c["flst"], c["ptn"])
        assert s.visualization_text() == oracle.reference_dot(c["ckt"], c["flst"], c["ptn"]), c["name"]


def test_generated_circuit_edges():
    """C1-shaped circuit: every gate and PO lists its fanins in pin order."""
    ckt, flst, ptn = workloads.texts("c1")
    s = _host(ckt, flst, ptn)
    text = s.visualization_text()
    o = oracle.Oracle(ckt, flst, ptn)
    P, O, G = o.P, o.O, o.G
    pos = o.order_map().astype(np.int64)  # file id -> position
    vals = [int(x) for x in ckt.split()]
    W = vals[3]
    fan = {}
    i = 4
    for _ in range(W):
        src, k = vals[i], vals[i + 2]
        for j in range(k):
            dst, pin = vals[i + 3 + 2 * j], vals[i + 4 + 2 * j]
            fan.setdefault(int(pos[dst]), []).append((3 if pin == 5 else pin, int(pos[src])))
        i += 3 + 2 * k
    types = vals[i:i + G]
    type_of = {}
    for gid, t in zip(range(P + O, P + O + G), types):
        type_of[int(pos[gid])] = ["INV", "AND", "OR", "XOR", "NAND", "NOR", "XNOR", "MUX", "CLKBUF"][t]

    def name(q):
        return f"PI_{q}" if q < P else f"Gate_{q - P}_{type_of[q]}"

    want = []
    for p in list(range(P, P + G)) + list(range(P + G, P + G + O)):
        now = f"Gate_{p - P}_{type_of[p]}" if p < P + G else f"PO_{p - P - G}"
        for _, q in sorted(fan.get(p, [])):
            want.append(f"{name(q)} -> {now};")
    got = re.findall(r"^\t\t(\S+ -> \S+;)$", text, flags=re.M)
    assert got == want
    assert text.count("[shape=Msquare]") == P + O
