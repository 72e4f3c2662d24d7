"""Structural diagnostics per level (the idea of
fsim/cuda_MA_partition/cuda_MA_partition.cu:227-346 at wave64): pinned by a
hand-computed circuit and by the restatement oracle/diagnostics.py on a
generated circuit, over fanins rebuilt from the .ckt wires."""
import numpy as np
import pytest

import lsim_amd
import oracle
from lsim_amd import workloads
from oracle import diagnostics


def test_one_gate_known_answer():
    s = lsim_amd.Simulator(device=-1)
    s.read_text("2 1 1 3 1 0 1 3 1 0 0 1 3 2 3 0 1 2 1 1", "0", "4 6 10")
    d = s.level_diagnostics()
    # level 0: two PIs; level 1: the AND over positions 0, 1; level 2: the PO over position 2
    assert list(d["nodes"]) == [2, 1, 1]
    assert list(d["avg_dist"]) == [0.0, 1.0, 1.0]
    assert list(d["max_span"]) == [0, 1, 1]
    assert list(d["avg_range"]) == [0.0, 1.0, 0.0]
    assert list(d["avg_unique"]) == [0.0, 2.0, 1.0]


def _fanins(ckt, order_map, N):
    vals = [int(x) for x in ckt.split()]
    fan = [[] for _ in range(N)]
    i = 4
    for _ in range(vals[3]):
        src, k = vals[i], vals[i + 2]
        for j in range(k):
            dst, pin = vals[i + 3 + 2 * j], vals[i + 4 + 2 * j]
            fan[int(order_map[dst])].append((3 if pin == 5 else pin, int(order_map[src])))
        i += 3 + 2 * k
    return [[q for _, q in sorted(f)] for f in fan]


@pytest.mark.parametrize("wave", [64, 32])
def test_matches_restatement(wave):
    ckt, flst, ptn = workloads.texts("c1")
    s = lsim_amd.Simulator(device=-1)
    s.read_text(ckt, flst, ptn)
    o = oracle.Oracle(ckt, flst, ptn)
    N = o.N
    levels = [int(x) for x in s.levels()]
    want = diagnostics.level_diagnostics(levels, _fanins(ckt, o.order_map(), N), wave)
    got = s.level_diagnostics(wave)
    for k in want:
        assert np.allclose(got[k], want[k], rtol=1e-12, atol=0), k
    assert int(got["nodes"].sum()) == N
