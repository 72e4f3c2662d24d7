"""Restatement of the reference netlist producer's loop removal (TEST
INFRASTRUCTURE ONLY): the checker for lsim_break_loops.

fsim/parser.py:163-211 takes nodes from a deque seeded with the PIs
(`popleft` for 'bfs', `pop` for 'dfs'), skips nodes already taken, and for a
taken node drops every connection to an already-taken destination and
appends the others. Here on .ckt text: node ids instead of cell:pin names, a
node's connections are its first wire record.
"""
from __future__ import annotations

from collections import deque
from typing import List, Tuple


def parse_ckt(text: bytes):
    toks = [int(x) for x in text.split()]
    P, O, G, W = toks[:4]
    i = 4
    recs = []
    for _ in range(W):
        src, ypin, k = toks[i], toks[i + 1], toks[i + 2]
        i += 3
        pairs = [(toks[i + 2 * j], toks[i + 2 * j + 1]) for j in range(k)]
        i += 2 * k
        recs.append([src, ypin, pairs])
    types = toks[i:i + G]
    return P, O, G, recs, types


def break_loops(text: bytes, mode: str = "bfs") -> Tuple[bytes, int, List[int]]:
    """(.ckt text, dropped entries, order in which nodes were taken)."""
    P, O, G, recs, types = parse_ckt(text)
    first = {}
    for idx, r in enumerate(recs):
        first.setdefault(r[0], idx)
    source = deque(range(P))
    visited = set()
    order = []
    removed = 0
    while source:
        cur = source.popleft() if mode == "bfs" else source.pop()   # fsim/parser.py:169-172
        if cur in visited:                                          # :176-177
            continue
        visited.add(cur)
        order.append(cur)
        if cur not in first:                                        # :184-192 (PO or no wires)
            continue
        rec = recs[first[cur]]
        outputs = [d for d, _ in rec[2]]                            # :201
        for out in outputs:                                         # :202-208
            if out in visited:
                before = len(rec[2])
                rec[2] = [(d, p) for d, p in rec[2] if d != out]
                removed += before - len(rec[2])
            else:
                source.append(out)
    lines = [f"{P} {O} {G} {len(recs)}"]
    for src, ypin, pairs in recs:
        lines.append(" ".join([str(src), str(ypin), str(len(pairs))] + [f"{d} {p}" for d, p in pairs]))
    lines.append(" ".join(map(str, types)))
    return ("\n".join(lines) + "\n").encode(), removed, order
