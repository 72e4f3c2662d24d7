"""Restatement of the reference's netlist converter (fsim/parser.py:1-227):
OpenTimer timing-graph dump -> .ckt text.

TEST INFRASTRUCTURE ONLY: the checker for lsim_convert_timing_dump. It follows
parser.py step by step, with one difference the product shares: a source's
destinations are kept in dump order (parser.py keeps them in a Python set,
whose iteration order is not fixed from run to run).

Each step cites the parser.py lines it follows."""
from collections import deque

# parser.py:36-54
PIN_D = {
    'CLKBUF': {'A': 1, 'Z': 0},
    'INV': {'A': 1, 'ZN': 0},
    'AND2': {'A': 1, 'B': 2, 'Y': 0},
    'NAND2': {'A1': 1, 'A2': 2, 'ZN': 0},
    'NAND3': {'A1': 1, 'A2': 2, 'A3': 3, 'ZN': 0},
    'NAND4': {'A1': 1, 'A2': 2, 'A3': 3, 'A4': 4, 'ZN': 0},
    'NOR2': {'A1': 1, 'A2': 2, 'ZN': 0},
    'NOR3': {'A1': 1, 'A2': 2, 'A3': 3, 'ZN': 0},
    'NOR4': {'A1': 1, 'A2': 2, 'A3': 3, 'A4': 4, 'ZN': 0},
    'AOI21': {'A': 1, 'B1': 2, 'B2': 3, 'ZN': 0},
    'AOI22': {'A1': 1, 'A2': 2, 'B1': 3, 'B2': 4, 'ZN': 0},
    'DFF': {'D': 1, 'CK': 2, 'Q': 0},
    'OAI21': {'A': 1, 'B1': 2, 'B2': 3, 'ZN': 0},
    'OAI22': {'A1': 1, 'A2': 2, 'B1': 3, 'B2': 4, 'ZN': 0},
    'AND2X1': {'A': 1, 'B': 2, 'Y': 0},
    'INVX1': {'A': 1, 'Y': 0},
    'NAND3X1': {'A': 1, 'B': 2, 'C': 3, 'Y': 0},
}

# parser.py:58-74 (dict order is the prefix search order)
GATE_ENUM = {'INV': 0, 'AND': 1, 'OR': 2, 'XOR': 3, 'NAND': 4, 'NOR': 5, 'XNOR': 6, 'MUX': 7, 'CLKBUF': 8,
             'AOI21': 4, 'AOI22': 4, 'DFF': 4, 'OAI21': 5, 'OAI22': 5}


def convert(text: str, mode: str = "bfs") -> str:
    lines = text.split("\n")
    # parser.py:11-25
    i_pi = lines.index("// Start PIs")
    i_po = lines.index("// Start POs")
    i_wire = lines.index("// Start wires")
    i_cell = lines.index("// Start cells")
    i_dg = lines.index("digraph TimingGraph {")
    pis = [x.split()[1][:-1] for x in lines[i_pi + 1:i_po - 1]]
    pos = [x.split()[1][:-1] for x in lines[i_po + 1:i_wire - 1]]
    _wires = [x.split()[1][:-1] for x in lines[i_wire + 1:i_cell - 1]]
    cells = [(x.split()[1], x.split()[0]) for x in lines[i_cell + 1:i_dg - 2]]
    connections = [tuple(y.strip(' ";') for y in x.split(" -> ")) for x in lines[i_dg + 1:-2] if "->" in x]
    # parser.py:77, 87-97
    pin_key = {c[0]: c[1].split('_')[0] for c in cells}

    def enum_id(s):
        for e, v in GATE_ENUM.items():
            if s.startswith(e):
                return v
        return 1

    types = [enum_id(c[1]) for c in cells]
    # parser.py:96-101 (dict of dicts as ordered sets)
    conn = {}
    for x in connections:
        if x[0].split(":")[0] == x[1].split(":")[0]:
            continue
        conn.setdefault(x[0], {})[x[1]] = None
    # parser.py:104
    gate_d = {y: i for i, y in enumerate(pis + pos + [c[0] for c in cells])}

    # parser.py:138-158
    def gate_and_pin(s):
        if ":" in s:
            g, p = s.split(":")
            if g in pin_key:
                k = pin_key[g]
                if k in PIN_D and p in PIN_D[k]:
                    return gate_d[g], PIN_D[k][p]
                return gate_d[g], 0
            return gate_d[g], 0
        if s in pis:
            return gate_d[s], 0
        if s in pos:
            return gate_d[s], 1
        raise KeyError(s)

    # parser.py:163-211
    source = deque(pis)
    visited = set()
    while source:
        cur = source.popleft() if mode == "bfs" else source.pop()
        if cur in visited:
            continue
        visited.add(cur)
        for post in ['', ':Z', ':ZN', ':Q', ':Y', ':QN']:
            if cur + post in conn:
                outputs = conn[cur + post]
                break
        else:
            continue
        outputs = [x.split(':')[0] for x in outputs]
        for out in outputs:
            if out in visited:
                conn[cur + post] = {x: None for x in conn[cur + post] if not x.startswith(out)}
            else:
                source.append(out)
    # parser.py:213-227
    out = [str(len(pis)), str(len(pos)), str(len(cells)), str(len(conn))]
    for k, v in conn.items():
        a, b = gate_and_pin(k)
        rec = [str(a), str(b), str(len(v))]
        for x in v:
            c, d = gate_and_pin(x)
            rec += [str(c), str(d)]
        out.append(" ".join(rec))
    out.append(" ".join(map(str, types)))
    return "\n".join(out) + "\n"
