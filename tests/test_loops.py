"""Loop removal by BFS/DFS traversal from the PIs (fsim/parser.py:163-211):
hand-worked cases, the restatement in oracle/loops.py on generated circuits,
and the property that every kept wire runs forward in traversal order."""
import random

import pytest

import lsim_amd
from oracle import loops as ref


def _norm(t: bytes) -> list:
    return [int(x) for x in t.split()]


# PIs 0, 1; PO 2; gates 3, 4. 0->3.1, 1->4.1, 4->3.2, 3->2.1 (reconvergence, no cycle)
RECONV = b"2 1 2 4\n0 0 1 3 1\n1 0 1 4 1\n4 0 1 3 2\n3 0 1 2 1\n1 0\n"
# 0->3.1, 3->4.1 and 3->2.1, 4->3.2: the cycle 3 -> 4 -> 3
CYCLE = b"1 1 2 3\n0 0 1 2 1\n2 0 2 3 1 1 1\n3 0 1 2 2\n1 0\n"


def test_bfs_drops_wire_into_taken_gate():
    # BFS takes 0, 1, 3, 4, 2: when 4 is taken, 3 already is, so 4->3 goes.
    out, removed = lsim_amd.break_loops(RECONV, "bfs")
    assert removed == 1
    assert _norm(out) == _norm(b"2 1 2 4\n0 0 1 3 1\n1 0 1 4 1\n4 0 0\n3 0 1 2 1\n1 0\n")


def test_dfs_takes_from_the_back():
    # DFS takes 1, 4, 3, 2, 0: when 0 is taken, 3 already is, so 0->3 goes.
    out, removed = lsim_amd.break_loops(RECONV, "dfs")
    assert removed == 1
    assert _norm(out) == _norm(b"2 1 2 4\n0 0 0\n1 0 1 4 1\n4 0 1 3 2\n3 0 1 2 1\n1 0\n")


def test_cycle_is_cut():
    # ids: PI 0, PO 1, gates 2, 3; 0->2.1, 2->3.1, 2->1.1, 3->2.2
    out, removed = lsim_amd.break_loops(CYCLE, "bfs")
    assert removed == 1
    assert _norm(out) == _norm(b"1 1 2 3\n0 0 1 2 1\n2 0 2 3 1 1 1\n3 0 0\n1 0\n")
    # the cyclic netlist is rejected by the reader, the cut one is read
    s = lsim_amd.Simulator(device=-1)
    with pytest.raises(lsim_amd.LsimError):
        s.read_text(CYCLE, b"0\n", b"0\n")
    s.read_text(out, b"0\n", b"0\n")
    # levels of pi(0), gate(0), gate(1), po(0): 0, 1, 2 and the POs' level 3
    assert list(s.levels()) == [0, 1, 2, 3]


def test_malformed_text_is_rejected():
    with pytest.raises(lsim_amd.LsimError):
        lsim_amd.break_loops(b"2 1 2 4\n0 0 1 3", "bfs")
    with pytest.raises(lsim_amd.LsimError):
        lsim_amd.break_loops(b"2 1 2 1\n0 0 1 99 1\n1 0\n", "dfs")


def _random_ckt(rng: random.Random, P: int, O: int, G: int, back: float) -> bytes:
    """Random netlist with forward wires and, with probability `back`, wires
    from a gate back to an earlier gate (cycles)."""
    N = P + O + G
    pins = {g: 0 for g in range(P + O, N)}
    wires = {u: [] for u in range(N)}
    for g in range(P + O, N):
        k = rng.randint(1, 2)
        for _ in range(k):
            lo = [x for x in range(P)] + [x for x in range(P + O, g)]
            src = rng.choice(lo)
            pins[g] += 1
            wires[src].append((g, pins[g]))
        if g > P + O + 1 and rng.random() < back and pins[g] < 4:
            dst = rng.randrange(P + O, g)
            if pins[dst] < 4:
                pins[dst] += 1
                wires[g].append((dst, pins[dst]))
    for o in range(P, P + O):
        wires[rng.randrange(P + O, N)].append((o, 1))
    recs = [(u, w) for u, w in wires.items() if w]
    rng.shuffle(recs)
    lines = [f"{P} {O} {G} {len(recs)}"]
    for u, w in recs:
        lines.append(" ".join([str(u), "0", str(len(w))] + [f"{d} {p}" for d, p in w]))
    lines.append(" ".join(str(rng.choice([0, 1, 2, 3, 4, 5, 6, 8])) for _ in range(G)))
    return ("\n".join(lines) + "\n").encode()


@pytest.mark.parametrize("mode", ["bfs", "dfs"])
def test_matches_restatement_and_runs_forward(mode):
    rng = random.Random(163)
    for case in range(40):
        ckt = _random_ckt(rng, rng.randint(1, 6), rng.randint(1, 4), rng.randint(2, 60), back=0.3)
        got, removed = lsim_amd.break_loops(ckt, mode)
        want, want_removed, order = ref.break_loops(ckt, mode)
        assert _norm(got) == _norm(want) and removed == want_removed, case
        # every kept wire between taken nodes runs from an earlier-taken node to a later one
        pos = {u: i for i, u in enumerate(order)}
        _, _, _, recs, _ = ref.parse_ckt(got)
        firsts = {}
        for r in recs:
            firsts.setdefault(r[0], r)
        for u, r in firsts.items():
            if u in pos:
                for d, _ in r[2]:
                    assert d in pos and pos[d] > pos[u], (case, u, d)


def test_generated_acyclic_circuits_keep_their_nodes():
    """On the generator's (acyclic, levelized) circuits the traversal takes
    every node, and the result has no more wires than the input."""
    for seed in range(4):
        ckt, _, _ = lsim_amd.generate(seed=seed, num_pis=8, num_pos=4, num_gates=200, num_levels=10,
                                      num_patterns=64)
        for mode in ("bfs", "dfs"):
            out, removed = lsim_amd.break_loops(ckt, mode)
            _, _, order = ref.break_loops(ckt, mode)
            assert len(order) == 8 + 4 + 200
            assert len(_norm(out)) == len(_norm(ckt)) - 2 * removed
