"""Writes tests/golden/ref_vectors.json from the reference simulator.

Runs oracle/_ref/fsim_ref (the reference fsim::Simulator compiled from its own
sources by `make -C oracle ref`) on small seeded circuits that exercise what
the reference's own unit tests leave open: the MUX word-level `!s`, 3- and
4-input folds, several rounds, a partial last round whose words carry set bits
above the pattern count, PI and PO fault sites, duplicate fanins, shuffled
wire order (Kahn tie order), zero patterns, an empty fault list, and gates
whose pins are not dense 1..k (an unfilled slot reads node id 0, a pin wired
twice keeps the later wire; the levels then need not ascend in topological
order).

Per-round detection d(f, r) is read from the reference by running it once per
round on a pattern file holding only that round: its report then holds
d(f, r) (fsim/simulation.cpp:142-144). The full run gives the literal table,
the good values of the last round, the order map and the levels.

Usage: python tests/golden/make_ref_vectors.py   (needs oracle/_ref/fsim_ref)
"""
from __future__ import annotations

import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

INV, AND, OR, XOR, NAND, NOR, XNOR, MUX, CLKBUF = range(9)


def rand_circuit(rng: random.Random, P: int, O: int, G: int, mux: float, wide: float, dup: float,
                 gap: float = 0.0) -> str:
    """Random levelized netlist in .ckt text with shuffled ids and wire order.
    With `gap` > 0 a gate's pins are, with that probability, a non-dense set
    (a pin slot left unfilled, which the reference reads as node id 0, and pins
    above the wire count, which it does not read) or repeat a pin (the later
    wire overwrites the earlier one) (fsim/simulation.cpp:676-711,851-877)."""
    N = P + O + G
    gate_ids = list(range(P + O, N))
    rng.shuffle(gate_ids)  # creation order != id order
    avail = list(range(P))  # nodes that may drive a gate (PIs and created gates)
    fanins = {}
    types = {}
    for gid in gate_ids:
        u = rng.random()
        if u < mux:
            t, k = MUX, 3
        elif u < mux + 0.12:
            t, k = rng.choice([INV, CLKBUF]), 1
        else:
            t = rng.choice([AND, OR, XOR, NAND, NOR, XNOR])
            k = rng.choice([3, 4]) if rng.random() < wide else 2
        srcs = []
        for _ in range(k):
            if srcs and rng.random() < dup:
                srcs.append(rng.choice(srcs))
            else:
                # mostly recent nodes, sometimes anything earlier
                srcs.append(avail[-1 - rng.randrange(min(len(avail), 6))] if rng.random() < 0.7 else rng.choice(avail))
        pins = [1, 2, 5] if t == MUX else list(range(1, k + 1))
        if gap and rng.random() < gap:
            if rng.random() < 0.25:
                pins[rng.randrange(k)] = pins[rng.randrange(k)]  # a pin wired twice
            else:
                pins = sorted(rng.sample([1, 2, 3, 4], k)) if t != MUX else rng.choice([[2, 4, 5], [1, 4, 5], [2, 1, 4]])
        fanins[gid] = list(zip(srcs, pins))
        types[gid] = t
        avail.append(gid)
    drivers = rng.sample(gate_ids[len(gate_ids) // 2:], O)
    for i, d in enumerate(drivers):
        fanins[P + i] = [(d, 1)]
    # wire records: one or two records per source, shuffled, pairs shuffled
    by_src = {}
    for dst, lst in fanins.items():
        for src, pin in lst:
            by_src.setdefault(src, []).append((dst, pin))
    records = []
    for src, outs in by_src.items():
        rng.shuffle(outs)
        if len(outs) > 1 and rng.random() < 0.3:
            cut = rng.randrange(1, len(outs))
            records.append((src, outs[:cut]))
            records.append((src, outs[cut:]))
        else:
            records.append((src, outs))
    rng.shuffle(records)
    lines = [f"{P} {O} {G} {len(records)}"]
    for src, outs in records:
        lines.append(f"{src} 0 {len(outs)} " + " ".join(f"{d} {p}" for d, p in outs))
    lines.append(" ".join(str(types[g]) for g in range(P + O, N)))
    return "\n".join(lines) + "\n"


def rand_faults(rng: random.Random, N: int, F: int) -> str:
    pairs = [(rng.randrange(N), rng.randrange(2)) for _ in range(F)]
    return f"{F}\n" + "".join(f"{i} {v}\n" for i, v in pairs)


def rand_patterns(rng: random.Random, P: int, n: int, garbage: bool) -> list:
    R = (n + 63) // 64
    rounds = []
    for r in range(R):
        words = [rng.getrandbits(64) for _ in range(P)]
        if r == R - 1 and n % 64 and not garbage:
            words = [w & ((1 << (n % 64)) - 1) for w in words]
        rounds.append(words)
    return rounds


def ptn_text(n: int, rounds: list) -> str:
    return f"{n}\n" + "".join(" ".join(str(w) for w in ws) + "\n" for ws in rounds)


def make_case(name: str, seed: int, P: int, O: int, G: int, F: int, n: int, mux=0.0, wide=0.0, dup=0.0,
              garbage=False, gap=0.0) -> dict:
    rng = random.Random(seed)
    ckt = rand_circuit(rng, P, O, G, mux, wide, dup, gap)
    flst = rand_faults(rng, P + O + G, F)
    rounds = rand_patterns(rng, P, n, garbage)
    ptn = ptn_text(n, rounds)
    full, _ = oracle.run_reference(ckt, flst, ptn, mode=0, threads=1)
    R = len(rounds)
    det = []
    for r in range(R):
        nr = 64 if (r < R - 1 or n % 64 == 0) else n % 64
        one, _ = oracle.run_reference(ckt, flst, ptn_text(nr, [rounds[r]]), mode=0, threads=1)
        det.append([int(x) for x in one["table"][0::2]])
    det_fr = [[det[r][f] for r in range(R)] for f in range(F)]
    return {
        "name": name, "ckt": ckt, "flst": flst, "ptn": ptn,
        "order_map": [int(x) for x in full["order_map"]],
        "gates_per_level": [int(x) for x in full["gates_per_level"]],
        "levels": [int(x) for x in full["levels"]],
        "table": [int(x) for x in full["table"]],
        "g_pi": [int(x) for x in full["g_pi"]], "g_gate": [int(x) for x in full["g_gate"]],
        "g_po": [int(x) for x in full["g_po"]],
        "det": det_fr,
    }


CASES = [
    # name, seed, P, O, G, F, n, mux, wide, dup, garbage
    ("two_input_multi_round", 11, 8, 3, 60, 40, 4 * 64, 0.0, 0.0, 0.0, False),
    ("partial_last_round", 12, 6, 2, 50, 40, 2 * 64 + 37, 0.0, 0.0, 0.0, False),
    ("partial_round_high_bits", 13, 6, 2, 50, 40, 64 + 5, 0.0, 0.0, 0.0, True),
    ("wide_gates", 14, 10, 4, 80, 60, 3 * 64, 0.0, 0.5, 0.0, False),
    ("mux_quirk", 15, 8, 3, 70, 60, 2 * 64 + 9, 0.25, 0.2, 0.0, True),
    ("mux_dense_single_round", 16, 5, 2, 40, 50, 17, 0.5, 0.0, 0.0, True),
    ("duplicate_fanins", 17, 6, 3, 60, 50, 3 * 64, 0.0, 0.3, 0.35, False),
    ("pi_po_faults_many_rounds", 18, 12, 5, 90, 80, 12 * 64, 0.05, 0.2, 0.1, False),
    ("one_round_exact", 19, 4, 1, 20, 30, 64, 0.0, 0.0, 0.0, False),
    ("single_pattern", 20, 4, 2, 30, 30, 1, 0.1, 0.1, 0.0, True),
    ("empty_fault_list", 21, 4, 2, 30, 0, 128, 0.0, 0.0, 0.0, False),
    ("zero_patterns", 22, 4, 2, 30, 10, 0, 0.0, 0.0, 0.0, False),
    ("larger_mixed", 23, 16, 8, 400, 300, 5 * 64 + 50, 0.05, 0.2, 0.05, True),
    # name, seed, P, O, G, F, n, mux, wide, dup, garbage, gap
    ("nondense_pins", 24, 8, 3, 80, 80, 3 * 64 + 11, 0.0, 0.3, 0.0, False, 0.35),
    ("nondense_pins_mux", 25, 8, 3, 90, 90, 2 * 64, 0.2, 0.3, 0.05, True, 0.3),
]


def main() -> None:
    if not oracle.ref_available():
        sys.exit("oracle/_ref/fsim_ref is missing: run `make -C oracle ref`")
    cases = [make_case(*c) for c in CASES]
    doc = {
        "source": "reference fsim::Simulator compiled from fsim/simulation.cpp (oracle/_ref/fsim_ref), "
                  "mode SEQUENTIAL; written by tests/golden/make_ref_vectors.py",
        "note": "det[f][r] = d(f, r) from one reference run per round; table/g_* from the full run; "
                "levels = pi(i), gate(i), po(i) levels in topological order.",
        "cases": cases,
    }
    out = os.path.join(HERE, "ref_vectors.json")
    with open(out, "w") as fh:
        json.dump(doc, fh, separators=(",", ":"))
    print("wrote", out, len(cases), "cases")


if __name__ == "__main__":
    main()
