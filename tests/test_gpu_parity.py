"""GPU parity: the HIP engine against the reference's vectors and the CPU oracle.

Bit-exact on every detection bit d(f, r), on the reference's literal report
found_fault_to_pattern (fsim/simulation.cpp:142-144) and on the good-machine
values of the last round (fsim/simulation.cpp:163-178).
"""
import numpy as np
import pytest

import lsim_amd
import oracle
from lsim_amd import workloads

pytestmark = pytest.mark.gpu

# Engines: the per-level form, and the persistent event form with each of its
# two sweep kernels (flattened (node, lane) pairs; one node per wave). The pair
# sweep evaluates dense level passes one node per wave: by default where the
# nodes average 24 active lanes, and forced never (pairs_flat) or always
# (pairs_dense). Dense steps walk every node of a level instead of pending
# flags: by default, and forced never (pairs_pend) or in every step
# (pairs_nopend). With pending flags a staged node's first fanout is loaded
# while it is staged: by default, and not (pairs_pend_nopre). Word steps are
# planned on the device by default, and on the host (pairs_host).
MODES = [(lsim_amd.Simulator.Mode.PLAIN, {}), (lsim_amd.Simulator.Mode.PARTITIONED, {}),
         (lsim_amd.Simulator.Mode.PARTITIONED, {"sweep_kernel": 1}),
         (lsim_amd.Simulator.Mode.PARTITIONED, {"pair_dense": 0}),
         (lsim_amd.Simulator.Mode.PARTITIONED, {"pair_dense": 1}),
         (lsim_amd.Simulator.Mode.PARTITIONED, {"pair_nopend": 0}),
         (lsim_amd.Simulator.Mode.PARTITIONED, {"pair_nopend": 2}),
         (lsim_amd.Simulator.Mode.PARTITIONED, {"pair_nopend": 0, "pair_fo_pre": 0}),
         (lsim_amd.Simulator.Mode.PARTITIONED, {"device_steps": 0})]
MODE_IDS = ["plain", "pairs", "nodewave", "pairs_flat", "pairs_dense", "pairs_pend", "pairs_nopend",
            "pairs_pend_nopre", "pairs_host"]


def _sim(ckt, flst, ptn, mode=None, **opts):
    s = lsim_amd.Simulator(device=0)
    for k, v in dict(mode[1] if mode else {}, **opts).items():
        s.set_option(k, v)
    s.read_text(ckt, flst, ptn)
    return s


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_reference_unit_vectors(unit_vectors, mode):
    """unittests/test_basics.cpp:528-605: good values and fault table."""
    for c in unit_vectors:
        s = _sim(c["ckt"], c["flst"], c["ptn"], mode)
        s.run(mode[0], 4)
        pi, gate, po = s.good_results()
        assert list(pi) == c["pi"], c["name"]
        assert list(gate) == c["gate"], c["name"]
        assert list(po) == c["po"], c["name"]
        assert list(s.found_fault_to_pattern()) == c["table"], c["name"]


@pytest.mark.parametrize("direct", [0, 1], ids=["stem_tracing", "direct"])
@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_reference_made_vectors(ref_vectors, mode, direct):
    """Per-round detection and report against vectors written by the reference."""
    for c in ref_vectors:
        s = _sim(c["ckt"], c["flst"], c["ptn"], mode, direct=direct)
        s.run(mode[0])
        F, R = s.num_faults(), s.num_rounds()
        det = np.array(c["det"], dtype=np.uint8).reshape(F, R)
        assert np.array_equal(s.detect_matrix(), det), c["name"]
        assert list(s.found_fault_to_pattern()) == c["table"], c["name"]
        pi, gate, po = s.good_results()
        assert list(gate) == c["g_gate"] and list(po) == c["g_po"] and list(pi) == c["g_pi"], c["name"]
        first = s.first_detect_round()
        want = np.where(det.any(1), det.argmax(1), -1) if R else -np.ones(F, dtype=np.int64)
        assert np.array_equal(first, want), c["name"]


@pytest.fixture(scope="module")
def c1():
    ckt, flst, ptn = workloads.texts("c1")
    ref = oracle.Oracle(ckt, flst, ptn).run_literal(threads=0)
    return ckt, flst, ptn, ref


@pytest.mark.parametrize("direct", [0, 1], ids=["stem_tracing", "direct"])
@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_c1_full_against_oracle(c1, mode, direct):
    """5k gates, 10k faults, 16 rounds: every d(f, r) against the literal restatement."""
    ckt, flst, ptn, ref = c1
    s = _sim(ckt, flst, ptn, mode, direct=direct)
    s.run(mode[0])
    assert np.array_equal(s.detect_matrix(), ref["det"])
    assert np.array_equal(s.found_fault_to_pattern(), ref["table"])
    pi, gate, po = s.good_results()
    assert np.array_equal(gate, ref["g_gate"]) and np.array_equal(po, ref["g_po"])


def _gate_forms_texts():
    """Every gate type of at most two inputs with one and with two wires:
    AND/OR/XOR and their inversions with a single input (the fold of one
    word), INV/CLKBUF with a second, unread pin. Two layers over four PIs,
    stems that reach several POs, SA0/SA1 on every node, three rounds with a
    partial last one."""
    P = 4
    types1 = [0, 1, 2, 3, 4, 5, 6, 8]
    layer1 = [(t, k) for t in types1 for k in (1, 2)]
    layer2 = [(t, 2) for t in (1, 2, 3, 4, 5, 6)] + [(t, 1) for t in types1]
    taps = [0, 5, 10, 15]  # layer-1 gates that also drive a PO
    O = len(layer2) + len(taps)
    G = len(layer1) + len(layer2)
    g1 = [P + O + j for j in range(len(layer1))]
    g2 = [P + O + len(layer1) + m for m in range(len(layer2))]
    wires = []  # (src, dst, pin)
    for j, (t, k) in enumerate(layer1):
        for pin, src in enumerate([j % P, (j + 1) % P][:k], 1):
            wires.append((src, g1[j], pin))
    for m, (t, k) in enumerate(layer2):
        for pin, src in enumerate([g1[m % len(g1)], g1[(5 * m + 3) % len(g1)]][:k], 1):
            wires.append((src, g2[m], pin))
    for i in range(len(layer2)):
        wires.append((g2[i], P + i, 1))
    for i, j in enumerate(taps):
        wires.append((g1[j], P + len(layer2) + i, 1))
    recs = " ".join(f"{a} 0 1 {b} {pin}" for a, b, pin in wires)
    ckt = f"{P} {O} {G} {len(wires)} {recs} " + " ".join(str(t) for t, _ in layer1 + layer2)
    N = P + O + G
    flst = f"{2 * N} " + " ".join(f"{v} {sa}" for v in range(N) for sa in (0, 1))
    rng = np.random.default_rng(7)
    words = rng.integers(0, 2**63, size=(3, P), dtype=np.int64)
    ptn = f"{3 * 64 - 5}\n" + "\n".join(" ".join(str(int(w)) for w in row) for row in words) + "\n"
    return ckt.encode(), flst.encode(), ptn.encode()


@pytest.mark.parametrize("direct", [0, 1], ids=["stem_tracing", "direct"])
@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_gate_forms_against_oracle(mode, direct):
    """One- and two-wire forms of every gate type (the sweep evaluates gates of
    at most two inputs from 4-bit truth tables): every d(f, r), the literal
    report and the last round's good values against the restatement."""
    ckt, flst, ptn = _gate_forms_texts()
    ref = oracle.Oracle(ckt, flst, ptn).run_literal(threads=1)
    s = _sim(ckt, flst, ptn, mode, direct=direct)
    s.run(mode[0])
    assert np.array_equal(s.detect_matrix(), ref["det"])
    assert np.array_equal(s.found_fault_to_pattern(), ref["table"])
    pi, gate, po = s.good_results()
    assert np.array_equal(gate, ref["g_gate"]) and np.array_equal(po, ref["g_po"])
    assert ref["det"].any(), "no fault detected: the circuit does not exercise the sweep"


def test_reference_modes_run_persistent_engine(c1):
    """The reference's three modes (fsim/simulation.hpp:45) all run the
    persistent pair sweep and give identical, oracle-exact results; the
    per-level form runs only when asked for by its own mode value."""
    ckt, flst, ptn, ref = c1
    M = lsim_amd.Simulator.Mode
    for mode in (M.SEQUENTIAL, M.PARALLEL_TF, M.PARALLEL_OMP, M.PARTITIONED, M.PLAIN):
        s = _sim(ckt, flst, ptn)
        s.run(mode, 8)
        assert s.last_run_stats()["engine"] == (0 if mode == M.PLAIN else 1), mode
        assert np.array_equal(s.detect_matrix(), ref["det"]), mode
        assert np.array_equal(s.found_fault_to_pattern(), ref["table"]), mode
        s.close()
    s = _sim(ckt, flst, ptn)
    with pytest.raises(lsim_amd.LsimError):
        s.run(3)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_c1_fault_dropping(c1, mode):
    """With dropping, the first detecting round and the detected set equal the full matrix's."""
    ckt, flst, ptn, ref = c1
    s = _sim(ckt, flst, ptn, mode, drop=1, report_last_round=1)
    s.run(mode[0])
    det = ref["det"]
    want = np.where(det.any(1), det.argmax(1), -1)
    assert np.array_equal(s.first_detect_round(), want)
    assert np.array_equal(s.detected(), det.any(1).astype(np.uint8))
    assert np.array_equal(s.found_fault_to_pattern(), ref["table"])
    with pytest.raises(lsim_amd.LsimError):
        s.detect_matrix()


@pytest.fixture(scope="module")
def c2():
    ckt, flst, ptn = workloads.texts("c2")
    return ckt, flst, ptn


def test_c2_sampled_against_oracle(c2):
    """100k gates, 200k faults, 128 rounds: all 128 rounds of a seeded fault
    sample against the event-driven restatement, and the engines against each
    other on every fault."""
    ckt, flst, ptn = c2
    runs = {}
    for name, mode, direct in (("plain", MODES[0], 0), ("pairs", MODES[1], 0), ("nodewave", MODES[2], 0),
                               ("direct", MODES[1], 1)):
        s = _sim(ckt, flst, ptn, mode, direct=direct)
        s.run(mode[0])
        runs[name] = s.detect_matrix()
        if name == "plain":
            table = s.found_fault_to_pattern()
            good = s.good_results()
        s.close()
    assert np.array_equal(runs["plain"], runs["pairs"])
    assert np.array_equal(runs["plain"], runs["nodewave"])
    assert np.array_equal(runs["plain"], runs["direct"])
    # With dropping, in steps of 32 then 96 words (a stem's words of one step
    # span two tiles), in the default steps, and with pending marks in every
    # step on 1 % of the sweep slots (each slot sweeps hundreds of tiles, so
    # the 255 pin-mark tags repeat and stale marks match by chance): the
    # matrix's first rounds.
    det = runs["plain"]
    want = np.where(det.any(1), det.argmax(1), -1)
    for opts in ({"first_words": 32, "word_growth": 4, "max_step_words": 128}, {},
                 {"pair_nopend": 0, "slot_pct": 1}):
        s = _sim(ckt, flst, ptn, MODES[1], drop=1, **opts)
        s.run(MODES[1][0])
        assert np.array_equal(s.first_detect_round(), want), opts
        s.close()
    o = oracle.Oracle(ckt, flst, ptn)
    rng = np.random.default_rng(2002)
    sample = np.sort(rng.choice(o.F, size=160, replace=False))
    ev = o.run_event(threads=0, faults=sample)
    assert np.array_equal(runs["plain"][sample], ev[sample])
    R = o.R
    assert np.array_equal(table[0::2][sample], ev[sample, R - 1])
    assert np.all(table[1::2] == R - 1)


def test_good_machine_graph_replays(c1):
    """The captured good-machine graph, replayed for several runs on one
    circuit with eager work in between, gives the eager launches' results."""
    ckt, flst, ptn, ref = c1
    eager = _sim(ckt, flst, ptn, good_graph=0)
    eager.run(lsim_amd.Simulator.Mode.PARTITIONED)
    want_good = eager.good_results()
    want_det = eager.detect_matrix()
    g = _sim(ckt, flst, ptn)
    for _ in range(3):
        g.run(lsim_amd.Simulator.Mode.PARTITIONED)
        got = g.good_results()
        assert all(np.array_equal(a, b) for a, b in zip(got, want_good))
        assert np.array_equal(g.detect_matrix(), want_det)
        eager.run(lsim_amd.Simulator.Mode.PLAIN)  # eager launches on another engine between replays
        assert np.array_equal(eager.detect_matrix(), want_det)
    assert np.array_equal(want_det, ref["det"])


@pytest.mark.parametrize("drop", [0, 1], ids=["matrix", "dropping"])
def test_pool_overflow_reruns(c1, drop):
    """Pools capped at 64 words per slot: most tiles outgrow theirs, are given
    up and swept again in slots that hold any tile; results are unchanged."""
    ckt, flst, ptn, ref = c1
    s = _sim(ckt, flst, ptn, drop=drop, pool_cap=64)
    s.run(lsim_amd.Simulator.Mode.PARTITIONED)
    st = s.last_run_stats()
    assert st["sweep_reruns"] > 0
    det = ref["det"]
    assert np.array_equal(s.first_detect_round(), np.where(det.any(1), det.argmax(1), -1))
    assert np.array_equal(s.found_fault_to_pattern(), ref["table"])
    if not drop:
        assert np.array_equal(s.detect_matrix(), det)


@pytest.mark.parametrize("form,levels,cap", [(0, 4, 0), (1, 4, 0), (1, 8, 0), (1, 2, 48), (1, 8, 40)],
                         ids=["per_level", "4lv_lds", "8lv_lds", "2lv_48", "8lv_40"])
def test_good_forms_against_oracle_c1(c1, form, levels, cap):
    """Both good-machine forms: one launch per level, and the partitioned
    form (LDS-sized level windows; cap 0 sizes the blocks from the device's
    LDS, small caps force windows cut short): every d(f, r), the report and
    the last-round good values against the literal restatement (5k gates, 16
    rounds)."""
    ckt, flst, ptn, ref = c1
    s = _sim(ckt, flst, ptn, good_form=form, win_levels=levels, win_cap=cap)
    s.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
    assert np.array_equal(s.detect_matrix(), ref["det"])
    assert np.array_equal(s.found_fault_to_pattern(), ref["table"])
    pi, gate, po = s.good_results()
    assert np.array_equal(pi, ref["g_pi"]) and np.array_equal(gate, ref["g_gate"]) and np.array_equal(po, ref["g_po"])


def test_window_form_against_oracle_c2(c2):
    """100k gates, 120 levels, 128 rounds: the partitioned good machine's
    last-round good values against one oracle pass over that round, its
    detections against the per-level form on every fault and round, and a
    seeded fault sample against the event-driven restatement."""
    ckt, flst, ptn = c2
    s = _sim(ckt, flst, ptn, good_form=1)
    s.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
    det = s.detect_matrix()
    good = s.good_results()
    other = _sim(ckt, flst, ptn, good_form=0)
    other.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
    assert np.array_equal(det, other.detect_matrix())
    o = oracle.Oracle(ckt, flst, ptn)
    want = o.good_round(o.R - 1)
    assert all(np.array_equal(a, b) for a, b in zip(good, want))
    sample = np.sort(np.random.default_rng(2012).choice(o.F, size=200, replace=False))
    assert np.array_equal(det[sample], o.run_event_range(sample, 0, o.R))


def test_repeated_runs_keep_inputs_resident(c1):
    """run() again on the same inputs reuses the resident fault list and
    patterns: every run, also after option changes, gives the same answer."""
    ckt, flst, ptn, ref = c1
    s = _sim(ckt, flst, ptn)
    mode = lsim_amd.Simulator.Mode.PARTITIONED
    s.run(mode)
    det, table = ref["det"], ref["table"]
    assert np.array_equal(s.detect_matrix(), det) and np.array_equal(s.found_fault_to_pattern(), table)
    first = np.where(det.any(1), det.argmax(1), -1)
    s.run(mode)
    assert np.array_equal(s.detect_matrix(), det) and np.array_equal(s.found_fault_to_pattern(), table)
    s.set_option("direct", 1)
    s.run(mode)
    assert np.array_equal(s.detect_matrix(), det)
    s.set_option("drop", 1)
    s.run(mode)
    assert np.array_equal(s.first_detect_round(), first)
    assert np.array_equal(s.found_fault_to_pattern(), table)
    s.set_option("direct", 0)
    s.run(mode)
    assert np.array_equal(s.first_detect_round(), first)
    s.set_option("drop", 0)
    s.run(lsim_amd.Simulator.Mode.PLAIN)
    assert np.array_equal(s.detect_matrix(), det)


def _with_muxes(ckt: bytes, seed: int, count: int) -> bytes:
    """The circuit with `count` 2-input gates turned into MUXes: the gate's
    pins 1 and 2 become the data inputs and a new wire from a PI drives the
    select (pin 5). PIs sit on level 0, so the order stays acyclic."""
    import random

    toks = ckt.split()
    P, O, G, W = (int(x) for x in toks[:4])
    rng = random.Random(seed)
    i, recs, pins = 4, [], {}
    for _ in range(W):
        src, py, k = int(toks[i]), int(toks[i + 1]), int(toks[i + 2])
        outs = [(int(toks[i + 3 + 2 * j]), int(toks[i + 4 + 2 * j])) for j in range(k)]
        recs.append((src, py, outs))
        for d, p in outs:
            pins.setdefault(d, []).append(p)
        i += 3 + 2 * k
    types = [int(x) for x in toks[i:i + G]]
    two = [g for g in range(G) if types[g] in (1, 2, 3, 4, 5, 6) and sorted(pins.get(P + O + g, [])) == [1, 2]]
    for g in rng.sample(two, count):
        types[g] = 7
        recs.append((rng.randrange(P), 0, [(P + O + g, 5)]))
    lines = [f"{P} {O} {G} {len(recs)}"]
    lines += [f"{s_} {py} {len(outs)} " + " ".join(f"{d} {p}" for d, p in outs) for s_, py, outs in recs]
    lines.append(" ".join(map(str, types)))
    return ("\n".join(lines) + "\n").encode()


@pytest.mark.parametrize("drop", [0, 1], ids=["matrix", "dropping"])
def test_mux_circuit_splits_faults_against_oracle(drop):
    """A 5,000-gate circuit with 40 MUXes: faults whose cone reaches a MUX
    select are injected directly, the others keep stem tracing (their
    patterns propagate independently through MUX data inputs). Every
    d(f, r) (or first round with dropping) and the report equal the literal
    restatement, and stem tracing is used."""
    ckt, flst, ptn = lsim_amd.generate(seed=4242, num_pis=64, num_pos=32, num_gates=5000, num_levels=40,
                                       num_patterns=64 * 6 + 13, fanout_cap=40)
    ckt = _with_muxes(ckt, 7, 40)
    # PI faults too: a PI that drives a select reaches it, so both kinds occur
    lines = flst.decode().split()
    pairs = lines[1:] + [x for i in range(64) for x in (str(i), str(i % 2))]
    flst = (f"{len(pairs) // 2}\n" + "\n".join(f"{pairs[2 * k]} {pairs[2 * k + 1]}" for k in range(len(pairs) // 2))
            + "\n").encode()
    ref = oracle.Oracle(ckt, flst, ptn).run_literal(threads=0)
    det = ref["det"]
    s = _sim(ckt, flst, ptn, drop=drop)
    s.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
    st = s.last_run_stats()
    assert st["stems"] > 0
    assert np.array_equal(s.found_fault_to_pattern(), ref["table"])
    if drop:
        assert np.array_equal(s.first_detect_round(), np.where(det.any(1), det.argmax(1), -1))
    else:
        assert np.array_equal(s.detect_matrix(), det)
    pi, gate, po = s.good_results()
    assert np.array_equal(gate, ref["g_gate"]) and np.array_equal(po, ref["g_po"])
    # all faults injected directly: the same answer
    s.set_option("direct", 1)
    s.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
    assert s.last_run_stats()["stems"] == 0
    assert np.array_equal(s.found_fault_to_pattern(), ref["table"])
    s.close()
