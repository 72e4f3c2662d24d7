"""Host-side checks that need no GPU: the library loads and exports its ABI,
the reader/levelizer reproduces the reference's structural results, malformed
input is rejected, the generator follows its specification, and the
stem/region structure used by the kernels satisfies its invariants."""
import ctypes
import os

import numpy as np
import pytest

import lsim_amd
import oracle
from lsim_amd import workloads


def test_library_exports_every_declared_symbol():
    L = lsim_amd.lib()
    names = lsim_amd.exported_symbols()
    assert len(names) >= 28
    for n in names:
        assert hasattr(L, n), n


def test_run_without_device_fails_loudly():
    s = lsim_amd.Simulator(device=-1)
    s.read_text("1 1 1 2 0 0 1 2 1 2 0 1 1 1 0", "2 1 1 0 0", "2 1")
    with pytest.raises(lsim_amd.LsimError) as e:
        s.run(0)
    assert e.value.code == lsim_amd.LSIM_E_DEVICE


def _host(ckt, flst, ptn):
    s = lsim_amd.Simulator(device=-1)
    s.read_text(ckt, flst, ptn)
    return s


def test_unit_vector_structure(unit_vectors):
    """unittests/test_basics.cpp:460-526 on the product reader."""
    for c in unit_vectors:
        s = _host(c["ckt"], c["flst"], c["ptn"])
        P, O, G, W, R = c["num_vars"]
        assert (s.num_PIs(), s.num_POs(), s.num_gates(), s.num_wires(), s.num_rounds()) == (P, O, G, W, R)
        assert list(s.gate_idx_to_tpg_order_idx()) == c["tpg"]
        assert list(s.num_gates_per_level()) == c["per_level"]
        assert list(s.levels()) == c["pi_lvl"] + c["gate_lvl"] + c["po_lvl"]
        site, val = s.faults()
        for i in range(len(site)):
            assert site[i] == c["tpg"][c["faults"][2 * i]]
            assert val[i] == (2**64 - 1 if c["faults"][2 * i + 1] else 0)
        assert list(s.patterns()[0]) == c["ptn_val"]


def test_reference_made_structure(ref_vectors):
    for c in ref_vectors:
        s = _host(c["ckt"], c["flst"], c["ptn"])
        assert list(s.gate_idx_to_tpg_order_idx()) == c["order_map"], c["name"]
        assert list(s.levels()) == c["levels"], c["name"]
        assert list(s.num_gates_per_level()) == c["gates_per_level"], c["name"]


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_generated_structure_matches_oracle(name):
    ckt, flst, ptn = workloads.texts(name)
    s = _host(ckt, flst, ptn)
    o = oracle.Oracle(ckt, flst, ptn)
    assert np.array_equal(s.gate_idx_to_tpg_order_idx(), o.order_map())
    assert np.array_equal(s.levels(), o.levels())
    assert np.array_equal(s.num_gates_per_level(), o.gates_per_level())
    assert np.array_equal(s.patterns(), o.patterns())
    a, b = s.faults(), o.faults()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_generator_follows_spec():
    p = workloads.describe("c1")
    ckt, flst, ptn = workloads.texts("c1")
    assert (ckt, flst, ptn) == workloads.texts("c1")  # deterministic
    s = _host(ckt, flst, ptn)
    assert (s.num_PIs(), s.num_POs(), s.num_gates()) == (64, 32, 5000)
    assert s.num_faults() == 10000 and s.num_rounds() == 16 and s.num_patterns() == 1024
    per = s.num_gates_per_level()
    # PIs at level 0, gates evenly over 40 levels, POs one level above
    assert len(per) == p["num_levels"] + 2 and per[0] == 64 and per[-1] == 32
    assert set(per[1:-1]) == {5000 // 40}
    toks = ckt.split()
    types = [int(t) for t in toks[-5000:]]
    assert set(types) <= {0, 1, 2, 3, 4, 5, 6, 8}
    unary = sum(t in (0, 8) for t in types) / 5000
    assert 0.07 < unary < 0.13
    # fault list: (id, 0), (id, 1) for every inner gate, ascending
    f = [int(x) for x in flst.split()]
    assert f[0] == 10000 and f[1:5] == [96, 0, 96, 1] and f[-2:] == [5095, 1]
    # partial last round keeps only the live low bits
    _, _, ptn2 = lsim_amd.generate(seed=1, num_pis=4, num_pos=1, num_gates=10, num_levels=2, num_patterns=70)
    words = [int(x) for x in ptn2.split()[1:]]
    assert len(words) == 8 and all(w < 2**6 for w in words[4:])
    with pytest.raises(lsim_amd.LsimError):
        lsim_amd.generate(seed=1, num_pis=2, num_pos=1, num_gates=10, num_levels=2, num_patterns=200)


@pytest.mark.parametrize(
    "ckt,flst,ptn,what",
    [
        ("1 1 1 2 0 0 1 1 1 2 0 1 1 1 0", "0", "0", "PI drives a PO"),
        ("1 1 1 2 0 0 1 2 1 2 0 1 1 1 9", "0", "0", "gate type"),
        ("1 1 2 2 0 0 1 2 1 2 0 1 1 1 0 0", "0", "0", "unreachable gate"),
        ("1 1 1 2 0 0 1 2 6 2 0 1 1 1 0", "0", "0", "pin out of range"),
        ("1 1 1 2 0 0 1 2 1 2 0 1 1 2 0", "0", "0", "PO reads an unfilled pin"),
        ("1 1 1 2 0 0 1 2 1 2 0 1 1 1 0", "1 7 0", "0", "fault id out of range"),
        ("1 1 1 2 0 0 1 2 1 2 0 1 1 1 0", "0", "65 1 2", "more rounds than PIs"),
        ("1 1 1 2 0 0 1 2 1 2 0 1 1 1 0", "0", "2", "truncated pattern words"),
        ("1 1 1 2 0 0 1 2 1 2 0 1 1 x 0", "0", "0", "non-numeric"),
        ("1 1 1 3 0 0 1 2 1 2 0 1 1 1 1 0 1 2 2", "0", "0", "PO drives a node"),
        ("1 1 1 2 0 0 1 2 1 2 0 1 1 1 7", "0", "0", "MUX with one input"),
    ],
)
def test_rejects_malformed_input(ckt, flst, ptn, what):
    s = lsim_amd.Simulator(device=-1)
    with pytest.raises(lsim_amd.LsimError) as e:
        s.read_text(ckt, flst, ptn)
    assert e.value.code == lsim_amd.LSIM_E_INPUT, what


@pytest.mark.parametrize(
    "ckt,good",
    [
        # PI 1 drives pin 1 of AND gate 3 twice: the later wire overwrites, pin 2 reads node 0
        ("2 1 1 3 1 0 1 3 1 1 0 1 3 1 3 0 1 2 1 1", 10 & 12),
        # PI 1 drives pin 2 only: pin 1 reads node 0 (fsim/simulation.cpp:678-680,856-861)
        ("2 1 1 2 1 0 1 3 2 3 0 1 2 1 1", 12),
        # pins 1 and 3 of a 2-wire gate: pin 2 reads node 0, pin 3 is not read
        ("2 1 1 3 1 0 1 3 1 0 0 1 3 3 3 0 1 2 1 1", 10 & 12),
    ],
)
def test_unfilled_pins_read_node_zero(ckt, good):
    """The reference's pin table is zero-filled: a node with k wires reads pins
    1..k and an unfilled slot reads node id 0 (PI 0 = 12, PI 1 = 10 here)."""
    ptn = "64\n12 10\n"
    s = lsim_amd.Simulator(device=-1)
    s.read_text(ckt, "0", ptn)
    o = oracle.Oracle(ckt, "0", ptn)
    assert list(s.levels()) == list(o.levels())
    assert list(s.gate_idx_to_tpg_order_idx()) == list(o.order_map())
    assert int(o.run_literal(threads=1)["g_gate"][0]) == good


def test_reference_file_errors(tmp_path):
    s = lsim_amd.Simulator(device=-1)
    missing = str(tmp_path / "nope.ckt")
    with pytest.raises(lsim_amd.LsimError) as e:
        s.read(missing, missing, missing)
    assert e.value.code == lsim_amd.LSIM_E_IO
    assert e.value.message == "cannot open circut file " + missing  # fsim/simulation.cpp:25
    c = tmp_path / "a.ckt"
    f = tmp_path / "a.flst"
    p = tmp_path / "a.ptn"
    c.write_text("1 1 1 2 0 0 1 2 1 2 0 1 1 1 0")
    f.write_text("2 1 1 0 0")
    p.write_text("2 1")
    s.read(str(c), str(f), str(p))
    assert list(s.gate_idx_to_tpg_order_idx()) == [0, 2, 1]


def test_accessor_size_checks():
    s = _host("1 1 1 2 0 0 1 2 1 2 0 1 1 1 0", "2 1 1 0 0", "2 1")
    L = lsim_amd.lib()
    buf = (ctypes.c_uint64 * 2)()
    rc = L.lsim_gate_idx_to_tpg_order_idx(s._h, buf, 2)
    assert rc == lsim_amd.LSIM_E_ARG
    with pytest.raises(lsim_amd.LsimError) as e:
        s.found_fault_to_pattern()
    assert e.value.code == lsim_amd.LSIM_E_STATE


@pytest.mark.parametrize("key,good,bad", [
    ("first_words", [1, 64], [0, 65]),
    ("word_growth", [1, 64], [0, 65]),
    ("max_step_words", [1, 128, 4096], [0, 4097]),
    ("pair_nopend", [0, 1, 2], [-1, 3]),
    ("pair_dense", [0, 24, 64], [-1, 65]),
    ("pair_fo_pre", [0, 1], [-1, 2]),
    ("pair_nopend_pct", [0, 45, 100], [-1, 101]),
    ("pair_threads", [0, 256, 512], [128, 1024]),
    ("win_levels", [1, 4, 8], [0, 9]),
    ("win_cap", [0, 96, 4096], [-1, 4097]),
])
def test_option_ranges(key, good, bad):
    """Options are checked on the host: values in range are taken, others
    fail with LSIM_E_ARG and leave the handle usable."""
    s = lsim_amd.Simulator(device=-1)
    for v in good:
        s.set_option(key, v)
    for v in bad:
        with pytest.raises(lsim_amd.LsimError) as e:
            s.set_option(key, v)
        assert e.value.code == lsim_amd.LSIM_E_ARG
    s.set_option(key, good[0])
    with pytest.raises(lsim_amd.LsimError) as e:
        s.set_option("no_such_option", 1)
    assert e.value.code == lsim_amd.LSIM_E_ARG


def test_stem_regions_on_generated_circuit():
    """Every live node's fanout-free path ends at a PO or at a node with two or
    more live fanouts; dead nodes reach no PO. Checked by brute force."""
    ckt, flst, ptn = lsim_amd.generate(seed=21, num_pis=12, num_pos=6, num_gates=400, num_levels=15,
                                       num_patterns=64, fanout_cap=20)
    o = oracle.Oracle(ckt, flst, ptn)
    N, P, G, O = o.N, o.P, o.G, o.O
    toks = [int(x) for x in ckt.split()]
    W = toks[3]
    pos = o.order_map()
    fo = [set() for _ in range(N)]
    i = 4
    for _ in range(W):
        src, _pin, k = toks[i:i + 3]
        i += 3
        for _ in range(k):
            fo[int(pos[src])].add(int(pos[toks[i]]))
            i += 2
    reach = [False] * N
    for p in range(N - 1, -1, -1):
        reach[p] = p >= P + G or any(reach[q] for q in fo[p])
    live = sum(reach)
    assert 0 < live < N
    # the product library agrees on which faults can be detected at all: a
    # fault on a dead node can never differ at a PO
    for p in range(P, P + G):
        if not reach[p]:
            assert all(not reach[q] for q in fo[p])


def test_trace_flag_off_values(monkeypatch):
    """LSIM_TRACE=0 or empty leaves tracing off (the library reads it once per process)."""
    import subprocess, sys
    code = "import lsim_amd; print(lsim_amd.lib() is not None)"
    for v in ("0", ""):
        env = dict(__import__("os").environ, LSIM_TRACE=v)
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0 and "[lsim]" not in r.stderr
