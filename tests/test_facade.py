"""The C++ facade (include/lsim/simulator.hpp) used the way the reference's
own callers use fsim::Simulator: both read() overloads, the accessors, and on
a GPU run() plus found_fault_to_pattern() (unittests/test_basics.cpp:460-605)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "facade_check")


def _run(case, tmp_path, device, how, run, dot=None):
    paths = []
    for key, ext in (("ckt", "ckt"), ("flst", "flst"), ("ptn", "ptn")):
        p = tmp_path / f"case.{ext}"
        p.write_text(case[key])
        paths.append(str(p))
    r = subprocess.run([BIN, str(device)] + paths + [how, str(int(run))] + ([dot] if dot else []),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        name, *vals = line.split()
        out[name] = vals if name == "type_names" else [int(v) for v in vals]
    return out


def _check_structure(c, out):
    assert out["num_vars"] == c["num_vars"]
    assert out["tpg"] == c["tpg"]
    assert out["per_level"] == c["per_level"]
    assert out["levels"] == c["pi_lvl"] + c["gate_lvl"] + c["po_lvl"]
    want = []
    for i in range(c["fault_size"]):
        want += [c["tpg"][c["faults"][2 * i]], c["faults"][2 * i + 1]]
    assert out["faults"] == want
    assert out["ptn0"] == c["ptn_val"]
    # fsim/simulation.cpp:1126-1154
    assert out["type_names"] == ["INV", "AND", "OR", "XOR", "NAND", "NOR", "XNOR", "MUX", "CLKBUF", "PI", "PO",
                                 "UNKNOWN"]


@pytest.mark.parametrize("how", ["stream", "path"])
def test_facade_structure(unit_vectors, tmp_path, how):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_bin/facade_check is not built (lsim_amd/build.py builds it)")
    for c in unit_vectors:
        _check_structure(c, _run(c, tmp_path, -1, how, False))


def test_facade_visualization_graph(unit_vectors, tmp_path):
    """Simulator::visualization_graph through the facade writes the library's text."""
    import lsim_amd

    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_bin/facade_check is not built (lsim_amd/build.py builds it)")
    for c in unit_vectors:
        dot = tmp_path / "g.dot"
        _run(c, tmp_path, -1, "path", False, str(dot))
        s = lsim_amd.Simulator(device=-1)
        s.read_text(c["ckt"], c["flst"], c["ptn"])
        assert dot.read_text() == s.visualization_text(), c["name"]
    # an unopenable path: the message on stderr, no failure (fsim/simulation.cpp:1167-1170)
    paths = [str(tmp_path / f"case.{e}") for e in ("ckt", "flst", "ptn")]
    r = subprocess.run([BIN, "-1"] + paths + ["path", "0", str(tmp_path / "missing" / "g.dot")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Error opening output file" in r.stderr


@pytest.mark.gpu
def test_facade_run(unit_vectors, tmp_path):
    for c in unit_vectors:
        out = _run(c, tmp_path, 0, "stream", True)
        _check_structure(c, out)
        assert out["pi"] == c["pi"] and out["gate"] == c["gate"] and out["po"] == c["po"]
        assert out["table"] == c["table"]
        # the reference's SEQUENTIAL mode runs the persistent pair sweep (k_sweep_pairs)
        assert out["engine"] == [1], c["name"]
