"""Parity at the 2M-gate size (2,000,000 gates, 512 PI, 256 PO, 1,500 levels,
4,000,000 SA0/SA1 faults, 32,768 patterns = 512 rounds, fault dropping).

Against the oracle:
- 2,048 seeded faults (1,024 uniform, 512 the GPU first detects after round 0,
  512 it never detects) have the first detecting round of the event-driven
  restatement over all 512 rounds;
- those faults' literal report {found in round 511, 511} equals the oracle's
  d(f, 511) (fsim/simulation.cpp:142-144);
- all 2,000,768 last-round good values equal one good-machine pass of the
  oracle over round 511 (fsim/simulation.cpp:163-178).
Between kernels: the two persistent sweep kernels, and the pair sweep with and
without its one-node-per-wave dense passes and with and without pending flags
in its dense first step, agree on every fault.
"""
import numpy as np
import pytest

import lsim_amd
import oracle
from lsim_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    return workloads.texts("c3")


def _run(ckt, flst, ptn, good=False, **opts):
    s = lsim_amd.Simulator(device=0)
    for k, v in opts.items():
        s.set_option(k, v)
    s.read_text(ckt, flst, ptn)
    s.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
    out = [s.first_detect_round(), s.found_fault_to_pattern()]
    if good:
        out.append(s.good_results())
    s.close()
    return out


def test_c3_against_oracle_and_kernels_agree(c3):
    ckt, flst, ptn = c3
    first, table, good = _run(ckt, flst, ptn, good=True, drop=1)
    F = first.size
    assert F == 4_000_000
    assert np.all(table[1::2] == 511)
    o = oracle.Oracle(ckt, flst, ptn)
    assert o.R == 512
    # every last-round good value
    g_pi, g_gate, g_po = o.good_round(511)
    assert np.array_equal(good[0], g_pi) and np.array_equal(good[1], g_gate) and np.array_equal(good[2], g_po)
    # a seeded fault sample over all rounds
    rng = np.random.default_rng(3003)
    late = np.flatnonzero(first >= 1)
    never = np.flatnonzero(first < 0)
    sample = np.unique(np.concatenate([rng.choice(F, 1024, replace=False), rng.choice(late, 512, replace=False),
                                       rng.choice(never, 512, replace=False)]))
    want = o.first_rounds(sample, chunk=32)
    bad = np.flatnonzero(first[sample] != want)
    assert bad.size == 0, [(int(sample[i]), int(first[sample[i]]), int(want[i])) for i in bad[:10]]
    # the literal report of those faults: d(f, 511)
    last = o.run_event_range(sample, 511, 512)[:, 0]
    assert np.array_equal(table[0::2][sample], last.astype(np.uint64))
    assert np.all(table[0::2] <= (first >= 0))
    del o
    # the other kernels and schedules on every fault
    for opts in ({"sweep_kernel": 1}, {"pair_dense": 0}, {"pair_nopend": 0}):
        f2, t2 = _run(ckt, flst, ptn, drop=1, **opts)
        assert np.array_equal(first, f2), opts
        assert np.array_equal(table, t2), opts
