"""Parity at the 2M-gate size (C3: 2,000,000 gates, 4,000,000 faults,
32,768 patterns, fault dropping): the two persistent sweep kernels, and the
pair sweep with and without its one-node-per-wave dense passes and with and
without pending flags in its dense first step, agree on every fault, and a seeded fault sample agrees with the event-driven
restatement on the first four rounds (a fault the GPU first detects in round
g < 4 is undetected before g and detected in g; every other sampled fault is
undetected in rounds 0-3)."""
import numpy as np
import pytest

import lsim_amd
import oracle
from lsim_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3():
    return workloads.texts("c3")


def _run(ckt, flst, ptn, **opts):
    s = lsim_amd.Simulator(device=0)
    for k, v in opts.items():
        s.set_option(k, v)
    s.read_text(ckt, flst, ptn)
    s.run(lsim_amd.Simulator.Mode.PARTITIONED)
    out = s.first_detect_round(), s.found_fault_to_pattern()
    s.close()
    return out


def test_c3_kernels_agree_and_sample_matches_oracle(c3):
    ckt, flst, ptn = c3
    first, table = _run(ckt, flst, ptn, drop=1)
    first_nw, table_nw = _run(ckt, flst, ptn, drop=1, sweep_kernel=1)
    assert np.array_equal(first, first_nw)
    assert np.array_equal(table, table_nw)
    first_fl, table_fl = _run(ckt, flst, ptn, drop=1, pair_dense=0)
    assert np.array_equal(first, first_fl)
    assert np.array_equal(table, table_fl)
    first_pd, table_pd = _run(ckt, flst, ptn, drop=1, pair_nopend=0)
    assert np.array_equal(first, first_pd)
    assert np.array_equal(table, table_pd)
    F = first.size
    assert F == 4_000_000
    # table = {found in the last round, last round} for every fault
    assert np.all(table[1::2] == 511)
    assert np.all(table[0::2] <= (first >= 0))
    # the first four rounds of the pattern file
    lines = ptn.split(b"\n")
    sub_ptn = b"256\n" + b"\n".join(lines[1:5]) + b"\n"
    rng = np.random.default_rng(3003)
    early = np.flatnonzero((first >= 0) & (first < 4))
    other = np.flatnonzero((first < 0) | (first >= 4))
    sample = np.sort(np.concatenate([rng.choice(early, 48, replace=False), rng.choice(other, 16, replace=False)]))
    o = oracle.Oracle(ckt, flst, sub_ptn)
    det = o.run_event(threads=0, faults=sample)
    for f in sample:
        g = int(first[f])
        if 0 <= g < 4:
            assert not det[f, :g].any() and det[f, g] == 1, (f, g, det[f])
        else:
            assert not det[f, :4].any(), (f, g, det[f])
