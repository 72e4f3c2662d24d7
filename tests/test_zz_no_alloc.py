"""lsim_run_noalloc (the entry point that allocates nothing) against lsim_run.

Two calls in a row on one handle, the second with a smaller input, each
bit-exact with lsim_run on the same input (detection matrix or first rounds,
the literal report, the last-round good values) and each reporting no
device allocation; a run too large for the fixed scratch is rejected with
LSIM_E_NOMEM before any device work; lsim_run on the same handle still works
afterwards."""
import numpy as np
import pytest

import lsim_amd
from lsim_amd import workloads

pytestmark = pytest.mark.gpu

M = lsim_amd.Simulator.Mode


def _result(s, drop):
    out = {"table": s.found_fault_to_pattern(), "good": s.good_results(), "first": s.first_detect_round()}
    if not drop:
        out["det"] = s.detect_matrix()
    return out


def _same(a, b):
    assert set(a) == set(b)
    for k in a:
        if k == "good":
            assert all(np.array_equal(x, y) for x, y in zip(a[k], b[k])), k
        else:
            assert np.array_equal(a[k], b[k]), k


def _reference_run(ckt, flst, ptn, drop, mode):
    s = lsim_amd.Simulator(device=0)
    s.set_option("drop", int(drop))
    s.read_text(ckt, flst, ptn)
    s.run(mode)
    out = _result(s, drop)
    s.close()
    return out


CASES = [
    # 5,000 gates, 10,000 faults, 16 rounds, full d(f, r) matrix
    (workloads.texts("c1"), False, M.SEQUENTIAL),
    # 600 gates, 1,200 faults, 3 rounds with a partial last one, dropping
    (lsim_amd.generate(seed=41, num_pis=16, num_pos=8, num_gates=600, num_levels=14, num_patterns=150,
                       fanout_cap=20), True, M.SEQUENTIAL),
]


def test_noalloc_two_calls_match_run():
    s = lsim_amd.Simulator(device=0)
    for (ckt, flst, ptn), drop, mode in CASES:
        want = _reference_run(ckt, flst, ptn, drop, mode)
        s.set_option("drop", int(drop))
        s.read_text(ckt, flst, ptn)
        s.run_noalloc(mode)
        _same(_result(s, drop), want)
        st = s.last_run_stats()
        assert st["device_allocs"] == 0, st
        assert st["engine"] == 1
    # the plain per-level form through the same entry point
    s.run_noalloc(M.PLAIN)
    _same(_result(s, True), want)
    assert s.last_run_stats()["device_allocs"] == 0
    # lsim_run on the same handle afterwards
    s.run(mode)
    _same(_result(s, True), want)
    s.close()


def test_noalloc_rejects_run_above_scratch():
    ckt, flst, ptn = workloads.texts("c2")  # 100k gates x 128 rounds: 100 MB of good values alone
    s = lsim_amd.Simulator(device=0)
    s.read_text(ckt, flst, ptn)
    with pytest.raises(lsim_amd.LsimError) as e:
        s.run_noalloc(M.SEQUENTIAL)
    assert e.value.code == lsim_amd.LSIM_E_NOMEM and "MiB" in e.value.message
    s.close()
