"""The batching and serving workloads at their configured sizes, against the oracle.

Batching (256 seeded circuits, 1k-200k gates, full SA0/SA1 fault lists,
2,048-16,384 patterns, one batch with fault dropping): for every circuit a
seeded fault sample has the first detecting round of the event-driven
restatement over all of the circuit's rounds, and the literal report of those
faults is the oracle's last-round detection (fsim/simulation.cpp:131-145).
The eight smallest circuits also run without dropping in the same batch: their
full d(f, r) matrices and reports equal the literal restatement.

Serving (64 resident circuits, 10k-2M gates): a burst of 40 jobs submitted at
once (1-100 % of a circuit's faults, 1,024-32,768 patterns) is admitted and
batched as memory allows; every job's sampled first rounds equal the oracle,
and a set of the jobs run again alone gives the same results as in the mix.

A scaled-down batching slice (48 circuits of both narrow and wide level
widths, with and without dropping) matches single-simulator runs on every
fault, including full matrices.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import lsim_amd
import oracle
from lsim_amd import workloads

pytestmark = pytest.mark.gpu


def _gen(fn, items):
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(fn, items))


@pytest.fixture(scope="module")
def c4():
    specs = workloads.c4_specs()
    ckts = _gen(workloads.circuit_text, specs)
    jobs = _gen(workloads.batch_job_texts, specs)
    return specs, ckts, jobs


def _sample(rng, first, n_uniform, n_late):
    F = first.size
    pick = [rng.choice(F, min(F, n_uniform), replace=False)]
    late = np.flatnonzero(first >= 1)
    if late.size:
        pick.append(rng.choice(late, min(late.size, n_late), replace=False))
    return np.unique(np.concatenate(pick))


def test_c4_batch_against_oracle(c4):
    specs, ckts, jobs = c4
    b = lsim_amd.Batch(device=0)
    cids = [b.add_circuit(c) for c in ckts]
    ids = [b.add_job(cid, f, p, drop=True) for cid, (f, p) in zip(cids, jobs)]
    small = sorted(range(len(specs)), key=lambda i: specs[i]["num_gates"])[:8]
    full_ids = {i: b.add_job(cids[i], jobs[i][0], jobs[i][1], drop=False) for i in small}
    b.run()
    rng = np.random.default_rng(4004)
    for i, spec in enumerate(specs):
        first = b.first_detect_round(ids[i])
        table = b.found_fault_to_pattern(ids[i])
        assert first.size == 2 * spec["num_gates"], i
        o = oracle.Oracle(ckts[i], jobs[i][0], jobs[i][1])
        R = o.R
        assert np.all(table[1::2] == R - 1), i
        sel = _sample(rng, first, 12, 12)
        want = o.first_rounds(sel, chunk=64)
        assert np.array_equal(first[sel], want), (i, spec, sel[first[sel] != want][:5])
        last = o.run_event_range(sel, R - 1, R)[:, 0]
        assert np.array_equal(table[0::2][sel], last.astype(np.uint64)), (i, spec)
    for i, jid in full_ids.items():
        ref = oracle.Oracle(ckts[i], jobs[i][0], jobs[i][1]).run_literal(threads=0)
        det = ref["det"]
        assert np.array_equal(b.detect_matrix(jid), det), (i, specs[i])
        assert np.array_equal(b.found_fault_to_pattern(jid), ref["table"]), i
        assert np.array_equal(b.first_detect_round(ids[i]), np.where(det.any(1), det.argmax(1), -1)), i
    st = b.last_run_stats()
    assert st["steps"] > 0 and st["engine"] == 1


def _single(ckt, flst, ptn, drop):
    s = lsim_amd.Simulator(device=0)
    s.set_option("drop", int(drop))
    s.read_text(ckt, flst, ptn)
    s.run(lsim_amd.Simulator.Mode.SEQUENTIAL)
    out = {"first": s.first_detect_round(), "table": s.found_fault_to_pattern()}
    if not drop:
        out["det"] = s.detect_matrix()
    s.close()
    return out


def test_c4_slice_matches_single_runs(c4):
    """48 circuits spread over the size range (levels from ~120 to ~210,
    narrow and wide), every other job without dropping: every job's first
    rounds, report and (without dropping) d(f, r) matrix equal a lone run."""
    specs, ckts, jobs = c4
    order = sorted(range(len(specs)), key=lambda i: specs[i]["num_gates"])
    pick = order[::len(order) // 48][:48]
    b = lsim_amd.Batch(device=0)
    ids = []
    for k, i in enumerate(pick):
        cid = b.add_circuit(ckts[i])
        drop = k % 2 == 0
        ids.append((b.add_job(cid, jobs[i][0], jobs[i][1], drop=drop), i, drop))
    b.run()
    for jid, i, drop in ids:
        want = _single(ckts[i], jobs[i][0], jobs[i][1], drop)
        assert np.array_equal(b.first_detect_round(jid), want["first"]), (i, drop)
        assert np.array_equal(b.found_fault_to_pattern(jid), want["table"]), (i, drop)
        if not drop:
            assert np.array_equal(b.detect_matrix(jid), want["det"]), i


@pytest.fixture(scope="module")
def c5():
    circuits = workloads.c5_circuit_specs()
    return circuits, _gen(workloads.circuit_text, circuits)


def test_c5_server_burst_against_oracle(c5):
    circuits, ckts = c5
    srv = lsim_amd.Server(device=0, job_budget_mb=64 * 1024)
    cids = [srv.add_circuit(c) for c in ckts]
    jobs = workloads.c5_jobs(circuits, 40)
    texts = _gen(lambda j: workloads.job_texts(circuits[j["circuit"]], j), jobs)
    ids = [srv.submit(cids[j["circuit"]], f, p, drop=True) for j, (f, p) in zip(jobs, texts)]
    for jid in ids:
        srv.wait(jid, 300)
    firsts = [srv.first_detect_round(jid) for jid in ids]
    cnt = srv.counters()
    for jid in ids:
        srv.release_job(jid)
    assert cnt["finished"] >= len(jobs) and cnt["failed"] == 0
    assert cnt["peak_admitted"] > 1  # jobs ran together
    rng = np.random.default_rng(5005)
    for k, (j, (f, p)) in enumerate(zip(jobs, texts)):
        first = firsts[k]
        o = oracle.Oracle(ckts[j["circuit"]], f, p)
        assert first.size == o.F
        sel = _sample(rng, first, 16, 8)
        want = o.first_rounds(sel, chunk=64)
        assert np.array_equal(first[sel], want), (k, j, sel[first[sel] != want][:5])
    # the same results alone as in the mix
    for k in range(0, len(jobs), 5):
        jid = srv.submit(cids[jobs[k]["circuit"]], texts[k][0], texts[k][1], drop=True)
        srv.wait(jid, 300)
        assert np.array_equal(srv.first_detect_round(jid), firsts[k]), k
        srv.release_job(jid)
    srv.close()
