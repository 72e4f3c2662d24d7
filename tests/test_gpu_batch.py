"""Batching and serving on the GPU: every job's result is the same as when the
job runs alone through the single-simulator path (and, for the reference-made
vectors, the same as the reference's per-round detections)."""
import threading

import numpy as np
import pytest

import lsim_amd
import oracle

pytestmark = pytest.mark.gpu


def _single(ckt, flst, ptn, drop):
    s = lsim_amd.Simulator(device=0)
    s.set_option("drop", int(drop))
    s.read_text(ckt, flst, ptn)
    s.run(lsim_amd.Simulator.Mode.PARTITIONED)
    out = {"first": s.first_detect_round(), "table": s.found_fault_to_pattern()}
    if not drop:
        out["det"] = s.detect_matrix()
    s.close()
    return out


# (seed, P, O, G, L, n, fanout cap, fault fraction)
SPECS = [
    (31, 16, 8, 600, 12, 200, 0, 1.0),
    (32, 32, 16, 3000, 24, 1024, 40, 1.0),
    (33, 64, 32, 12000, 60, 4000 + 17, 100, 0.5),
    (34, 24, 12, 1500, 30, 640, 0, 1.0),
    (35, 128, 64, 30000, 90, 2048, 200, 0.2),
]


def _texts(spec, pattern_seed=None, fraction=None):
    seed, P, O, G, L, n, cap, frac = spec
    ckt, flst, ptn = lsim_amd.generate(seed=seed, num_pis=P, num_pos=O, num_gates=G, num_levels=L, num_patterns=n,
                                       fanout_cap=cap, fault_fraction=frac if fraction is None else fraction)
    if pattern_seed is not None:
        _, flst, ptn = lsim_amd.generate(seed=pattern_seed, num_pis=P, num_pos=O, num_gates=G, num_levels=L,
                                         num_patterns=n, fanout_cap=cap,
                                         fault_fraction=frac if fraction is None else fraction)
    return ckt, flst, ptn


@pytest.mark.parametrize("kernel", [0, 1], ids=["pairs", "nodewave"])
def test_batch_matches_single_runs(kernel):
    b = lsim_amd.Batch(device=0, sweep_kernel=kernel)
    jobs = []
    for i, spec in enumerate(SPECS):
        ckt, flst, ptn = _texts(spec)
        cid = b.add_circuit(ckt)
        drop = i % 2 == 1
        jobs.append((b.add_job(cid, flst, ptn, drop=drop), ckt, flst, ptn, drop))
        # a second job on the same circuit with other patterns and faults
        _, flst2, ptn2 = _texts(spec, pattern_seed=spec[0] + 100, fraction=0.3)
        jobs.append((b.add_job(cid, flst2, ptn2, drop=not drop), ckt, flst2, ptn2, not drop))
    b.run()
    for jid, ckt, flst, ptn, drop in jobs:
        want = _single(ckt, flst, ptn, drop)
        assert np.array_equal(b.first_detect_round(jid), want["first"]), jid
        assert np.array_equal(b.found_fault_to_pattern(jid), want["table"]), jid
        if not drop:
            assert np.array_equal(b.detect_matrix(jid), want["det"]), jid
    st = b.last_run_stats()
    assert st["steps"] > 0 and st["stems"] > 0


def test_batch_reference_made_vectors(ref_vectors):
    """Jobs of different circuits, with and without MUX, with an empty fault
    list and with no patterns, in one batch: each case once with the full
    d(f, r) matrix and once with fault dropping."""
    b = lsim_amd.Batch(device=0)
    ids = []
    for c in ref_vectors:
        cid = b.add_circuit(c["ckt"])
        ids.append((b.add_job(cid, c["flst"], c["ptn"], drop=False), b.add_job(cid, c["flst"], c["ptn"], drop=True), c))
    b.run()
    for jid, jdrop, c in ids:
        F = b._L.lsim_batch_job_faults(b._h, jid)
        R = b._L.lsim_batch_job_rounds(b._h, jid)
        det = np.array(c["det"], dtype=np.uint8).reshape(F, R)
        assert np.array_equal(b.detect_matrix(jid), det), c["name"]
        assert list(b.found_fault_to_pattern(jid)) == c["table"], c["name"]
        want_first = np.where(det.any(1), det.argmax(1), -1) if R else -np.ones(F, dtype=np.int64)
        assert np.array_equal(b.first_detect_round(jdrop), want_first), c["name"]


def test_batch_plain_mode_matches():
    spec = SPECS[1]
    ckt, flst, ptn = _texts(spec)
    o = oracle.Oracle(ckt, flst, ptn)
    want = o.run_event(threads=0)
    b = lsim_amd.Batch(device=0, mode=0)
    cid = b.add_circuit(ckt)
    j1 = b.add_job(cid, flst, ptn, drop=False)
    j2 = b.add_job(cid, flst, ptn, drop=True)
    b.run()
    assert np.array_equal(b.detect_matrix(j1), want)
    assert np.array_equal(b.first_detect_round(j2), np.where(want.any(1), want.argmax(1), -1))


def test_server_same_result_alone_and_in_mix():
    srv = lsim_amd.Server(device=0, job_budget_mb=16 * 1024, scratch_mb=4096)
    circuits = []
    for spec in SPECS:
        ckt, flst, ptn = _texts(spec)
        circuits.append((srv.add_circuit(ckt), ckt, spec))
    cid, ckt, spec = circuits[2]
    _, flst, ptn = _texts(spec)
    alone = srv.submit(cid, flst, ptn, drop=True)
    srv.wait(alone, 120)
    first_alone = srv.first_detect_round(alone)
    # A burst from several threads; the same job again in the middle of it.
    submitted = []
    lock = threading.Lock()

    def client(k):
        for i, (c, ck, sp) in enumerate(circuits):
            if (i + k) % 2:
                continue
            _, f2, p2 = _texts(sp, pattern_seed=sp[0] + 7 * k + 1, fraction=0.4)
            jid = srv.submit(c, f2, p2, drop=True)
            with lock:
                submitted.append((jid, ck, f2, p2))

    threads = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    again = srv.submit(cid, flst, ptn, drop=True)
    for t in threads:
        t.join()
    srv.wait(again, 300)
    assert np.array_equal(srv.first_detect_round(again), first_alone)
    assert np.array_equal(first_alone, _single(ckt, flst, ptn, True)["first"])
    for jid, ck, f2, p2 in submitted:
        srv.wait(jid, 300)
        assert np.array_equal(srv.first_detect_round(jid), _single(ck, f2, p2, True)["first"]), jid
        t_sub, t_adm, t_fin = srv.job_times(jid)
        assert t_sub <= t_adm <= t_fin
    cnt = srv.counters()
    assert cnt["finished"] == len(submitted) + 2 and cnt["failed"] == 0
    assert cnt["peak_admitted"] >= 2
    srv.close()


def test_server_rejects_job_above_budget():
    srv = lsim_amd.Server(device=0, job_budget_mb=1, scratch_mb=256)
    ckt, flst, ptn = _texts(SPECS[4])
    cid = srv.add_circuit(ckt)
    jid = srv.submit(cid, flst, ptn)
    with pytest.raises(lsim_amd.LsimError) as e:
        srv.wait(jid, 60)
    assert e.value.code == lsim_amd.LSIM_E_NOMEM
    # the server keeps serving
    ckt2, flst2, ptn2 = _texts(SPECS[0])
    cid2 = srv.add_circuit(ckt2)
    j2 = srv.submit(cid2, flst2, ptn2)
    srv.wait(j2, 60)
    srv.close()


def test_autotuned_batch_matches_oracle():
    """With the autotuner on, successive batches over the same circuits each
    measure one sweep configuration (coordinate descent over word schedule,
    pair workgroup size, dense-pass threshold, step cap and kernel: 12
    distinct configurations for dropping jobs; a run that allocated sweep
    scratch is not timed, so its configuration runs again), then keep the
    fastest; every batch's first detecting rounds equal the oracle's over all
    rounds, for every fault."""
    b = lsim_amd.Batch(device=0, autotune=1)
    cids, texts = [], []
    for spec in SPECS[:3]:
        ckt, flst, ptn = _texts(spec)
        cids.append(b.add_circuit(ckt))
        texts.append((ckt, flst, ptn))
    want = []
    for ck, f, p in texts:
        o = oracle.Oracle(ck, f, p)
        want.append(o.first_rounds(np.arange(o.F)))
    seen = []
    trials = 0
    for rnd in range(20):
        jobs = [b.add_job(c, f, p, drop=True) for c, (ck, f, p) in zip(cids, texts)]
        b.run()
        st = b.last_run_stats()
        seen.append(st["tuned"])
        assert trials <= st["tune_trials"] <= trials + 1
        trials = st["tune_trials"]
        for jid, w in zip(jobs, want):
            assert np.array_equal(b.first_detect_round(jid), w), (rnd, jid, st["tuned"])
            b.release_job(jid)
    assert st["tune_candidates"] == 12 and trials == 12
    assert len(set(seen)) == 12
    assert len(set(seen[-3:])) == 1  # explored: the fastest is kept


def test_server_two_lanes_match_single_runs():
    """Jobs split between the batching lane and the one-at-a-time lane (the
    larger circuits' jobs by node-words), submitted together from several
    threads: every result equals the single-simulator run."""
    srv = lsim_amd.Server(device=0, job_budget_mb=16 * 1024, scratch_mb=4096)
    # jobs on the two largest circuits (nodes x rounds >= 720,000) take the large-job lane
    srv.set_option("big_node_words", 12000 * 60)
    circuits = []
    for spec in SPECS:
        ckt, flst, ptn = _texts(spec)
        circuits.append((srv.add_circuit(ckt), ckt, spec))
    submitted = []
    lock = threading.Lock()

    def client(k):
        for c, ck, sp in circuits:
            _, f2, p2 = _texts(sp, pattern_seed=sp[0] + 11 * k + 3, fraction=0.5)
            jid = srv.submit(c, f2, p2, drop=True)
            with lock:
                submitted.append((jid, ck, f2, p2))

    threads = [threading.Thread(target=client, args=(k,)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for jid, ck, f2, p2 in submitted:
        srv.wait(jid, 300)
        assert np.array_equal(srv.first_detect_round(jid), _single(ck, f2, p2, True)["first"]), jid
    cnt = srv.counters()
    assert cnt["finished"] == len(submitted) and cnt["failed"] == 0
    srv.close()


def test_batched_good_machine_graph_replays():
    """The batched good machine is captured in a graph once a job set
    recurs and replayed after that, and so are the word steps' launches
    (plan, sweep, detection) once a step's launch signature recurs; four
    batches of the same set, with an eager batch of another set on another
    engine between them, all give the oracle's first detecting rounds, the
    good-machine graph is replayed from the second batch on, and every word
    step of the third and fourth batches replays its captured graphs (the
    step buffers grow during the first batch, so the second one's launch
    signatures are partly new)."""
    b = lsim_amd.Batch(device=0)
    other = lsim_amd.Batch(device=0, good_graph=0)
    cids, texts, want = [], [], []
    for spec in SPECS[:4]:
        ckt, flst, ptn = _texts(spec)
        cids.append(b.add_circuit(ckt))
        texts.append((ckt, flst, ptn))
        o = oracle.Oracle(ckt, flst, ptn)
        want.append(o.first_rounds(np.arange(o.F)))
    ocid = other.add_circuit(texts[0][0])
    replayed, steps = [], []
    for rnd in range(4):
        jobs = [b.add_job(c, f, p, drop=True) for c, (ck, f, p) in zip(cids, texts)]
        b.run()
        st = b.last_run_stats()
        replayed.append(st["good_replayed"])
        steps.append((st["steps_replayed"], st["steps"]))
        for jid, w in zip(jobs, want):
            assert np.array_equal(b.first_detect_round(jid), w), (rnd, jid)
            b.release_job(jid)
        # eager work of another engine between the replays
        oj = [other.add_job(ocid, texts[0][1], texts[0][2], drop=False), other.add_job(ocid, texts[0][1], texts[0][2],
                                                                                       drop=True)]
        other.run()
        assert other.last_run_stats()["good_replayed"] == 0
        assert np.array_equal(other.first_detect_round(oj[1]), want[0])
        for j in oj:
            other.release_job(j)
    assert replayed == [0, 1, 1, 1]
    assert steps[0][0] < steps[0][1] and all(r == n and n > 0 for r, n in steps[2:]), steps


def test_server_failure_fails_only_its_job():
    """A job whose circuit is too large for the lane's sweep scratch (32 MiB:
    a tile slot of a 600k-gate circuit does not fit) fails with
    LSIM_E_NOMEM; a small job batched in the same lane at the same time is
    run again alone and finishes with the single-simulator result."""
    srv = lsim_amd.Server(device=0, job_budget_mb=16 * 1024, scratch_mb=100)
    srv.set_option("big_evals", 0)  # one lane for both jobs
    small = _texts(SPECS[0])
    big_ckt, big_flst, big_ptn = lsim_amd.generate(seed=77, num_pis=128, num_pos=64, num_gates=600000,
                                                   num_levels=200, num_patterns=256, fanout_cap=200,
                                                   fault_fraction=0.001)
    cs = srv.add_circuit(small[0])
    cb = srv.add_circuit(big_ckt)
    jb = srv.submit(cb, big_flst, big_ptn)
    js = srv.submit(cs, small[1], small[2])
    with pytest.raises(lsim_amd.LsimError) as e:
        srv.wait(jb, 120)
    assert e.value.code == lsim_amd.LSIM_E_NOMEM
    srv.wait(js, 120)
    want = _single(small[0], small[1], small[2], True)["first"]
    assert np.array_equal(srv.first_detect_round(js), want)
    cnt = srv.counters()
    assert cnt["failed"] == 1 and cnt["finished"] >= 1
    srv.close()
