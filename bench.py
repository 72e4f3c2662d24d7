"""Benchmark: fault x pattern evaluations per second on one MI355X.

A step is one run() of the fault-simulation path over the resident workload:
good machine over every round, then the fault sweeps and detection (with fault
dropping where the configuration asks for it). Circuit parsing and the upload
to HBM happen before the timed region.

The default run measures the 2M-gate workload (the headline line) and, in the
same JSON line, the two systems forms on their own workloads: `batching` (256
circuits in one batch, autotuned, evals/s and ms per batch) and `serving` (64
resident circuits, Poisson job arrivals: p50/p99 latency beside unloaded
latency, throughput, and whether every job's result is the same alone and in
the mix).

Prints one JSON line (rank 0). `roofline.achieved` is the sweep kernels' own
count of the bytes they read and write (lsim_run_stats.sweep_bytes) over the
HIP-event time of their launches; `roofline.traffic` and the request rate come
from the committed rocprofv3 summary of the same workload under profiles/.
`cpu_baseline` times the oracle's restatement of the reference CPU simulator
on a seeded sample of faults and rounds of the same circuit; its cost is linear
in faults x rounds.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def profiled_value(config: str, kernel: str, key: str = "traffic_bytes_per_launch"):
    """A per-launch figure of `kernel` on `config` from its committed rocprofv3
    summary: `traffic_bytes_per_launch` (HBM bytes from the read and write
    request counts by size, TCC_EA0_RDREQ_128B/64B/32B and TCC_EA0_WRREQ_64B)
    or `transactions_per_launch` (TCC_EA0_RDREQ + TCC_EA0_WRREQ requests);
    None when absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{config}.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    return None


def _probe_lib():
    import ctypes

    path = os.path.join(ROOT, "lsim_amd", "liblsim_probe.so")
    if not os.path.exists(path):
        raise RuntimeError("lsim_amd/liblsim_probe.so is not built (python lsim_amd/build.py)")
    return ctypes.CDLL(path)


def copy_probe(nbytes: int, reps: int = 3) -> dict:
    """A plain device copy of nbytes, reps times (the known-size kernel
    k_probe_copy that memory counters are checked against)."""
    import ctypes

    fn = _probe_lib().lsim_bench_copy
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    rc = fn(0, int(nbytes), int(reps), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"copy probe failed ({rc})")
    return {"copy_bytes": int(nbytes), "reps": reps, "ms_per_copy": ms.value,
            "gbs": 2 * nbytes / (ms.value * 1e-3) / 1e9}


def probe_random(nbytes: int = 16 << 30) -> dict:
    """Random 8-byte reads per second the device serves, from the measurement
    library lsim_amd/liblsim_probe.so (not part of the simulator library)."""
    import ctypes

    fn = _probe_lib().lsim_bench_probe_random_reads
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    rate, lat = ctypes.c_double(), ctypes.c_double()
    rc = fn(0, int(nbytes), ctypes.byref(rate), ctypes.byref(lat))
    if rc != 0:
        raise RuntimeError(f"random-read probe failed ({rc})")
    return {"reads_per_s": rate.value, "latency_ns": lat.value, "buffer_bytes": int(nbytes)}


def cpu_baseline(ckt: bytes, flst: bytes, ptn: bytes, preset: dict, budget_s: float) -> dict:
    """Time the oracle's restatement of the reference CPU simulator
    (oracle_run_literal: a whole-circuit faulty pass per fault and round,
    fsim/simulation.cpp:104-228) on a bounded sample of the same workload: a
    seeded subset of faults over the first rounds, on one core and on the
    host's cores (faults split over threads). Its cost is F x R x N node
    evaluations whatever the detections, so the sample's rate stands for the
    whole job."""
    import numpy as np

    import oracle

    lines = flst.split(b"\n")
    F = int(lines[0])
    N = preset["num_pis"] + preset["num_pos"] + preset["num_gates"]
    rounds_total = (preset["num_patterns"] + 63) // 64
    # The restatement evaluates about 7e7 nodes/s on one core; size the sample
    # to about budget_s seconds.
    fr_budget = max(4, int(budget_s * 7e7 / N))
    n_rounds = max(1, min(rounds_total, 4, fr_budget // 4))
    n_faults = max(1, min(F, fr_budget // n_rounds))
    rng = np.random.default_rng(4242)
    pick = np.sort(rng.choice(F, size=n_faults, replace=False))
    sub_flst = (f"{n_faults}\n".encode() + b"\n".join(lines[1 + i] for i in pick) + b"\n")
    ptn_lines = ptn.split(b"\n")
    n_pat = min(preset["num_patterns"], 64 * n_rounds)
    sub_ptn = f"{n_pat}\n".encode() + b"\n".join(ptn_lines[1:1 + n_rounds]) + b"\n"
    sample = (f"{n_faults} seeded faults x {n_rounds} rounds ({n_pat} patterns) of the full circuit "
              f"({N} nodes)")
    log(f"[bench] cpu baseline: {sample}")
    # Host cores for the threaded run: the box's CPU share (16 per GPU) at most.
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))
    o = oracle.Oracle(ckt, sub_flst, sub_ptn)
    t0 = time.perf_counter()
    o.run_literal(threads=1)
    secs = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.run_literal(threads=threads)
    psecs = time.perf_counter() - t0
    return {"value": n_faults * n_pat / secs, "unit": "fault*pattern evals/s", "cores": 1, "kind": "port",
            "sample": sample, "seconds": round(secs, 3), "cpu": _cpu_model(),
            "algorithm": "restatement of the reference SEQUENTIAL mode (oracle_run_literal)",
            "threaded": {"value": n_faults * n_pat / psecs, "cores": threads, "seconds": round(psecs, 3),
                         "split": "faults over OpenMP threads"}}


def plain_form_sample(ckt: bytes, flst: bytes, ptn: bytes, stride: int) -> dict:
    """Every `stride`-th fault of the workload through the plain per-level
    form (LSIM_MODE_PLAIN: one launch per level for the good machine and for
    every sweep) and through the persistent form, both with dropping, each
    timed on its second run: evals/s of both on the same sample."""
    import lsim_amd

    lines = flst.split(b"\n")
    F = int(lines[0])
    pick = list(range(0, F, stride))
    sub = f"{len(pick)}\n".encode() + b"\n".join(lines[1 + i] for i in pick) + b"\n"
    n = int(ptn.split(b"\n", 1)[0])
    sim = lsim_amd.Simulator(device=0)
    sim.set_option("drop", 1)
    sim.set_option("report_last_round", 0)
    sim.read_text(ckt, sub, ptn)
    timed = {}
    for name, mode in (("plain", lsim_amd.Simulator.Mode.PLAIN), ("persistent", lsim_amd.Simulator.Mode.SEQUENTIAL)):
        sim.run(mode)  # warm: scratch allocation
        t0 = time.perf_counter()
        sim.run(mode)
        timed[name] = (time.perf_counter() - t0, sim.last_run_stats())
    sim.close()
    secs, st = timed["plain"]
    v = len(pick) * n / secs
    vp = len(pick) * n / timed["persistent"][0]
    return {"value": v, "unit": "fault*pattern evals/s", "engine": "k_sweep_level" if st["engine"] == 0 else st["engine"],
            "sample": f"every {stride}th fault ({len(pick)} faults), all {n} patterns, dropping on",
            "seconds": round(secs, 3), "sweep_launches": st["sweep_launches"],
            "persistent_value": vp, "persistent_seconds": round(timed["persistent"][0], 4),
            "persistent_over_plain": vp / v}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _roofline(kernel: str, agg: dict, config: str, probe: bool = True) -> dict:
    launches = max(1, agg["sweep_launches"])
    t_launch_s = agg["t_sweep_ms"] / 1000.0 / launches
    bytes_launch = agg["sweep_bytes"] / launches
    achieved = bytes_launch / t_launch_s / 1e9 if t_launch_s > 0 else 0.0
    traffic = profiled_value(config, kernel)
    out = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
           "bytes_per_launch": bytes_launch, "avg_launch_ms": t_launch_s * 1000.0}
    if traffic and t_launch_s > 0:
        # rocprof-measured HBM bytes per launch over this run's launch time
        out["traffic_gbs"] = traffic / t_launch_s / 1e9
        out["traffic_frac"] = out["traffic_gbs"] / HBM_PEAK_GBS
    # The sweeps' accesses are small records at random node positions: their
    # memory requests per second against the rate the device serves random
    # 8-byte reads (measured live on this device).
    tx = profiled_value(config, kernel, "transactions_per_launch")
    if not probe:
        if tx and t_launch_s > 0:
            out["transactions_per_s"] = tx / t_launch_s
        return out
    try:
        pr = probe_random()
    except Exception as e:  # reported, never fatal for the GPU figure
        pr = {"error": str(e)}
    ra = {"probe_reads_per_s": pr.get("reads_per_s"), "probe_latency_ns": pr.get("latency_ns"),
          "transactions_per_launch": tx}
    if "error" in pr:
        ra["probe_error"] = pr["error"]
    if tx and t_launch_s > 0:
        ra["transactions_per_s"] = tx / t_launch_s
        if pr.get("reads_per_s"):
            ra["frac"] = ra["transactions_per_s"] / pr["reads_per_s"]
    out["random_access"] = ra
    return out


def _kernel_name(args, tuned=None) -> str:
    if args.mode == "plain":
        return "k_sweep_level"
    kernel = (tuned & 0xff) if tuned is not None else (0 if args.sweep_kernel == "pairs" else 1)
    return "k_sweep_pairs" if kernel == 0 else "k_sweep_events"


def _tuned(st: dict) -> dict:
    t = int(st["tuned"])
    return {"sweep_kernel": "pairs" if (t & 0xff) == 0 else "nodewave", "first_step_words": (t >> 8) & 0xff,
            "word_growth": (t >> 16) & 0xff, "pair_threads": ((t >> 24) & 0xff) * 256,
            "pair_dense": (t >> 32) & 0xff, "max_step_words": (t >> 40) & 0xffff}


def _batch_options(batch, args) -> None:
    for key in ("pair_dense", "pair_nopend", "pair_fo_pre", "pair_nopend_pct", "max_step_words", "device_steps"):
        v = getattr(args, key)
        if v >= 0:
            batch.set_option(key, v)


def batching_leg(args, steps: int, warmup: int) -> dict:
    """256 circuits of uneven size (1k-200k gates), each with its full fault
    list and its own pattern block (2,048-16,384 patterns), run together
    through the batch API: one persistent sweep launch per word step over all
    jobs, fault dropping on. A step adds the 256 jobs (parse and pattern
    upload, untimed) and times run() to the last result on the host. With the
    autotuner on, the warmup runs explore its candidates."""
    import lsim_amd
    from lsim_amd import workloads

    specs = workloads.c4_specs()
    t0 = time.perf_counter()
    batch = lsim_amd.Batch(device=0, mode=1 if args.mode == "partitioned" else 0,
                           sweep_kernel=0 if args.sweep_kernel == "pairs" else 1, autotune=args.c4_autotune,
                           pair_threads=args.pair_threads)
    _batch_options(batch, args)
    cids, jobs_text = [], []
    for i, sp in enumerate(specs):
        cids.append(batch.add_circuit(workloads.circuit_text(sp)))
        jobs_text.append(workloads.batch_job_texts(sp))
        if i % 64 == 63:
            log(f"[bench] c4: {i + 1} circuits resident ({time.perf_counter() - t0:.1f}s)")
    F_tot = sum(2 * sp["num_gates"] for sp in specs)
    evals = sum(2 * sp["num_gates"] * sp["num_patterns"] for sp in specs)
    times, agg, detected, st = [], None, 0, None
    it, timed_runs, kept_runs = 0, 0, 0
    while timed_runs < steps:
        # Timed once the warmup is done and, with the autotuner, once a run
        # has reported every candidate measured (it then keeps the fastest)
        # and one run with the kept configuration has captured its step graphs.
        explored = st is not None and st["tune_trials"] >= st["tune_candidates"]
        timed = it >= warmup and (not args.c4_autotune or kept_runs >= 1 or it > 60)
        kept_runs += 1 if explored else 0
        ids = [batch.add_job(c, f, p, drop=True) for c, (f, p) in zip(cids, jobs_text)]
        t0 = time.perf_counter()
        batch.run()
        dt = time.perf_counter() - t0
        st = batch.last_run_stats()
        it += 1
        if timed:
            timed_runs += 1
            times.append(dt)
            if agg is None:
                agg = {k: 0 for k in st}
            for k, v in st.items():
                agg[k] += v
            detected = sum(int((batch.first_detect_round(j) >= 0).sum()) for j in ids)
        log(f"[bench] c4 {'step' if timed else 'warmup'} {dt:.3f}s sweep {st['t_sweep_ms']:.1f}ms "
            f"detect {st['t_detect_ms']:.1f}ms good {st['t_good_ms']:.1f}ms steps {st['steps']} "
            f"launches {st['sweep_launches']} tuned {_tuned(st)}")
        for j in ids:
            batch.release_job(j)
    batch.close()
    elapsed = sum(times)
    kernel = _kernel_name(args, st["tuned"])
    return {
        "metric": "fault x pattern evals/sec",
        "value": evals / (elapsed / steps),
        "unit": "fault*pattern evals/s",
        "ms_per_batch": 1000.0 * elapsed / steps,
        "steps": steps, "warmup": it - steps,
        "workload": f"c4: {len(specs)} seeded circuits, {sum(sp['num_gates'] for sp in specs)} gates in all "
                    f"(1k-200k each), {F_tot} SA0/SA1 faults, 2,048-16,384 patterns per circuit, "
                    "fault dropping on, one batch",
        "faults": F_tot, "evals_per_batch": evals, "detected": detected,
        "kernel": kernel,
        "sweep": _roofline(kernel, agg, "c4", probe=False),
        "autotune": dict(_tuned(st), enabled=bool(args.c4_autotune), candidates_measured=st["tune_trials"],
                         candidates=st["tune_candidates"]),
        "breakdown_ms": {k: agg[k] / steps for k in ("t_good_ms", "t_sweep_ms", "t_detect_ms", "t_total_ms")},
    }


def bench_batch(args) -> int:
    leg = batching_leg(args, args.steps, args.warmup)
    out = {
        "metric": leg["metric"], "value": leg["value"], "unit": leg["unit"],
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": leg["ms_per_batch"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": leg["workload"], "engine": args.mode, "faults": leg["faults"],
                   "evals_per_step": leg["evals_per_batch"], "detected": leg["detected"]},
        "roofline": leg["sweep"], "autotune": leg["autotune"], "breakdown_ms": leg["breakdown_ms"],
    }
    print(json.dumps(out), flush=True)
    return 0


def _pct(xs, q):
    import numpy as np

    return float(np.percentile(np.asarray(xs), q)) if len(xs) else 0.0


def serving_leg(args, n_jobs: int) -> dict:
    """Serving over 64 resident circuits. Phase 1: every job alone (unloaded
    latency). Phase 2: all jobs at once (saturated job rate). Phase 3: the same
    jobs as a Poisson stream at `--c5-load` x the saturated rate; reports
    p50/p99 latency (submit to result), jobs/s and evals/s, and checks that
    every job's first-detection rounds equal its unloaded run's."""
    import random

    import numpy as np

    import lsim_amd
    from lsim_amd import workloads

    circuits = workloads.c5_circuit_specs()
    t0 = time.perf_counter()
    srv = lsim_amd.Server(device=0, job_budget_mb=64 * 1024)
    srv.set_option("word_growth", args.c5_growth)
    if args.device_steps >= 0:
        srv.set_option("device_steps", args.device_steps)
    cids = []
    for i, sp in enumerate(circuits):
        cids.append(srv.add_circuit(workloads.circuit_text(sp)))
        if i % 8 == 7:
            log(f"[bench] c5: {i + 1} circuits resident ({time.perf_counter() - t0:.1f}s)")
    jobs = workloads.c5_jobs(circuits, n_jobs)
    texts = [workloads.job_texts(circuits[j["circuit"]], j) for j in jobs]
    F = [int(f.split(b"\n", 1)[0]) for f, _ in texts]
    evals = sum(f * j["num_patterns"] for f, j in zip(F, jobs))

    def submit(k):
        f, p = texts[k]
        return srv.submit(cids[jobs[k]["circuit"]], f, p, drop=True)

    # One untimed job first: the sweep scratch is allocated on first use.
    w = submit(0)
    srv.wait(w, 600)
    srv.release_job(w)

    # Phase 1: unloaded.
    alone, first = [], []
    for k in range(len(jobs)):
        jid = submit(k)
        srv.wait(jid, 600)
        ts, ta, tf = srv.job_times(jid)
        alone.append(tf - ts)
        first.append(srv.first_detect_round(jid))
        srv.release_job(jid)
    log(f"[bench] c5 unloaded: p50 {_pct(alone, 50) * 1e3:.1f}ms p99 {_pct(alone, 99) * 1e3:.1f}ms "
        f"sum {sum(alone):.2f}s")
    # Phase 2: saturated.
    t0 = time.perf_counter()
    ids = [submit(k) for k in range(len(jobs))]
    for jid in ids:
        srv.wait(jid, 600)
    t_sat = time.perf_counter() - t0
    exact = all(np.array_equal(srv.first_detect_round(jid), first[k]) for k, jid in enumerate(ids))
    for jid in ids:
        srv.release_job(jid)
    rate_sat = len(jobs) / t_sat
    log(f"[bench] c5 saturated: {len(jobs)} jobs in {t_sat:.2f}s ({rate_sat:.2f} jobs/s)")
    # Phase 3: Poisson arrivals at a fixed rate.
    lam = args.c5_load * rate_sat
    rng = random.Random(2005)
    ids = []
    t_start = time.perf_counter()
    t_next = t_start
    for k in range(len(jobs)):
        t_next += rng.expovariate(lam)
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        ids.append(submit(k))
    for jid in ids:
        srv.wait(jid, 600)
    t_run = time.perf_counter() - t_start
    lat, slow = [], []
    for k, jid in enumerate(ids):
        ts, ta, tf = srv.job_times(jid)
        lat.append(tf - ts)
        slow.append((tf - ts) / max(alone[k], 1e-9))
        exact = exact and np.array_equal(srv.first_detect_round(jid), first[k])
    cnt = srv.counters()
    srv.close()
    slowest = sorted(range(len(jobs)), key=lambda k: -lat[k])[:5]
    tail = [{"latency_ms": round(lat[k] * 1e3, 1), "unloaded_ms": round(alone[k] * 1e3, 1),
             "faults": F[k], "patterns": jobs[k]["num_patterns"],
             "gates": circuits[jobs[k]["circuit"]]["num_gates"]} for k in slowest]
    p99_alone = _pct(alone, 99)
    return {
        "metric": "p99 latency at fixed QPS", "value": _pct(lat, 99) * 1e3, "unit": "ms",
        "workload": f"c5: {len(circuits)} resident seeded circuits (10k-2M gates, "
                    f"{sum(sp['num_gates'] for sp in circuits)} in all), {len(jobs)} jobs (1-100% of faults, "
                    "1,024-32,768 patterns), fault dropping on, Poisson arrivals",
        "jobs": len(jobs), "qps": lam, "load": args.c5_load,
        "latency_ms": {"p50": _pct(lat, 50) * 1e3, "p99": _pct(lat, 99) * 1e3,
                       "unloaded_p50": _pct(alone, 50) * 1e3, "unloaded_p99": p99_alone * 1e3,
                       "p99_over_unloaded_p99": _pct(lat, 99) / max(p99_alone, 1e-9),
                       "slowdown_p50": _pct(slow, 50), "slowdown_p99": _pct(slow, 99)},
        "throughput": {"jobs_per_s": len(jobs) / t_run, "evals_per_s": evals / t_run,
                       "saturated_jobs_per_s": rate_sat, "saturated_evals_per_s": evals / t_sat},
        "bit_exact_alone_vs_mix": bool(exact),
        "slowest_jobs": tail,
        "server": cnt,
        "ms_per_job": 1e3 * t_run / len(jobs),
    }


def bench_serving(args) -> int:
    leg = serving_leg(args, args.c5_jobs)
    out = {
        "metric": leg["metric"], "value": leg["value"], "unit": leg["unit"],
        "n_gpus": 1, "steps": leg["jobs"], "warmup": 0, "ms_per_step": leg["ms_per_job"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": leg["workload"], "qps": leg["qps"], "load": leg["load"]},
    }
    out.update({k: leg[k] for k in ("latency_ms", "throughput", "bit_exact_alone_vs_mix", "slowest_jobs", "server")})
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--c5-jobs", type=int, default=64, help="jobs per phase of the serving run")
    ap.add_argument("--c5-load", type=float, default=0.5, help="arrival rate as a fraction of the saturated rate")
    ap.add_argument("--c5-growth", type=int, default=4, help="word growth per step of serving jobs")
    ap.add_argument("--mode", default="partitioned", choices=["plain", "partitioned"])
    ap.add_argument("--sweep-kernel", default="pairs", choices=["pairs", "nodewave"],
                    help="persistent sweep kernel of the partitioned engine (without --autotune)")
    ap.add_argument("--pair-threads", type=int, default=0, choices=[0, 256, 512],
                    help="workgroup size of the pair sweep (0: 512 where most tiles walk levels wider than 256 nodes, 256 "
                         "otherwise)")
    ap.add_argument("--pair-dense", type=int, default=-1,
                    help="pair sweep: passes whose nodes average this many active lanes go one node per wave "
                         "(0: never; -1: the engine default)")
    ap.add_argument("--pair-nopend", type=int, default=-1,
                    help="pair sweep without pending flags: 0 never, 1 dense steps, 2 every step (-1: the engine "
                         "default)")
    ap.add_argument("--pair-nopend-pct", type=int, default=-1,
                    help="with pair_nopend 1: flag-free after a step that staged more than this %% of walked nodes "
                         "(-1: the engine default)")
    ap.add_argument("--pair-fo-pre", type=int, default=-1, choices=[-1, 0, 1],
                    help="pair sweep with pending flags: load a staged node's first fanout in the scan (-1: the "
                         "engine default)")
    ap.add_argument("--max-step-words", type=int, default=-1,
                    help="with fault dropping: most words of one step (-1: the engine default)")
    ap.add_argument("--device-steps", type=int, default=-1, choices=[-1, 0, 1],
                    help="plan word steps on the device (1) or the host (0); -1: the library default")
    ap.add_argument("--first-words", type=int, default=0,
                    help="with fault dropping: words of the first step (0: the engine default)")
    ap.add_argument("--word-growth", type=int, default=0,
                    help="with fault dropping: factor of words per step (0: the engine default)")
    ap.add_argument("--good-form", type=int, default=-1, choices=[-1, 0, 1],
                    help="good machine: 0 one launch per level, 1 LDS-sized level windows (-1: the engine default, 1)")
    ap.add_argument("--win-levels", type=int, default=0, help="level windows: most levels per window (0: default)")
    ap.add_argument("--win-cap", type=int, default=0, help="level windows: nodes per block (0: from the LDS)")
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: the engine's autotuner tries one kernel/word-schedule candidate per run on the "
                         "first runs (give --warmup >= 8 to finish exploring) and then keeps the fastest")
    ap.add_argument("--c4-autotune", type=int, default=1,
                    help="batching leg: 1 lets the autotuner explore its candidates in the warmup runs and keep "
                         "the fastest")
    ap.add_argument("--c4-steps", type=int, default=3, help="timed batches of the batching leg")
    ap.add_argument("--c4-warmup", type=int, default=2,
                    help="untimed batches of the batching leg (with the autotuner: at least until every candidate "
                         "was measured)")
    ap.add_argument("--no-legs", action="store_true",
                    help="the 2M-gate line only, without the batching and serving legs")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of one-core CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe-random", type=int, default=0, metavar="MIB",
                    help="only run the random-read probe over a MIB MiB buffer and print its figures")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the known-size kernel for memory "
                         "counters) and print its figures")
    args = ap.parse_args()
    if args.probe_random:
        print(json.dumps(probe_random(args.probe_random << 20)), flush=True)
        return 0
    if args.copy_probe:
        print(json.dumps(copy_probe(args.copy_probe << 20)), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the fault set would shard "
              "across GPUs by fault, each GPU holding the circuit)", file=sys.stderr)
        return 2

    if args.config == "c4":
        return bench_batch(args)
    if args.config == "c5":
        return bench_serving(args)
    import lsim_amd
    from lsim_amd import workloads

    preset = workloads.describe(args.config)
    t0 = time.perf_counter()
    ckt, flst, ptn = workloads.texts(args.config)
    log(f"[bench] generated {args.config} in {time.perf_counter() - t0:.1f}s")
    sim = lsim_amd.Simulator(device=0)
    sim.set_option("drop", 1 if preset["drop"] else 0)
    sim.set_option("report_last_round", 0)
    sim.set_option("sweep_kernel", 0 if args.sweep_kernel == "pairs" else 1)
    sim.set_option("autotune", args.autotune)
    sim.set_option("pair_threads", args.pair_threads)
    if args.good_form >= 0:
        sim.set_option("good_form", args.good_form)
    if args.win_levels:
        sim.set_option("win_levels", args.win_levels)
    sim.set_option("win_cap", args.win_cap)
    if args.pair_dense >= 0:
        sim.set_option("pair_dense", args.pair_dense)
    if args.pair_nopend >= 0:
        sim.set_option("pair_nopend", args.pair_nopend)
    if args.pair_fo_pre >= 0:
        sim.set_option("pair_fo_pre", args.pair_fo_pre)
    if args.pair_nopend_pct >= 0:
        sim.set_option("pair_nopend_pct", args.pair_nopend_pct)
    if args.max_step_words >= 0:
        sim.set_option("max_step_words", args.max_step_words)
    if args.device_steps >= 0:
        sim.set_option("device_steps", args.device_steps)
    if args.first_words > 0:
        sim.set_option("first_words", args.first_words)
    if args.word_growth > 0:
        sim.set_option("word_growth", args.word_growth)
    t0 = time.perf_counter()
    sim.read_text(ckt, flst, ptn)
    log(f"[bench] read + upload in {time.perf_counter() - t0:.1f}s")
    mode = lsim_amd.Simulator.Mode.PLAIN if args.mode == "plain" else lsim_amd.Simulator.Mode.PARTITIONED
    for i in range(args.warmup):
        t0 = time.perf_counter()
        sim.run(mode)
        wst = sim.last_run_stats()
        log(f"[bench] warmup {i} {time.perf_counter() - t0:.3f}s config {_tuned(wst)} "
            f"({wst['tune_trials']} autotune candidates measured)")
    times = []
    agg = None
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        sim.run(mode)  # returns after the device work of the step has completed
        times.append(time.perf_counter() - t0)
        st = sim.last_run_stats()
        if agg is None:
            agg = {k: 0 for k in st}
        for k, v in st.items():
            agg[k] += v
        log(f"[bench] step {i} {times[-1]:.3f}s sweep {st['t_sweep_ms']:.1f}ms detect {st['t_detect_ms']:.1f}ms "
            f"good {st['t_good_ms']:.1f}ms launches {st['sweep_launches']} stems {st['stems']}")
    elapsed = time.perf_counter() - t_all
    F = sim.num_faults()
    n = sim.num_patterns()
    ms = 1000.0 * elapsed / args.steps
    value = F * n / (elapsed / args.steps)
    detected = int(sim.detected().sum())
    kernel = _kernel_name(args, st["tuned"])
    out = {
        "metric": "fault x pattern evals/sec",
        "value": value,
        "unit": "fault*pattern evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: seeded DAG {preset['num_gates']} gates / {preset['num_pis']} PI / "
                        f"{preset['num_pos']} PO / {preset['num_levels']} levels, {F} SA0/SA1 faults, "
                        f"{n} patterns, fault dropping {'on' if preset['drop'] else 'off'}",
            "engine": args.mode,
            "faults": F,
            "patterns": n,
            "detected": detected,
        },
        "roofline": _roofline(kernel, agg, args.config),
        "autotune": dict(_tuned(st), enabled=bool(args.autotune), candidates_measured=st["tune_trials"]),
        "breakdown_ms": {k: agg[k] / args.steps for k in ("t_good_ms", "t_sweep_ms", "t_detect_ms", "t_total_ms")},
    }
    sim.close()
    if not args.no_legs and args.mode == "partitioned":
        # The plain per-level form on a 1 % fault sample of the same workload
        # (its cost grows with the faults swept): the ratio the persistent
        # form reaches over it.
        t0 = time.perf_counter()
        try:
            out["plain_form"] = plain_form_sample(ckt, flst, ptn, stride=400)
        except Exception as e:  # reported, never fatal for the headline figure
            out["plain_form"] = {"error": f"{type(e).__name__}: {e}"}
        log(f"[bench] plain-form sample in {time.perf_counter() - t0:.1f}s")
    if not args.no_legs and args.config == "c3":
        # The two systems forms on their own workloads, in the same run.
        for name, fn in (("batching", lambda: batching_leg(args, args.c4_steps, args.c4_warmup)),
                         ("serving", lambda: serving_leg(args, args.c5_jobs))):
            t0 = time.perf_counter()
            try:
                out[name] = fn()
            except Exception as e:  # reported, never fatal for the headline figure
                out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] {name} leg in {time.perf_counter() - t0:.1f}s")
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(ckt, flst, ptn, preset, args.cpu_budget)
        except Exception as e:  # reported, never fatal for the GPU figure
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
