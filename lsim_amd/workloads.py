"""Seeded synthetic workloads of the benchmark configurations.

Each preset fixes the generator parameters (seed, counts, levels, fanout law)
of one configuration; `texts(name)` returns its (.ckt, .flst, .ptn) texts
produced by the library's generator (lsim_generate).
"""
from __future__ import annotations

import math
import random
from typing import Dict, List, Tuple

from . import generate, generate_faults, generate_patterns

PRESETS: Dict[str, dict] = {
    # 5k gates / 64 PI / 32 PO / 40 levels / 10k faults / 1,024 patterns
    "c1": dict(seed=1001, num_pis=64, num_pos=32, num_gates=5000, num_levels=40, num_patterns=1024,
               fanout_cap=0, drop=False),
    # 100k gates / 128 PI / 64 PO / 120 levels / 200k faults / 8,192 patterns, power-law fanout cap 200
    "c2": dict(seed=1002, num_pis=128, num_pos=64, num_gates=100000, num_levels=120, num_patterns=8192,
               fanout_cap=200, drop=False),
    # 2M gates / 512 PI / 256 PO / 1,500 levels / 4M faults / 32,768 patterns, cap 500, fault dropping
    "c3": dict(seed=1003, num_pis=512, num_pos=256, num_gates=2000000, num_levels=1500, num_patterns=32768,
               fanout_cap=500, drop=True),
}


def describe(name: str) -> dict:
    p = dict(PRESETS[name])
    p["num_faults"] = 2 * p["num_gates"]
    return p


def texts(name: str) -> Tuple[bytes, bytes, bytes]:
    p = dict(PRESETS[name])
    p.pop("drop")
    return generate(**p)


def c4_specs(count: int = 256) -> List[dict]:
    """Batching workload: `count` independent circuits, gate counts
    log-uniform in [1k, 200k], levels ~ 12 log2(gates), 256 PI / 128 PO,
    power-law fanout (cap 200), full SA0/SA1 fault list, 2,048 to 16,384
    patterns (uniform over powers of two). Circuit i uses seed 1004*1000+i."""
    out = []
    for i in range(count):
        seed = 1004 * 1000 + i
        r = random.Random(seed)
        G = int(round(math.exp(r.uniform(math.log(1e3), math.log(2e5)))))
        L = max(2, int(round(12 * math.log2(G))))
        n = r.choice([2048, 4096, 8192, 16384])
        out.append(dict(seed=seed, num_pis=256, num_pos=128, num_gates=G, num_levels=L, num_patterns=n,
                        fanout_cap=200))
    return out


def c5_circuit_specs(count: int = 64) -> List[dict]:
    """Serving pool: `count` circuits, gate counts log-uniform in [10k, 2M],
    levels ~ 12 log2(gates), 512 PI / 256 PO, power-law fanout (cap 500).
    Circuit i uses seed 1005*1000+i."""
    out = []
    for i in range(count):
        seed = 1005 * 1000 + i
        r = random.Random(seed)
        G = int(round(math.exp(r.uniform(math.log(1e4), math.log(2e6)))))
        L = max(2, int(round(12 * math.log2(G))))
        out.append(dict(seed=seed, num_pis=512, num_pos=256, num_gates=G, num_levels=L, num_patterns=64,
                        fanout_cap=500))
    return out


def c5_jobs(circuits: List[dict], count: int, seed: int = 2005) -> List[dict]:
    """Job stream over the serving pool: circuit uniform, fault fraction
    log-uniform in [0.01, 1], pattern count log-uniform integer in
    [1,024, 32,768]; each job's faults and patterns have their own seed."""
    r = random.Random(seed)
    jobs = []
    for k in range(count):
        c = r.randrange(len(circuits))
        frac = math.exp(r.uniform(math.log(0.01), 0.0))
        n = int(round(math.exp(r.uniform(math.log(1024), math.log(32768)))))
        jobs.append(dict(circuit=c, fraction=frac, num_patterns=n, seed=seed * 1000003 + k))
    return jobs


def job_texts(circuit_spec: dict, job: dict) -> Tuple[bytes, bytes]:
    """(.flst, .ptn) of one serving job against its circuit."""
    c = circuit_spec
    flst = generate_faults(job["seed"], c["num_pis"], c["num_pos"], c["num_gates"], job["fraction"])
    ptn = generate_patterns(job["seed"], c["num_pis"], job["num_patterns"])
    return flst, ptn


def circuit_text(spec: dict) -> bytes:
    """.ckt text of one circuit spec (c4_specs / c5_circuit_specs)."""
    ckt, _, _ = generate(**dict(spec, num_patterns=64))
    return ckt


def batch_job_texts(spec: dict) -> Tuple[bytes, bytes]:
    """Full fault list and patterns of one batching circuit."""
    flst = generate_faults(spec["seed"], spec["num_pis"], spec["num_pos"], spec["num_gates"], 1.0)
    ptn = generate_patterns(spec["seed"], spec["num_pis"], spec["num_patterns"])
    return flst, ptn
