"""Seeded synthetic workloads of the benchmark configurations.

Each preset fixes the generator parameters (seed, counts, levels, fanout law)
of one configuration; `texts(name)` returns its (.ckt, .flst, .ptn) texts
produced by the library's generator (lsim_generate).
"""
from __future__ import annotations

from typing import Dict, Tuple

from . import generate

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
