"""lsim_amd: MI355X-native stuck-at fault simulator.

Python view of the C ABI in include/lsim.h. `Simulator` mirrors the reference's
fsim::Simulator (fsim/simulation.hpp:39-188): read the .ckt/.flst/.ptn inputs,
run, then read the same accessors (found_fault_to_pattern, good values, order
map, levels, faults, patterns). All simulation runs in the HIP library
liblsim.so; there is no CPU execution path, and a missing library raises.
"""
from __future__ import annotations

import ctypes
import enum
import os
from typing import Optional, Tuple, Union

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblsim.so")

LSIM_OK = 0
LSIM_E_ARG = -1
LSIM_E_INPUT = -2
LSIM_E_IO = -3
LSIM_E_STATE = -4
LSIM_E_DEVICE = -5
LSIM_E_NOMEM = -6

_lib: Optional[ctypes.CDLL] = None


class LsimError(RuntimeError):
    """Error returned by the native library (code is an LSIM_E_* value)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class RunStats(ctypes.Structure):
    _fields_ = [
        ("t_total_ms", ctypes.c_double),
        ("t_good_ms", ctypes.c_double),
        ("t_sweep_ms", ctypes.c_double),
        ("t_detect_ms", ctypes.c_double),
        ("sweep_launches", ctypes.c_int64),
        ("sweep_tiles", ctypes.c_int64),
        ("sweep_node_steps", ctypes.c_int64),
        ("sweep_bytes", ctypes.c_int64),
        ("good_bytes", ctypes.c_int64),
        ("live_faults", ctypes.c_int64),
        ("stems", ctypes.c_int64),
        ("steps", ctypes.c_int64),
        ("slots", ctypes.c_int64),
        ("live_nodes", ctypes.c_int64),
        ("sweep_reruns", ctypes.c_int64),
        ("tune_trials", ctypes.c_int64),
        ("tuned", ctypes.c_int64),
        ("engine", ctypes.c_int64),
        ("tune_candidates", ctypes.c_int64),
        ("good_replayed", ctypes.c_int64),
        ("device_allocs", ctypes.c_int64),
        ("steps_replayed", ctypes.c_int64),
    ]


class GenSpec(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("num_pis", ctypes.c_int64),
        ("num_pos", ctypes.c_int64),
        ("num_gates", ctypes.c_int64),
        ("num_levels", ctypes.c_int64),
        ("fanout_cap", ctypes.c_int32),
        ("power_exp", ctypes.c_double),
        ("locality", ctypes.c_double),
        ("window", ctypes.c_int32),
        ("unary_frac", ctypes.c_double),
        ("num_patterns", ctypes.c_int64),
        ("fault_fraction", ctypes.c_double),
    ]


def lib() -> ctypes.CDLL:
    """Load liblsim.so (built by lsim_amd/build.py); raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u64p = ctypes.POINTER(ctypes.c_uint64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    sz = ctypes.c_size_t
    sig = {
        "lsim_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "lsim_destroy": (None, [P]),
        "lsim_last_error": (ctypes.c_char_p, [P]),
        "lsim_read_files": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
        "lsim_read_buffers": (ctypes.c_int, [P, ctypes.c_char_p, sz, ctypes.c_char_p, sz, ctypes.c_char_p, sz]),
        "lsim_set_option": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_int64]),
        "lsim_run": (ctypes.c_int, [P, ctypes.c_int, sz]),
        "lsim_run_noalloc": (ctypes.c_int, [P, ctypes.c_int, sz]),
        "lsim_num_pis": (ctypes.c_int64, [P]),
        "lsim_num_pos": (ctypes.c_int64, [P]),
        "lsim_num_gates": (ctypes.c_int64, [P]),
        "lsim_num_wires": (ctypes.c_int64, [P]),
        "lsim_num_rounds": (ctypes.c_int64, [P]),
        "lsim_num_faults": (ctypes.c_int64, [P]),
        "lsim_num_levels": (ctypes.c_int64, [P]),
        "lsim_num_patterns": (ctypes.c_int64, [P]),
        "lsim_gate_idx_to_tpg_order_idx": (ctypes.c_int, [P, u64p, sz]),
        "lsim_num_gates_per_level": (ctypes.c_int, [P, u64p, sz]),
        "lsim_levels": (ctypes.c_int, [P, u64p, sz]),
        "lsim_visualization_graph": (ctypes.c_int, [P, ctypes.c_char_p]),
        "lsim_level_diagnostics": (ctypes.c_int, [P, ctypes.c_int, i64p, ctypes.POINTER(ctypes.c_double), i64p,
                                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                  sz]),
        "lsim_visualization_text": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]),
        "lsim_faults": (ctypes.c_int, [P, u64p, u64p, sz]),
        "lsim_patterns": (ctypes.c_int, [P, u64p, sz]),
        "lsim_good_results": (ctypes.c_int, [P, u64p, sz, u64p, sz, u64p, sz]),
        "lsim_found_fault_to_pattern": (ctypes.c_int, [P, u64p, sz]),
        "lsim_detected": (ctypes.c_int, [P, u8p, sz]),
        "lsim_first_detect_round": (ctypes.c_int, [P, i64p, sz]),
        "lsim_detect_matrix": (ctypes.c_int, [P, u8p, sz]),
        "lsim_last_run_stats": (ctypes.c_int, [P, ctypes.POINTER(RunStats)]),
        "lsim_generate": (
            ctypes.c_int,
            [
                ctypes.POINTER(GenSpec),
                ctypes.POINTER(ctypes.c_void_p),
                ctypes.POINTER(sz),
                ctypes.POINTER(ctypes.c_void_p),
                ctypes.POINTER(sz),
                ctypes.POINTER(ctypes.c_void_p),
                ctypes.POINTER(sz),
            ],
        ),
        "lsim_free_text": (None, [ctypes.c_void_p]),
        "lsim_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "lsim_batch_destroy": (None, [P]),
        "lsim_batch_last_error": (ctypes.c_char_p, [P]),
        "lsim_batch_set_option": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_int64]),
        "lsim_batch_add_circuit": (ctypes.c_int, [P, ctypes.c_char_p, sz, i64p]),
        "lsim_batch_add_job": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_char_p, sz, ctypes.c_char_p, sz,
                                              ctypes.c_int, i64p]),
        "lsim_batch_run": (ctypes.c_int, [P]),
        "lsim_batch_job_faults": (ctypes.c_int64, [P, ctypes.c_int64]),
        "lsim_batch_job_rounds": (ctypes.c_int64, [P, ctypes.c_int64]),
        "lsim_batch_first_detect_round": (ctypes.c_int, [P, ctypes.c_int64, i64p, sz]),
        "lsim_batch_detect_matrix": (ctypes.c_int, [P, ctypes.c_int64, u8p, sz]),
        "lsim_batch_found_fault_to_pattern": (ctypes.c_int, [P, ctypes.c_int64, u64p, sz]),
        "lsim_batch_last_run_stats": (ctypes.c_int, [P, ctypes.POINTER(RunStats)]),
        "lsim_batch_release_job": (ctypes.c_int, [P, ctypes.c_int64]),
        "lsim_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(P)]),
        "lsim_server_destroy": (None, [P]),
        "lsim_server_last_error": (ctypes.c_char_p, [P]),
        "lsim_server_add_circuit": (ctypes.c_int, [P, ctypes.c_char_p, sz, i64p]),
        "lsim_server_submit": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_char_p, sz, ctypes.c_char_p, sz,
                                              ctypes.c_int, i64p]),
        "lsim_server_wait": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_double]),
        "lsim_server_first_detect_round": (ctypes.c_int, [P, ctypes.c_int64, i64p, sz]),
        "lsim_server_job_faults": (ctypes.c_int64, [P, ctypes.c_int64]),
        "lsim_server_job_times": (ctypes.c_int, [P, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]),
        "lsim_server_release_job": (ctypes.c_int, [P, ctypes.c_int64]),
        "lsim_server_counters": (ctypes.c_int, [P, i64p]),
        "lsim_server_set_option": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_int64]),
        "lsim_break_loops": (ctypes.c_int, [ctypes.c_char_p, sz, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(sz), i64p]),
        "lsim_generate_faults": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                ctypes.c_double, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]),
        "lsim_generate_patterns": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                                  ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]),
        "lsim_convert_timing_dump": (ctypes.c_int, [ctypes.c_char_p, sz, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                                    ctypes.POINTER(sz), ctypes.POINTER(IngestStats)]),
        "lsim_produce_faults": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]),
        "lsim_pack_patterns": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]),
        "lsim_produce_patterns": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(sz)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols() -> list:
    """Names of the C ABI entry points declared in include/lsim.h."""
    import re

    hdr = os.path.join(os.path.dirname(_HERE), "include", "lsim.h")
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"LSIM_API[^;(]*?\b(lsim_\w+)\s*\(", text)))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _as_bytes(x: Union[str, bytes]) -> bytes:
    return x.encode() if isinstance(x, str) else bytes(x)


class Simulator:
    """Mirror of fsim::Simulator backed by the GPU engine.

    device=-1 gives a host-only handle: read() and the structural accessors
    work, run() raises LsimError(LSIM_E_DEVICE)."""

    class Mode(enum.IntEnum):
        # reference names (fsim/simulation.hpp:45): all three run the
        # persistent engine; the two engine names choose an engine explicitly
        SEQUENTIAL = 0
        PARALLEL_TF = 1
        PARALLEL_OMP = 2
        PLAIN = 16
        PARTITIONED = 17

    def __init__(self, device: int = 0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.lsim_create(device, ctypes.byref(h))
        if rc != LSIM_OK:
            raise LsimError(rc, self._L.lsim_last_error(None).decode())
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.lsim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int) -> None:
        if rc != LSIM_OK:
            raise LsimError(rc, self._L.lsim_last_error(self._h).decode())

    # -- input (fsim/simulation.cpp:12-63)
    def read(self, ckt_path: str, flst_path: str, ptn_path: str) -> None:
        self._check(self._L.lsim_read_files(self._h, ckt_path.encode(), flst_path.encode(), ptn_path.encode()))

    def read_text(self, ckt, flst, ptn) -> None:
        c, f, p = _as_bytes(ckt), _as_bytes(flst), _as_bytes(ptn)
        self._check(self._L.lsim_read_buffers(self._h, c, len(c), f, len(f), p, len(p)))

    def set_option(self, key: str, value: int) -> None:
        self._check(self._L.lsim_set_option(self._h, key.encode(), int(value)))

    # -- run (fsim/simulation.cpp:65-92)
    def run(self, mode: int = 0, num_threads: int = 1) -> None:
        self._check(self._L.lsim_run(self._h, int(mode), int(num_threads)))

    def run_noalloc(self, mode: int = 0, num_threads: int = 1) -> None:
        """run() with every device buffer carved from the library's fixed
        per-device scratch (lsim_run_noalloc): no device allocation."""
        self._check(self._L.lsim_run_noalloc(self._h, int(mode), int(num_threads)))

    # -- accessors (fsim/simulation.hpp:58-89)
    def num_PIs(self) -> int:
        return self._L.lsim_num_pis(self._h)

    def num_POs(self) -> int:
        return self._L.lsim_num_pos(self._h)

    def num_gates(self) -> int:
        return self._L.lsim_num_gates(self._h)

    def num_wires(self) -> int:
        return self._L.lsim_num_wires(self._h)

    def num_rounds(self) -> int:
        return self._L.lsim_num_rounds(self._h)

    def num_faults(self) -> int:
        return self._L.lsim_num_faults(self._h)

    def num_patterns(self) -> int:
        return self._L.lsim_num_patterns(self._h)

    def _u64(self, fn, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.uint64)
        self._check(fn(self._h, _ptr(out, ctypes.c_uint64), n))
        return out

    def gate_idx_to_tpg_order_idx(self) -> np.ndarray:
        n = self.num_PIs() + self.num_POs() + self.num_gates()
        return self._u64(self._L.lsim_gate_idx_to_tpg_order_idx, n)

    def num_gates_per_level(self) -> np.ndarray:
        return self._u64(self._L.lsim_num_gates_per_level, self._L.lsim_num_levels(self._h))

    def levels(self) -> np.ndarray:
        """Levels of pi(i), gate(i), po(i) concatenated (topological order)."""
        n = self.num_PIs() + self.num_POs() + self.num_gates()
        return self._u64(self._L.lsim_levels, n)

    def visualization_graph(self, path: str) -> None:
        """Graphviz digraph of the circuit written to `path`
        (fsim/simulation.cpp:1157-1225)."""
        self._check(self._L.lsim_visualization_graph(self._h, path.encode()))

    def level_diagnostics(self, wave: int = 64) -> dict:
        """Per level: node count, mean and largest fanin level distance, and per
        group of `wave` consecutive nodes the fanin position range and the
        number of distinct fanins (means over the level's groups)."""
        L = int(self.levels().max()) + 1 if self.num_PIs() + self.num_gates() + self.num_POs() else 0
        nodes = np.zeros(L, dtype=np.int64)
        span = np.zeros(L, dtype=np.int64)
        dist, rng_, uniq = (np.zeros(L, dtype=np.float64) for _ in range(3))
        self._check(self._L.lsim_level_diagnostics(self._h, wave, _ptr(nodes, ctypes.c_int64), _ptr(dist, ctypes.c_double),
                                                   _ptr(span, ctypes.c_int64), _ptr(rng_, ctypes.c_double),
                                                   _ptr(uniq, ctypes.c_double), L))
        return {"nodes": nodes, "avg_dist": dist, "max_span": span, "avg_range": rng_, "avg_unique": uniq}

    def visualization_text(self) -> str:
        b, n = ctypes.c_void_p(), ctypes.c_size_t()
        self._check(self._L.lsim_visualization_text(self._h, ctypes.byref(b), ctypes.byref(n)))
        out = ctypes.string_at(b.value, n.value)
        self._L.lsim_free_text(b)
        return out.decode()

    def faults(self) -> Tuple[np.ndarray, np.ndarray]:
        n = self.num_faults()
        site = np.zeros(n, dtype=np.uint64)
        val = np.zeros(n, dtype=np.uint64)
        self._check(self._L.lsim_faults(self._h, _ptr(site, ctypes.c_uint64), _ptr(val, ctypes.c_uint64), n))
        return site, val

    def patterns(self) -> np.ndarray:
        R, P = self.num_rounds(), self.num_PIs()
        return self._u64(self._L.lsim_patterns, R * P).reshape(R, P)

    def good_results(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        P, G, O = self.num_PIs(), self.num_gates(), self.num_POs()
        pi, gate, po = (np.zeros(n, dtype=np.uint64) for n in (P, G, O))
        u = ctypes.c_uint64
        self._check(self._L.lsim_good_results(self._h, _ptr(pi, u), P, _ptr(gate, u), G, _ptr(po, u), O))
        return pi, gate, po

    def g_pi_results(self) -> np.ndarray:
        return self.good_results()[0]

    def g_gate_results(self) -> np.ndarray:
        return self.good_results()[1]

    def g_po_results(self) -> np.ndarray:
        return self.good_results()[2]

    def found_fault_to_pattern(self) -> np.ndarray:
        return self._u64(self._L.lsim_found_fault_to_pattern, 2 * self.num_faults())

    def detected(self) -> np.ndarray:
        out = np.zeros(self.num_faults(), dtype=np.uint8)
        self._check(self._L.lsim_detected(self._h, _ptr(out, ctypes.c_uint8), out.size))
        return out

    def first_detect_round(self) -> np.ndarray:
        out = np.zeros(self.num_faults(), dtype=np.int64)
        self._check(self._L.lsim_first_detect_round(self._h, _ptr(out, ctypes.c_int64), out.size))
        return out

    def detect_matrix(self) -> np.ndarray:
        F, R = self.num_faults(), self.num_rounds()
        out = np.zeros(F * R, dtype=np.uint8)
        self._check(self._L.lsim_detect_matrix(self._h, _ptr(out, ctypes.c_uint8), out.size))
        return out.reshape(F, R)

    def last_run_stats(self) -> dict:
        st = RunStats()
        self._check(self._L.lsim_last_run_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in RunStats._fields_}


def generate(
    seed: int,
    num_pis: int,
    num_pos: int,
    num_gates: int,
    num_levels: int,
    num_patterns: int,
    fanout_cap: int = 0,
    power_exp: float = 2.2,
    locality: float = 0.7,
    window: int = 3,
    unary_frac: float = 0.1,
    fault_fraction: float = 1.0,
) -> Tuple[bytes, bytes, bytes]:
    """Seeded synthetic (.ckt, .flst, .ptn) texts, produced by the library."""
    L = lib()
    spec = GenSpec(seed, num_pis, num_pos, num_gates, num_levels, fanout_cap, power_exp, locality, window,
                   unary_frac, num_patterns, fault_fraction)
    bufs = [ctypes.c_void_p() for _ in range(3)]
    lens = [ctypes.c_size_t() for _ in range(3)]
    rc = L.lsim_generate(ctypes.byref(spec), ctypes.byref(bufs[0]), ctypes.byref(lens[0]), ctypes.byref(bufs[1]),
                         ctypes.byref(lens[1]), ctypes.byref(bufs[2]), ctypes.byref(lens[2]))
    if rc != LSIM_OK:
        raise LsimError(rc, L.lsim_last_error(None).decode() or "generator rejected the spec")
    out = []
    for b, n in zip(bufs, lens):
        out.append(ctypes.string_at(b.value, n.value))
        L.lsim_free_text(b)
    return out[0], out[1], out[2]


def _take_text(rc: int, buf: ctypes.c_void_p, n: ctypes.c_size_t) -> bytes:
    L = lib()
    if rc != LSIM_OK:
        raise LsimError(rc, L.lsim_last_error(None).decode() or "generator rejected the spec")
    out = ctypes.string_at(buf.value, n.value)
    L.lsim_free_text(buf)
    return out


def break_loops(ckt, mode: str = "bfs"):
    """(.ckt text, dropped entries): wires into nodes already taken by a BFS or
    DFS traversal from the PIs are dropped (fsim/parser.py:163-211)."""
    c = _as_bytes(ckt)
    b, n = ctypes.c_void_p(), ctypes.c_size_t()
    removed = ctypes.c_int64()
    m = {"bfs": 0, "dfs": 1}[mode]
    rc = lib().lsim_break_loops(c, len(c), m, ctypes.byref(b), ctypes.byref(n), ctypes.byref(removed))
    return _take_text(rc, b, n), removed.value


def generate_faults(seed: int, num_pis: int, num_pos: int, num_gates: int, fraction: float = 1.0) -> bytes:
    """Seeded .flst text: SA0/SA1 over the inner gates, thinned to `fraction`."""
    b, n = ctypes.c_void_p(), ctypes.c_size_t()
    rc = lib().lsim_generate_faults(seed, num_pis, num_pos, num_gates, fraction, ctypes.byref(b), ctypes.byref(n))
    return _take_text(rc, b, n)


def generate_patterns(seed: int, num_pis: int, num_patterns: int) -> bytes:
    """Seeded .ptn text with `num_patterns` random patterns."""
    b, n = ctypes.c_void_p(), ctypes.c_size_t()
    rc = lib().lsim_generate_patterns(seed, num_pis, num_patterns, ctypes.byref(b), ctypes.byref(n))
    return _take_text(rc, b, n)


def produce_faults(seed: int, num_nodes: int, num_faults: int, num_pis: int, num_pos: int) -> bytes:
    """.flst of distinct random (wire, value) pairs, PO ids excluded
    (fsim/generate_fault_patterns.py:7-30)."""
    b, n = ctypes.c_void_p(), ctypes.c_size_t()
    rc = lib().lsim_produce_faults(seed, num_nodes, num_faults, num_pis, num_pos, ctypes.byref(b), ctypes.byref(n))
    return _take_text(rc, b, n)


def pack_patterns(patterns, bits_per_round: int = 64) -> bytes:
    """.ptn of explicit patterns (array [n][P] of 0/1) in the layout of
    fsim/generate_fault_patterns.py:56-94 (first pattern most significant)."""
    a = np.ascontiguousarray(np.asarray(patterns, dtype=np.uint8))
    if a.ndim != 2:
        raise ValueError("patterns must be a 2-D array [n][P]")
    b, n = ctypes.c_void_p(), ctypes.c_size_t()
    rc = lib().lsim_pack_patterns(a.shape[1], a.shape[0], bits_per_round, a.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.byref(b), ctypes.byref(n))
    return _take_text(rc, b, n)


def produce_patterns(seed: int, num_pis: int, num_patterns: int, bits_per_round: int = 64) -> bytes:
    """.ptn of distinct random patterns packed as pack_patterns does."""
    b, n = ctypes.c_void_p(), ctypes.c_size_t()
    rc = lib().lsim_produce_patterns(seed, num_pis, num_patterns, bits_per_round, ctypes.byref(b), ctypes.byref(n))
    return _take_text(rc, b, n)


class IngestStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("pis", "pos", "cells", "wires", "connections", "records", "removed",
                                              "unknown_cells", "unknown_pins", "unreached")]


def convert_timing_dump(dump, mode: str = "bfs"):
    """(.ckt text, stats) of an OpenTimer timing-graph dump (fsim/parser.py):
    cell and pin mapping, node ids, and loop removal by a BFS or DFS traversal
    from the PIs."""
    d = _as_bytes(dump)
    if mode not in ("bfs", "dfs"):
        raise ValueError("mode must be 'bfs' or 'dfs'")
    b, n, st = ctypes.c_void_p(), ctypes.c_size_t(), IngestStats()
    rc = lib().lsim_convert_timing_dump(d, len(d), 0 if mode == "bfs" else 1, ctypes.byref(b), ctypes.byref(n),
                                        ctypes.byref(st))
    text = _take_text(rc, b, n)
    return text, {name: getattr(st, name) for name, _ in IngestStats._fields_}


class Batch:
    """Many circuits with their own fault lists and pattern blocks, run
    together (one persistent sweep launch per word step over all jobs)."""

    def __init__(self, device: int = 0, **options):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.lsim_batch_create(device, ctypes.byref(h))
        if rc != LSIM_OK:
            raise LsimError(rc, self._L.lsim_batch_last_error(None).decode())
        self._h = h
        for k, v in options.items():
            self.set_option(k, v)

    def _check(self, rc: int) -> None:
        if rc != LSIM_OK:
            raise LsimError(rc, self._L.lsim_batch_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.lsim_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, key: str, value: int) -> None:
        self._check(self._L.lsim_batch_set_option(self._h, key.encode(), int(value)))

    def add_circuit(self, ckt) -> int:
        c = _as_bytes(ckt)
        cid = ctypes.c_int64()
        self._check(self._L.lsim_batch_add_circuit(self._h, c, len(c), ctypes.byref(cid)))
        return cid.value

    def add_job(self, circuit: int, flst, ptn, drop: bool = False) -> int:
        f, p = _as_bytes(flst), _as_bytes(ptn)
        jid = ctypes.c_int64()
        self._check(self._L.lsim_batch_add_job(self._h, circuit, f, len(f), p, len(p), int(drop), ctypes.byref(jid)))
        return jid.value

    def run(self) -> None:
        self._check(self._L.lsim_batch_run(self._h))

    def first_detect_round(self, job: int) -> np.ndarray:
        n = self._L.lsim_batch_job_faults(self._h, job)
        out = np.zeros(n, dtype=np.int64)
        self._check(self._L.lsim_batch_first_detect_round(self._h, job, _ptr(out, ctypes.c_int64), n))
        return out

    def detect_matrix(self, job: int) -> np.ndarray:
        F = self._L.lsim_batch_job_faults(self._h, job)
        R = self._L.lsim_batch_job_rounds(self._h, job)
        out = np.zeros(F * R, dtype=np.uint8)
        self._check(self._L.lsim_batch_detect_matrix(self._h, job, _ptr(out, ctypes.c_uint8), out.size))
        return out.reshape(F, R)

    def found_fault_to_pattern(self, job: int) -> np.ndarray:
        F = self._L.lsim_batch_job_faults(self._h, job)
        out = np.zeros(2 * F, dtype=np.uint64)
        self._check(self._L.lsim_batch_found_fault_to_pattern(self._h, job, _ptr(out, ctypes.c_uint64), out.size))
        return out

    def last_run_stats(self) -> dict:
        st = RunStats()
        self._check(self._L.lsim_batch_last_run_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in RunStats._fields_}

    def release_job(self, job: int) -> None:
        self._check(self._L.lsim_batch_release_job(self._h, job))


class Server:
    """Resident circuits and two lanes, each an engine with its own stream and
    worker thread: small jobs are batched continuously (a job joins at the
    next word step), large ones (fault x pattern evaluations from the
    `big_evals` option, default 4e9) run one at a time in arrival order.
    Admission in arrival order by HBM footprint."""

    def __init__(self, device: int = 0, job_budget_mb: int = 64 * 1024, scratch_mb: int = 0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.lsim_server_create(device, job_budget_mb, scratch_mb, ctypes.byref(h))
        if rc != LSIM_OK:
            raise LsimError(rc, self._L.lsim_server_last_error(None).decode())
        self._h = h

    def _check(self, rc: int) -> None:
        if rc != LSIM_OK:
            raise LsimError(rc, self._L.lsim_server_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.lsim_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_circuit(self, ckt) -> int:
        c = _as_bytes(ckt)
        cid = ctypes.c_int64()
        self._check(self._L.lsim_server_add_circuit(self._h, c, len(c), ctypes.byref(cid)))
        return cid.value

    def submit(self, circuit: int, flst, ptn, drop: bool = True) -> int:
        f, p = _as_bytes(flst), _as_bytes(ptn)
        jid = ctypes.c_int64()
        self._check(self._L.lsim_server_submit(self._h, circuit, f, len(f), p, len(p), int(drop), ctypes.byref(jid)))
        return jid.value

    def wait(self, job: int, timeout_s: float = -1.0) -> None:
        self._check(self._L.lsim_server_wait(self._h, job, float(timeout_s)))

    def first_detect_round(self, job: int) -> np.ndarray:
        n = self._L.lsim_server_job_faults(self._h, job)
        out = np.zeros(n, dtype=np.int64)
        self._check(self._L.lsim_server_first_detect_round(self._h, job, _ptr(out, ctypes.c_int64), n))
        return out

    def job_times(self, job: int):
        t = (ctypes.c_double * 3)()
        self._check(self._L.lsim_server_job_times(self._h, job, t))
        return t[0], t[1], t[2]

    def release_job(self, job: int) -> None:
        self._check(self._L.lsim_server_release_job(self._h, job))

    def set_option(self, key: str, value: int) -> None:
        self._check(self._L.lsim_server_set_option(self._h, key.encode(), int(value)))

    def counters(self) -> dict:
        c = (ctypes.c_int64 * 4)()
        self._check(self._L.lsim_server_counters(self._h, c))
        return {"iterations": c[0], "finished": c[1], "failed": c[2], "peak_admitted": c[3]}
