"""CPU oracle for the fault-simulation path (TEST INFRASTRUCTURE ONLY).

Python binding of oracle/liboracle.so, the plain-C restatement of the
reference simulator fsim::Simulator (fsim/simulation.{hpp,cpp}), and a runner
for oracle/_ref/fsim_ref, the reference itself compiled from its sources.
Only the tests, the smoke check and the CPU-baseline leg of bench.py use this
module, always as the checker; the product library never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "fsim_ref")

NUM_PIS, NUM_POS, NUM_GATES, NUM_WIRES, NUM_ROUNDS, NUM_FAULTS, NUM_NODES, NUM_LEVELS, NUM_PATTERNS = range(9)

_lib = None


def build(reference: Optional[str] = None) -> None:
    """Compile liboracle.so; with `reference` (a checkout of the reference
    sources) also compile _ref/fsim_ref from them."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if reference and os.path.isdir(os.path.join(reference, "fsim")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={reference}"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, u64p, i64p, u8p = (ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint8))
        L.oracle_read.restype = vp
        L.oracle_read.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p, sz, ctypes.c_char_p, sz, ctypes.c_char_p, sz]
        L.oracle_free.argtypes = [vp]
        L.oracle_count.restype = ctypes.c_int64
        L.oracle_count.argtypes = [vp, ctypes.c_int]
        for name in ("oracle_order_map", "oracle_gates_per_level", "oracle_levels", "oracle_patterns"):
            getattr(L, name).argtypes = [vp, u64p]
        L.oracle_faults.argtypes = [vp, u64p, u64p]
        L.oracle_run_literal.restype = ctypes.c_int
        L.oracle_run_literal.argtypes = [vp, ctypes.c_int, i64p, ctypes.c_int64, u64p, u8p, u64p, u64p, u64p]
        L.oracle_run_event.restype = ctypes.c_int
        L.oracle_run_event.argtypes = [vp, ctypes.c_int, i64p, ctypes.c_int64, u8p]
        L.oracle_run_event_range.restype = ctypes.c_int
        L.oracle_run_event_range.argtypes = [vp, ctypes.c_int, i64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.c_int64, ctypes.c_int, u8p]
        L.oracle_good_round.restype = ctypes.c_int
        L.oracle_good_round.argtypes = [vp, ctypes.c_int64, u64p, u64p, u64p]
        _lib = L
    return _lib


def _b(x) -> bytes:
    return x.encode() if isinstance(x, str) else bytes(x)


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class OracleInputError(ValueError):
    pass


class Oracle:
    """The reference's observable results for one (.ckt, .flst, .ptn) set."""

    def __init__(self, ckt, flst, ptn):
        L = lib()
        c, f, p = _b(ckt), _b(flst), _b(ptn)
        err = ctypes.create_string_buffer(256)
        self._h = L.oracle_read(c, len(c), f, len(f), p, len(p), err, 256)
        if not self._h:
            raise OracleInputError(err.value.decode())
        self._L = L

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.oracle_free(self._h)
            self._h = None

    def count(self, which: int) -> int:
        return self._L.oracle_count(self._h, which)

    @property
    def P(self):
        return self.count(NUM_PIS)

    @property
    def O(self):
        return self.count(NUM_POS)

    @property
    def G(self):
        return self.count(NUM_GATES)

    @property
    def R(self):
        return self.count(NUM_ROUNDS)

    @property
    def F(self):
        return self.count(NUM_FAULTS)

    @property
    def N(self):
        return self.count(NUM_NODES)

    def _arr(self, fn, n):
        out = np.zeros(n, dtype=np.uint64)
        fn(self._h, _p(out, ctypes.c_uint64))
        return out

    def order_map(self):
        return self._arr(self._L.oracle_order_map, self.N)

    def gates_per_level(self):
        return self._arr(self._L.oracle_gates_per_level, self.count(NUM_LEVELS))

    def levels(self):
        return self._arr(self._L.oracle_levels, self.N)

    def patterns(self):
        return self._arr(self._L.oracle_patterns, self.R * self.P).reshape(self.R, self.P)

    def faults(self):
        site = np.zeros(self.F, dtype=np.uint64)
        val = np.zeros(self.F, dtype=np.uint64)
        self._L.oracle_faults(self._h, _p(site, ctypes.c_uint64), _p(val, ctypes.c_uint64))
        return site, val

    def run_literal(self, threads: int = 0, faults: Optional[Sequence[int]] = None):
        """Reference algorithm. Returns dict(table, det[F,R] (uint8), g_pi, g_gate, g_po)."""
        F, R = self.F, self.R
        table = np.zeros(2 * F, dtype=np.uint64)
        det = np.zeros(F * R, dtype=np.uint8)
        g = [np.zeros(n, dtype=np.uint64) for n in (self.P, self.G, self.O)]
        sel, nsel = None, 0
        if faults is not None:
            arr = np.ascontiguousarray(np.asarray(faults, dtype=np.int64))
            sel, nsel = _p(arr, ctypes.c_int64), arr.size
        self._L.oracle_run_literal(self._h, threads, sel, nsel, _p(table, ctypes.c_uint64), _p(det, ctypes.c_uint8),
                                   _p(g[0], ctypes.c_uint64), _p(g[1], ctypes.c_uint64), _p(g[2], ctypes.c_uint64))
        return {"table": table, "det": det.reshape(F, R), "g_pi": g[0], "g_gate": g[1], "g_po": g[2]}

    def run_event(self, threads: int = 0, faults: Optional[Sequence[int]] = None) -> np.ndarray:
        """Event-driven restatement; returns det[F,R] (rows of unselected faults are 0)."""
        F, R = self.F, self.R
        det = np.zeros(F * R, dtype=np.uint8)
        sel, nsel = None, 0
        if faults is not None:
            arr = np.ascontiguousarray(np.asarray(faults, dtype=np.int64))
            sel, nsel = _p(arr, ctypes.c_int64), arr.size
        if self._L.oracle_run_event(self._h, threads, sel, nsel, _p(det, ctypes.c_uint8)) != 0:
            raise MemoryError("oracle_run_event failed")
        return det.reshape(F, R)

    def run_event_range(self, faults: Sequence[int], r0: int = 0, r1: Optional[int] = None, chunk: int = 64,
                        stop_at_first: bool = False, threads: int = 0) -> np.ndarray:
        """Event-driven restatement over rounds [r0, r1) for the given faults,
        the good machine `chunk` rounds at a time. Returns det[len(faults),
        r1 - r0]; with stop_at_first a fault's rounds after its first
        detection stay 0."""
        r1 = self.R if r1 is None else r1
        arr = np.ascontiguousarray(np.asarray(faults, dtype=np.int64))
        det = np.zeros(max(1, arr.size * (r1 - r0)), dtype=np.uint8)
        rc = self._L.oracle_run_event_range(self._h, threads, _p(arr, ctypes.c_int64), arr.size, r0, r1, chunk,
                                            int(stop_at_first), _p(det, ctypes.c_uint8))
        if rc != 0:
            raise ValueError("oracle_run_event_range failed (bad range or out of memory)")
        return det[:arr.size * (r1 - r0)].reshape(arr.size, r1 - r0)

    def first_rounds(self, faults: Sequence[int], chunk: int = 64, threads: int = 0) -> np.ndarray:
        """First detecting round of each given fault over all rounds, -1 if
        never (event-driven, each fault simulated up to its first detection)."""
        det = self.run_event_range(faults, 0, self.R, chunk=chunk, stop_at_first=True, threads=threads)
        if det.shape[1] == 0:
            return -np.ones(det.shape[0], dtype=np.int64)
        return np.where(det.any(1), det.argmax(1), -1).astype(np.int64)

    def good_round(self, rd: int):
        """Masked good values (g_pi, g_gate, g_po) of round rd alone."""
        g = [np.zeros(n, dtype=np.uint64) for n in (self.P, self.G, self.O)]
        if self._L.oracle_good_round(self._h, rd, *(_p(x, ctypes.c_uint64) for x in g)) != 0:
            raise ValueError("round out of range")
        return tuple(g)


def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def reference_dot(ckt, flst, ptn) -> str:
    """The reference's Simulator::visualization_graph text for the three texts
    (fsim/simulation.cpp:1157-1225), from the compiled reference."""
    with tempfile.TemporaryDirectory() as d:
        paths = []
        for name, data in (("c.ckt", ckt), ("f.flst", flst), ("p.ptn", ptn)):
            path = os.path.join(d, name)
            with open(path, "wb") as fh:
                fh.write(_b(data))
            paths.append(path)
        dot = os.path.join(d, "g.dot")
        subprocess.run([REF_BIN] + paths + ["0", "1", "-", dot], check=True, capture_output=True, text=True)
        with open(dot) as fh:
            return fh.read()


def run_reference(ckt, flst, ptn, mode: int = 0, threads: int = 1, want_output: bool = True):
    """Run the reference binary on the three texts. Returns (dict, seconds) where
    dict holds counts, order map, gates per level, levels, table and good values."""
    with tempfile.TemporaryDirectory() as d:
        paths = []
        for name, data in (("c.ckt", ckt), ("f.flst", flst), ("p.ptn", ptn)):
            path = os.path.join(d, name)
            with open(path, "wb") as fh:
                fh.write(_b(data))
            paths.append(path)
        out = os.path.join(d, "out.txt") if want_output else "-"
        r = subprocess.run([REF_BIN] + paths + [str(mode), str(threads), out], check=True, capture_output=True,
                           text=True)
        parts = r.stdout.split()
        secs = float(parts[parts.index("run_s") + 1])
        if not want_output:
            return None, secs
        with open(out) as fh:
            lines = fh.read().split("\n")
    def nums(line):
        return np.array([int(x) for x in line.split()], dtype=np.uint64)
    counts = [int(x) for x in lines[0].split()]
    res = {
        "P": counts[0], "O": counts[1], "G": counts[2], "W": counts[3], "R": counts[4], "F": counts[5],
        "order_map": nums(lines[1]), "gates_per_level": nums(lines[2]), "levels": nums(lines[3]),
        "table": nums(lines[4]), "g_pi": nums(lines[5]), "g_gate": nums(lines[6]), "g_po": nums(lines[7]),
    }
    return res, secs
