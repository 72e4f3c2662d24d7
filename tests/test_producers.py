"""Fault-list and pattern producers with the conventions of
fsim/generate_fault_patterns.py: PO-excluded distinct fault pairs, distinct
patterns, and the "new" writer's packing (first pattern most significant,
rounds of 32 or 64 patterns, a binary copy after the decimal words).

Pinned by the reference's own example (write_simple_testcase_to_file: PI
words 2, 4, 1 for the patterns 010 / 100 / 001 per PI), by a string-level
restatement of the writer (oracle/producers.py), and by reading produced
files back through the product reader."""
import numpy as np
import pytest

import lsim_amd
from oracle import producers


def test_reference_example():
    # fsim/generate_fault_patterns.py:114-124: PI_0 010, PI_1 100, PI_2 001 over 3 patterns.
    pats = [[0, 1, 0], [1, 0, 0], [0, 0, 1]]
    assert lsim_amd.pack_patterns(pats) == b"3\n2 4 1 \n\n010 100 001 \n\n"


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("n,P", [(1, 3), (17, 5), (32, 7), (64, 9), (65, 4), (200, 12), (1000, 40)])
def test_packing_matches_restatement(n, P, bits):
    rng = np.random.default_rng(n * 131 + P)
    pats = rng.integers(0, 2, size=(n, P), dtype=np.uint8)
    got = lsim_amd.pack_patterns(pats, bits).decode()
    assert got == producers.pack(pats.tolist(), bits)


def _unpack(words, n):
    """Patterns back from [R][P] words: in round r the first of its
    L = min(64, n - 64 r) patterns is bit L - 1."""
    R, P = words.shape
    out = []
    for r in range(R):
        L = min(64, n - 64 * r)
        for k in range(L):
            out.append([(int(words[r, i]) >> (L - 1 - k)) & 1 for i in range(P)])
    return np.array(out, dtype=np.uint8)


def _reader(P):
    # P PIs feeding an AND chain into one PO: any file with R <= P rounds reads.
    ids = list(range(P))
    g0 = P + 1
    recs = [f"{i} 0 1 {g0} {1 if i == 0 else 2}" if i < 2 else f"{i} 0 1 {g0 + i - 1} 2" for i in ids]
    recs += [f"{g0 + k} 0 1 {g0 + k + 1} 1" for k in range(P - 2)] + [f"{g0 + P - 2} 0 1 {P} 1"]
    ckt = f"{P} 1 {P - 1} {len(recs)}\n" + "\n".join(recs) + "\n" + " ".join(["1"] * (P - 1)) + "\n"
    return ckt


@pytest.mark.parametrize("n,P", [(5, 4), (64, 8), (100, 8), (1000, 64)])
def test_packed_patterns_read_back(n, P):
    rng = np.random.default_rng(n + P)
    pats = rng.integers(0, 2, size=(n, P), dtype=np.uint8)
    s = lsim_amd.Simulator(device=-1)
    s.read_text(_reader(P), "0", lsim_amd.pack_patterns(pats))
    assert np.array_equal(_unpack(s.patterns(), n), pats)


def test_produced_patterns_are_distinct_and_read_back():
    n, P = 3000, 64  # 47 rounds: the reader needs rounds <= PIs
    text = lsim_amd.produce_patterns(9, P, n)
    assert text == lsim_amd.produce_patterns(9, P, n)  # seeded
    s = lsim_amd.Simulator(device=-1)
    s.read_text(_reader(P), "0", text)
    pats = _unpack(s.patterns(), n)
    assert len({tuple(p) for p in pats}) == n
    # all 2^P patterns when n = 2^P; more is rejected (:40-41)
    full = lsim_amd.produce_patterns(1, 4, 16)
    s2 = lsim_amd.Simulator(device=-1)
    s2.read_text(_reader(4), "0", full)
    assert len({tuple(p) for p in _unpack(s2.patterns(), 16)}) == 16
    with pytest.raises(lsim_amd.LsimError):
        lsim_amd.produce_patterns(1, 4, 17)


@pytest.mark.parametrize("N,F,P,O", [(50, 20, 5, 5), (50, 90, 5, 5), (1000, 1990, 10, 5), (10, 0, 2, 3)])
def test_produced_faults(N, F, P, O):
    text = lsim_amd.produce_faults(3, N, F, P, O).decode().split("\n")
    assert int(text[0]) == F
    pairs = [tuple(map(int, line.split())) for line in text[1:] if line]
    assert len(pairs) == F and len(set(pairs)) == F
    assert all(0 <= w < N and not (P <= w < P + O) and v in (0, 1) for w, v in pairs)
    if F == 2 * (N - O):
        assert set(pairs) == {(w, v) for w in range(N) if not (P <= w < P + O) for v in (0, 1)}


def test_too_many_faults_rejected():
    with pytest.raises(lsim_amd.LsimError) as e:
        lsim_amd.produce_faults(3, 10, 15, 2, 3)
    assert "num_of_faults must be less than or equal" in e.value.message


def test_produced_faults_read_by_simulator():
    P, G = 6, 20
    ckt = _reader(P)
    N = P + 1 + (P - 1)
    flst = lsim_amd.produce_faults(5, N, 2 * (N - 1), P, 1)
    s = lsim_amd.Simulator(device=-1)
    s.read_text(ckt, flst, lsim_amd.produce_patterns(2, P, 64))
    assert s.num_faults() == 2 * (N - 1)
