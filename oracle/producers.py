"""Restatement of the reference's pattern writer
(fsim/generate_fault_patterns.py:56-94, write_patterns_to_file_new).

TEST INFRASTRUCTURE ONLY: the checker for lsim_pack_patterns. String-for-
string the steps of the reference writer over an explicit list of patterns
(the reference's random draws and set order are not reproduced)."""


def pack(patterns, bits_per_round: int = 64) -> str:
    n = len(patterns)
    P = len(patterns[0])
    out = [f"{n}\n"]
    bits = int(bits_per_round)
    team = 1
    if n > bits:  # :62-64
        team = -(-n // bits)
    for t in range(team):  # decimal section, :67-77
        for b in range(P):
            col = [str(p[b]) for p in patterns]
            a = t * bits
            out.append(str(int("".join(col[a:a + bits]), 2)) + " ")
        out.append("\n")
    out.append("\n")
    for t in range(team):  # binary section, :80-90
        for b in range(P):
            col = [str(p[b]) for p in patterns]
            a = t * bits
            out.append("".join(col[a:a + bits]) + " ")
        out.append("\n")
    out.append("\n")
    return "".join(out)
