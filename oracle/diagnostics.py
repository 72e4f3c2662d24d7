"""Restatement of the reference's structural diagnostics
(fsim/cuda_MA_partition/cuda_MA_partition.cu:227-346: _test_parent_dist,
_test_spatial_locality, _test_unique_parent_count), with the group size a
parameter (the reference uses 32-thread warps).

TEST INFRASTRUCTURE ONLY: the checker for lsim_level_diagnostics. Input:
per node (topological position) its level and its fanin positions."""


def level_diagnostics(levels, fanins, wave=64):
    L = max(levels) + 1
    by_level = [[] for _ in range(L)]
    for p, l in enumerate(levels):
        by_level[l].append(p)
    out = {"nodes": [], "avg_dist": [], "max_span": [], "avg_range": [], "avg_unique": []}
    for l in range(L):
        v = by_level[l]
        total, span = 0.0, 0  # :236-265
        for g in v:
            ps = fanins[g]
            if not ps:
                continue
            s = 0.0
            for q in ps:
                d = l - levels[q]
                s += d
                span = max(span, d)
            total += s / len(ps)
        rng_total, uniq_total, groups = 0.0, 0.0, 0  # :279-297, :310-333
        for i in range(0, len(v), wave):
            ids = [q for g in v[i:i + wave] for q in fanins[g]]
            groups += 1
            if ids:
                rng_total += max(ids) - min(ids)
            uniq_total += len(set(ids))
        out["nodes"].append(len(v))
        out["avg_dist"].append(total / len(v) if v else 0.0)
        out["max_span"].append(span)
        out["avg_range"].append(rng_total / groups if groups else 0.0)
        out["avg_unique"].append(uniq_total / groups if groups else 0.0)
    return out
