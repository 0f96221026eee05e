"""Workload definitions: the committed C5 request stream matches its
parameters and holds well-formed queries; C4/C5 parameter sets are fixed."""
import json

import numpy as np

import fsig_amd as F
from fsig_amd import workloads as W


def test_c5_stream_file_matches_parameters():
    with open(W.C5_STREAM_FILE) as f:
        rec = json.load(f)
    assert rec["key"] == W._stream_key(W.C5_STREAM)
    qs = rec["queries"]
    assert len(qs) == W.C5_STREAM["count"]
    for r in qs:
        k = r["k"]
        assert W.C5_STREAM["kmin"] <= k <= W.C5_STREAM["k_cap"]
        assert len(r["labels"]) == k and all(0 <= x < 64 for x in r["labels"])
        q = F.Query(k, r["labels"], r["edges"])
        # walk-induced: connected, simple, k - 1 <= |E| <= k(k-1)/2
        assert k - 1 <= q.num_edges <= k * (k - 1) // 2
        m = q.adjm()
        seen, todo = {0}, [0]
        while todo:
            u = todo.pop()
            for w in range(k):
                if (int(m[u]) >> w) & 1 and w not in seen:
                    seen.add(w)
                    todo.append(w)
        assert len(seen) == k


def test_c5_arrivals_are_seeded_poisson():
    # the arrival process depends only on its seed and the rate
    g = np.random.default_rng(W.C5_STREAM["arrival_seed"]).exponential(1.0 / 1000.0, size=128)
    t = np.cumsum(g) - g[0]
    assert t[0] == 0 and np.all(np.diff(t) > 0)
    assert 0.08 < t[-1] < 0.18  # 127 gaps of mean 1 ms
