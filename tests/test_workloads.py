"""Workload definitions: the C4 and C5 laws keep every drawn walk (no redraw
for size), draw k from the stated laws, and the C5 arrivals are a seeded
Poisson process."""
import numpy as np

import fsig_amd as F
from fsig_amd import workloads as W


def _connected(q):
    m = q.adjm()
    seen, todo = {0}, [0]
    while todo:
        u = todo.pop()
        for w in range(q.k):
            if (int(m[u]) >> w) & 1 and w not in seen:
                seen.add(w)
                todo.append(w)
    return len(seen) == q.k


def test_laws_are_fixed():
    assert W.C4_QUERIES == dict(count=256, kmin=4, kmax=16, seed=4, bound=10 ** 8, work_limit=2 * 10 ** 9)
    assert W.C5_STREAM["kmin"] == 4 and W.C5_STREAM["kmax"] == 32 and W.C5_STREAM["bound"] == 10 ** 8


def test_walk_queries_keep_every_draw_and_alternate_density():
    g = F.HostGraph.chung_lu(20000, 120000, 2.3, 3).set_labels("zipf", 8, 1.2, 4)
    info = {}
    qs = W.walk_queries(g, 40, 4, 16, 4, info=info)
    assert len(qs) == 40 and info["draws"] == 40 + info["no_walk"]
    for i, (kind, q) in enumerate(qs):
        assert kind == ("dense" if i % 2 == 0 else "sparse")
        assert 4 <= q.k <= 16 and _connected(q)
        assert (2 * q.num_edges >= 3 * q.k) == (kind == "dense")
    again = W.walk_queries(g, 40, 4, 16, 4)  # seeded: the same queries
    assert all(a.edges.tolist() == b.edges.tolist() for (_, a), (_, b) in zip(qs, again))


def test_c5_law_is_log_uniform_k_and_poisson_arrivals():
    g = F.HostGraph.rmat(14, 200000, seed=3).set_labels("uniform", 64, seed=4)
    info = {}
    stream = W.c5_stream(g, count=64, info=info)
    ks = [q.k for _, q in stream]
    assert len(stream) == 64 and min(ks) >= 4 and max(ks) <= 32
    assert max(ks) >= 12  # log-uniform to 32, not conditioned on small k
    t = np.array([a for a, _ in stream])
    assert t[0] == 0 and np.all(np.diff(t) > 0)
    gaps = np.random.default_rng(W.C5_STREAM["arrival_seed"]).exponential(1.0 / W.C5_STREAM["qps"], size=64)
    assert np.allclose(t, np.cumsum(gaps) - gaps[0])
