"""The fused count of the last two join levels equals the level-by-level
count and the oracle: labeled random walks, unlabeled patterns whose last
vertex has shared lists, an own list, both or only same-label conflicts,
with the shared intersection in LDS and (small tile cap) the group
fallback; plain path and lockstep batch."""
import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from fsig_amd import workloads as W

pytestmark = pytest.mark.gpu

PATTERNS = {
    "triangle": (3, [(0, 1), (1, 2), (0, 2)]),
    "path4": (4, [(0, 1), (1, 2), (2, 3)]),            # leaf conflicts with non-adjacent columns
    "cycle4": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
    "star4": (4, [(0, 1), (0, 2), (0, 3)]),            # last vertex: a shared list, a conflict with column p
    "clique4": (4, [(a, b) for a in range(4) for b in range(a + 1, 4)]),
    "house": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (1, 4)]),
    "diamond_tail": (5, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4)]),
}


def _count(ctx, dg, q, fused, tile=0, G=0):
    ctx.set_tune(G, 0, tile, 4, 0 if fused else 1)
    try:
        return F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count
    finally:
        ctx.set_tune(0, 0, 0, 4, 0)


def test_tail_unlabeled_patterns(gpu_ctx):
    g = F.HostGraph.rmat(10, 6000, seed=21)
    off, adj, _ = g.arrays()
    lab = np.zeros(g.n, np.int32)
    dg = gpu_ctx.upload(g)
    for name, (k, e) in PATTERNS.items():
        q = F.Query(k, np.zeros(k, np.int32), e)
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)
        got = {(fused, tile, G): _count(gpu_ctx, dg, q, fused, tile, G)
               for fused in (True, False) for tile in (0, 4) for G in (4, 16, 64)}
        print(name, want, got, flush=True)
        assert all(v == want for v in got.values()), name
    dg.free()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tail_labeled_walks(gpu_ctx, seed):
    g = F.HostGraph.chung_lu(6000, 30000, 2.3, 40 + seed).set_labels("zipf", 4, 1.2, 50 + seed)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    qs = W.walk_queries(g, 16, 3, 8, seed, dg=dg, max_count=2_000_000, max_rows=4_000_000)
    for kind, q in qs:
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, limit=10 ** 7)
        fused = _count(gpu_ctx, dg, q, True)
        capped = _count(gpu_ctx, dg, q, True, tile=2)
        plain = _count(gpu_ctx, dg, q, False)
        print(kind, q.k, q.num_edges, want, fused, capped, plain, flush=True)
        assert fused == plain == capped
        if want != O.TOO_MANY:
            assert fused == want
    dg.free()


def test_tail_lockstep_batch(gpu_ctx):
    g = F.HostGraph.chung_lu(8000, 40000, 2.3, 61).set_labels("zipf", 4, 1.2, 62)
    dg = gpu_ctx.upload(g)
    qs = [q for _, q in W.walk_queries(g, 24, 3, 9, 5, dg=dg, max_count=2_000_000, max_rows=4_000_000)]
    qs += [F.Query(k, np.zeros(k, np.int32), e) for k, e in (PATTERNS["path4"], PATTERNS["star4"])]
    want = [_count(gpu_ctx, dg, q, False) for q in qs]
    for G in (8, 16, 64):
        for fused in (0, 1):
            gpu_ctx.set_tune(G, 0, 0, 4, fused)
            try:
                got, _ = F.match_batch(dg, qs)
            finally:
                gpu_ctx.set_tune(0, 0, 0, 4, 0)
            assert got.tolist() == want, (G, fused)
    dg.free()


CHAINS = [
    ([0, 1, 0, 1, 0, 1, 0, 1], [(i, i + 1) for i in range(7)]),         # same-label columns all along
    ([0, 0, 0, 0, 0, 0], [(i, i + 1) for i in range(5)]),               # one label: every column conflicts
    ([1, 0, 1, 1, 0, 1], [(0, 1), (1, 2), (1, 3), (3, 4), (4, 5)]),     # a tree ending in a chain
    ([0, 1, 2, 0, 1], [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)]),  # not a chain end (control)
]


@pytest.mark.parametrize("graph_labels", [2, 3])
def test_chain_tail_equals_level_by_level_and_oracle(gpu_ctx, graph_labels):
    """Queries whose last vertex hangs off the previous position, which has
    one backward list (the chain tail): its count equals the level-by-level
    count and the oracle, with conflicts of the same and of other classes."""
    g = F.HostGraph.chung_lu(3000, 15000, 2.3, 70 + graph_labels).set_labels("uniform", graph_labels, seed=71)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    for labels, edges in CHAINS:
        q = F.Query(len(labels), [l % graph_labels for l in labels], edges)
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, limit=5 * 10 ** 8)
        fused = {G: _count(gpu_ctx, dg, q, True, 0, G) for G in (4, 8, 16)}
        plain = _count(gpu_ctx, dg, q, False)
        print(labels, want, fused, plain, flush=True)
        assert all(v == plain for v in fused.values())
        if want != O.TOO_MANY:
            assert plain == want
    dg.free()
