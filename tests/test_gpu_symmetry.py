"""Counts with symmetry breaking (one embedding per orbit of the query's
automorphism group, times the group's order) equal the counts of every
embedding and the oracle: symmetric unlabeled patterns (cycles, cliques,
stars, the house, K2,3, a 6-cycle with chords) and labeled ones whose
labels keep part of the group, alg 1 with the fused tail on and off and
several lane counts, alg 2, and bounded counts (min(count, bound) with the
bound scaled by the group's order inside).  The star's count is its closed
form (per center, ordered choices of distinct neighbours), the others the
oracle's."""
import numpy as np
import pytest

import fsig_amd as F
import oracle as O

pytestmark = pytest.mark.gpu

PATTERNS = {
    "triangle": (3, [(0, 1), (1, 2), (0, 2)]),
    "cycle4": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
    "clique4": (4, [(a, b) for a in range(4) for b in range(a + 1, 4)]),
    "house": (5, [(0, 1), (1, 2), (2, 3), (3, 0), (2, 4), (3, 4)]),
    "star1+4": (5, [(0, 1), (0, 2), (0, 3), (0, 4)]),
    "path5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)]),
    "cycle5": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]),
    "K2,3": (5, [(a, b) for a in (0, 1) for b in (2, 3, 4)]),
    "diamond": (4, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]),
    "cycle6+chords": (6, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0), (0, 3), (1, 4)]),
}


def _count(ctx, dg, q, alg=F.ALG_MMATCH, symmetry=0, fused=0, G=0, bound=0):
    ctx.set_tune(G, 0, 0, 0, fused, 0, 0, symmetry)
    try:
        return F.match(dg, q, alg, F.MODE_COUNT, max_embeddings=bound).count
    finally:
        ctx.set_tune()


def _star_count(off, adj, lab, labels):
    """Embeddings of the star center + 4 leaves (labels[0] the center's):
    per center, ordered choices of distinct neighbours with the leaves'
    labels (a closed form; the oracle would enumerate them)."""
    deg_by = {l: np.zeros(len(off) - 1, np.int64) for l in set(labels[1:])}
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    for l in deg_by:
        np.add.at(deg_by[l], src, (lab[adj] == l).astype(np.int64))
    total = np.ones(len(off) - 1, dtype=object)
    for l in deg_by:
        need = sum(1 for x in labels[1:] if x == l)
        d = deg_by[l].astype(object)
        for i in range(need):
            total = total * (d - i)
    return int(sum(t for t, c in zip(total, lab) if c == labels[0]))


def _want(name, q, off, adj, lab):
    if name == "star1+4":
        return _star_count(off, adj, lab, q.labels.tolist())
    return O.match(q.labels, q.adjm(), off, adj, lab, threads=8)[0]


def test_symmetric_unlabeled_patterns(gpu_ctx):
    big = F.HostGraph.rmat(9, 2500, seed=31)
    small = F.HostGraph.rmat(8, 1200, seed=31)  # the 8-edge pattern: the oracle's enumeration stays short
    for name, (k, e) in PATTERNS.items():
        g = small if name == "cycle6+chords" else big
        off, adj, _ = g.arrays()
        lab = np.zeros(g.n, np.int32)
        dg = gpu_ctx.upload(g, pcsr=True)
        q = F.Query(k, np.zeros(k, np.int32), e)
        want = _want(name, q, off, adj, lab)
        got = {("sym", fused, G): _count(gpu_ctx, dg, q, fused=fused, G=G) for fused in (0, 1) for G in (0, 4, 64)}
        got["all"] = _count(gpu_ctx, dg, q, symmetry=1)
        got["gsi sym"] = _count(gpu_ctx, dg, q, alg=F.ALG_GSI)
        got["gsi all"] = _count(gpu_ctx, dg, q, alg=F.ALG_GSI, symmetry=1)
        print(name, want, got, flush=True)
        assert all(v == want for v in got.values()), name
        dg.free()


@pytest.mark.parametrize("seed", [1, 2])
def test_symmetric_labeled_patterns(gpu_ctx, seed):
    """Labels a b a b ... on the patterns keep a subgroup of their
    automorphisms (and none for some)."""
    g = F.HostGraph.chung_lu(4000, 20000, 2.3, 60 + seed).set_labels("uniform", 2, seed=70 + seed)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True)
    for name, (k, e) in PATTERNS.items():
        q = F.Query(k, np.array([i % 2 for i in range(k)], np.int32), e)
        want = _want(name, q, off, adj, lab)
        got = {"sym": _count(gpu_ctx, dg, q), "all": _count(gpu_ctx, dg, q, symmetry=1),
               "sym level by level": _count(gpu_ctx, dg, q, fused=1), "gsi sym": _count(gpu_ctx, dg, q, alg=F.ALG_GSI)}
        print(name, want, got, flush=True)
        assert all(v == want for v in got.values()), name
    dg.free()


def test_symmetric_bounded_counts(gpu_ctx):
    """min(count, bound) for bounds below, at and above the count, and ones
    that the group's order does not divide."""
    g = F.HostGraph.rmat(10, 7000, seed=33)
    dg = gpu_ctx.upload(g)
    for name in ("triangle", "cycle4", "clique4", "star1+4"):
        k, e = PATTERNS[name]
        q = F.Query(k, np.zeros(k, np.int32), e)
        total = _count(gpu_ctx, dg, q, symmetry=1)
        for b in (1, 5, 23, 1000, total - 1, total, total + 1, 10 ** 15):
            if b <= 0:
                continue
            got, want = _count(gpu_ctx, dg, q, bound=b), min(total, b)
            assert got == want, (name, b, got, want)
    dg.free()
