"""Pins the oracle before it is trusted: closed forms, brute force over all
injective maps, networkx VF2 golden fixtures, and the invariants of the
refinement restatement."""
import itertools

import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from conftest import adjm_from_edges, csr_from_edges, load_golden


def brute_force(n, off, adj, labels, k, qedges, qlabels):
    nbr = [set(adj[off[v]:off[v + 1]].tolist()) for v in range(n)]
    rows = []
    for phi in itertools.permutations(range(n), k):
        if any(labels[phi[u]] != qlabels[u] for u in range(k)):
            continue
        if all(phi[b] in nbr[phi[a]] for a, b in qedges):
            rows.append(list(phi))
    return sorted(rows)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(6, 9))
    edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.45]
    labels = rng.integers(0, 2, n)
    off, adj, lab = csr_from_edges(n, edges, labels)
    for k, qe in [(2, [(0, 1)]), (3, [(0, 1), (1, 2)]), (3, [(0, 1), (1, 2), (0, 2)]),
                  (4, [(0, 1), (1, 2), (2, 3), (3, 0)]), (4, [(0, 1), (2, 3)])]:
        ql = rng.integers(0, 2, k)
        want = brute_force(n, off, adj, lab, k, qe, ql)
        cnt, rows = O.match(ql, adjm_from_edges(k, qe), off, adj, lab, cap=10 ** 6)
        assert cnt == len(want)
        assert sorted(rows.tolist()) == want


def test_oracle_equals_vf2_fixtures():
    for c in load_golden("vf2_cases.json"):
        off, adj, lab = csr_from_edges(c["n"], c["edges"], c["labels"])
        cnt, rows = O.match(c["qlabels"], adjm_from_edges(c["k"], c["qedges"]), off, adj, lab, cap=10 ** 6)
        assert cnt == c["count"], (c["graph"], c["query"])
        assert sorted(rows.tolist()) == c["rows"]


def _unlabeled(g):
    off, adj, _ = g.arrays()
    return off, adj, np.zeros(g.n, np.int32)


def test_closed_forms_triangles_and_four_cycles():
    g = F.HostGraph.rmat(10, 6000, seed=5)
    off, adj, lab = _unlabeled(g)
    n = g.n
    A = np.zeros((n, n), np.int64)
    for v in range(n):
        A[v, adj[off[v]:off[v + 1]]] = 1
    A2 = A @ A
    tri = int(np.trace(A2 @ A))
    deg = A.sum(1)
    m = int(A.sum() // 2)
    c4 = int(np.trace(A2 @ A2)) - 2 * int((deg ** 2).sum()) + 2 * m
    z = np.zeros(3, np.int32)
    assert O.match(z, adjm_from_edges(3, [(0, 1), (1, 2), (0, 2)]), off, adj, lab, threads=4)[0] == tri
    z4 = np.zeros(4, np.int32)
    assert O.match(z4, adjm_from_edges(4, [(0, 1), (1, 2), (2, 3), (3, 0)]), off, adj, lab, threads=4)[0] == c4
    # paths of length 2: sum d(d-1); stars with 3 leaves: sum d(d-1)(d-2)
    assert O.match(z, adjm_from_edges(3, [(0, 1), (1, 2)]), off, adj, lab)[0] == int((deg * (deg - 1)).sum())
    assert O.match(z4, adjm_from_edges(4, [(0, 1), (0, 2), (0, 3)]), off, adj, lab)[0] == \
        int((deg * (deg - 1) * (deg - 2)).sum())


def test_closed_forms_complete_graph():
    n = 9
    g = F.HostGraph.from_edges(n, [(u, v) for u in range(n) for v in range(u + 1, n)], [0] * n)
    off, adj, lab = g.arrays()
    K4 = [(a, b) for a in range(4) for b in range(a + 1, 4)]
    assert O.match(np.zeros(4, np.int32), adjm_from_edges(4, K4), off, adj, lab)[0] == 9 * 8 * 7 * 6
    assert O.match(np.zeros(3, np.int32), adjm_from_edges(3, [(0, 1), (1, 2), (0, 2)]), off, adj, lab)[0] == 504


def test_oracle_threads_agree():
    g = F.HostGraph.chung_lu(5000, 30000, 2.3, 3).set_labels("uniform", 3, seed=4)
    off, adj, lab = g.arrays()
    q = g.walk_query(5, 5, 10, seed=9)
    a = O.match(q.labels, q.adjm(), off, adj, lab, threads=1)[0]
    b = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)[0]
    assert a == b and a >= 1


def test_ldf_closed_form():
    g = F.HostGraph.chung_lu(4000, 16000, 2.5, 8).set_labels("zipf", 5, 1.2, 8)
    off, adj, lab = g.arrays()
    deg = np.diff(off)
    for label in range(5):
        for d in (0, 2, 7):
            want = np.nonzero((lab == label) & (deg >= d))[0]
            assert np.array_equal(O.ldf(off, lab, label, d), want)


def test_refine_invariants():
    """Every embedding's image lies in the refined sets (soundness), the sets
    shrink with more passes and stay inside the LDF sets."""
    g = F.HostGraph.chung_lu(3000, 15000, 2.5, 21).set_labels("uniform", 4, seed=22)
    off, adj, lab = g.arrays()
    deg = np.diff(off)
    for seed in range(4):
        q = g.walk_query(5, 5, 8, seed=100 + seed)
        _, rows = O.match(q.labels, q.adjm(), off, adj, lab, cap=10 ** 6)
        prev = None
        for passes in (0, 1, 2, 6):
            sizes, sets = O.refine(q.labels, q.adjm(), off, adj, lab, passes, want_sets=True)
            for u in range(q.k):
                assert sets[u][rows[:, u]].all()
                qd = int(bin(int(q.adjm()[u])).count("1"))
                ldf = (lab == q.labels[u]) & (deg >= qd)
                assert not (sets[u].astype(bool) & ~ldf).any()
            if prev is not None:
                assert (sizes <= prev).all()
            prev = sizes


def test_classes_are_equitable():
    # a 6-cycle with one chord: classes must give equal neighbour-class counts
    k = 6
    edges = [(i, (i + 1) % k) for i in range(k)] + [(0, 3)]
    m = adjm_from_edges(k, edges)
    cls = O.classes(np.zeros(k, np.int32), m)
    for a in range(k):
        for b in range(k):
            if cls[a] == cls[b]:
                ca = sorted(cls[w] for w in range(k) if (int(m[a]) >> w) & 1)
                cb = sorted(cls[w] for w in range(k) if (int(m[b]) >> w) & 1)
                assert ca == cb
    assert cls[0] == cls[3] and cls[1] == cls[2] == cls[4] == cls[5] and cls[0] != cls[1]
