"""Pins the oracle before it is trusted: closed forms, brute force over all
injective maps, networkx VF2 golden fixtures, and the invariants of the
refinement restatement."""
import itertools

import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from fsig_amd import workloads as W
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
    assert 6 * O.triangles(off, adj, threads=3) == tri
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


# ---------------------------------------------------------------------------
# The set digest (fsig_match_digest's checker) and the row checker.
def test_row_hash_equals_numpy_restatement():
    rng = np.random.default_rng(3)
    for k in (1, 3, 8, 32):
        rows = rng.integers(0, 2 ** 32, size=(200, k), dtype=np.uint64).astype(np.uint32)
        want = O.row_hashes_numpy(rows)
        got = [O.lib().oracle_row_hash(k, r.ctypes.data_as(O._u32p)) for r in np.ascontiguousarray(rows)]
        assert [int(x) for x in want] == got
    # a known answer: the empty chain is the seed, one word is one finaliser step
    x0 = (0x9e3779b97f4a7c15 * 2) & ((1 << 64) - 1)
    with np.errstate(over="ignore"):
        want0 = int(O._mix64_np(np.array([x0], np.uint64))[0])
    assert int(O.row_hashes_numpy(np.zeros((1, 1), np.uint32))[0]) == want0


def test_digest_equals_numpy_over_listed_rows():
    for c in load_golden("vf2_cases.json")[:40]:
        off, adj, lab = csr_from_edges(c["n"], c["edges"], c["labels"])
        m = adjm_from_edges(c["k"], c["qedges"])
        cnt, dig = O.match_digest(c["qlabels"], m, off, adj, lab, threads=2)
        assert cnt == c["count"]
        rows = np.asarray(c["rows"], np.uint32).reshape(-1, c["k"])
        assert dig == O.digest_numpy(rows), (c["graph"], c["query"])


def test_digest_separates_sets_and_ignores_order():
    g = F.HostGraph.chung_lu(3000, 15000, 2.5, 11).set_labels("uniform", 3, seed=12)
    off, adj, lab = g.arrays()
    q = g.walk_query(4, 3, 6, seed=5)
    cnt, rows = O.match(q.labels, q.adjm(), off, adj, lab, threads=4, cap=10 ** 6)
    assert 100 < cnt <= 10 ** 6
    cnt2, dig = O.match_digest(q.labels, q.adjm(), off, adj, lab, threads=4)
    assert cnt2 == cnt and dig == O.digest_numpy(rows)
    rng = np.random.default_rng(0)
    assert O.digest_numpy(rows[rng.permutation(len(rows))]) == dig  # order of rows
    changed = rows.copy()
    changed[7, 2] ^= 1  # one word
    assert O.digest_numpy(changed) != dig
    assert O.digest_numpy(rows[1:]) != dig  # a row dropped
    assert O.digest_numpy(np.vstack([rows, rows[:1]])) != dig  # a row repeated
    swapped = rows.copy()
    swapped[:, [0, 1]] = swapped[:, [1, 0]]  # columns transposed
    assert O.digest_numpy(swapped) != dig


def test_check_rows_flags_planted_mistakes():
    g = F.HostGraph.chung_lu(3000, 15000, 2.5, 13).set_labels("uniform", 3, seed=14)
    off, adj, lab = g.arrays()
    q = g.walk_query(4, 4, 6, seed=6)
    cnt, rows = O.match(q.labels, q.adjm(), off, adj, lab, threads=4, cap=10 ** 6)
    assert cnt == len(rows) > 10
    bad, h = O.check_rows(q.labels, q.adjm(), off, adj, lab, rows, threads=4, want_hashes=True)
    assert bad == 0 and np.array_equal(h, O.row_hashes_numpy(rows))
    a, b = (int(x) for x in q.edges[0])
    planted = rows[:4].copy()
    planted[0, 1] = planted[0, 0]                               # a repeated vertex
    wrong = np.nonzero(lab != q.labels[2])[0][0]
    planted[1, 2] = wrong                                       # a wrong label
    nb = set(adj[off[planted[2, a]]:off[planted[2, a] + 1]].tolist())
    same = [v for v in np.nonzero(lab == q.labels[b])[0] if v not in nb and v not in planted[2]]
    planted[2, b] = same[0]                                     # a query edge missing in the data
    planted[3, 0] = g.n                                         # not a vertex
    assert O.check_rows(q.labels, q.adjm(), off, adj, lab, planted, threads=2)[0] == 4
    assert O.check_rows(q.labels, q.adjm(), off, adj, lab, rows[4:], threads=2)[0] == 0


# ---------------------------------------------------------------------------
# The signature filter (filter_kernel, match/GSI_kernels.cuh:8-24) against
# what the mathematics fixes.
def _sig_slots(num_labels):
    """Effective 2-bit counter slot of each label in the 16-word signature:
    word 1 + h // 16 and bit shift 4 * (h % 8) with h = MurmurHash2(label) %
    240 (graph_format_transformer.cc:217-281 as x86 executes its shift)."""
    h = np.array([O.murmur2_u32(l) % 240 for l in range(num_labels)])
    return (h // 16) * 8 + (h % 8)  # 120 distinct (word, shift) slots


def _slot_counts(off, adj, lab, slots):
    """Per vertex and slot: neighbours whose label maps there, saturated at 2."""
    n = len(lab)
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    cnt = np.zeros((n, 120), np.int64)
    np.add.at(cnt, (src, slots[lab[adj]]), 1)
    return np.minimum(cnt, 2)


def test_sig_filter_closed_form():
    """v passes for u iff same label and, in every counter slot, min(count, 2)
    of v's neighbours is at least u's: the subset test of 2-bit saturating
    counters (0 -> 01 -> 11).  Labels may share a slot (their counts add),
    which the closed form models through the same slot map."""
    L = 48
    g = F.HostGraph.chung_lu(6000, 40000, 2.3, 17).set_labels("uniform", L, seed=18)
    off, adj, lab = g.arrays()
    dsig = O.signature(off, adj, lab, True)
    slots = _sig_slots(L)
    sat = _slot_counts(off, adj, lab, slots)
    fails_only_in_word = set()
    for s in range(12):
        q = g.walk_query(5, 4, 8, seed=300 + s)
        qoff, qadj = q.csr()
        qsig = O.signature(qoff, qadj, q.labels, False).reshape(q.k, 16)
        qsat = _slot_counts(qoff, qadj, q.labels, slots)
        for u in range(q.k):
            ok_slot = sat >= qsat[u][None, :]
            want = np.nonzero((lab == q.labels[u]) & ok_slot.all(axis=1))[0]
            assert np.array_equal(O.sig_filter(dsig, qsig[u]), want), (s, u)
            # record words that alone reject some vertex of the label: the
            # comparison then covers every word a planted skip would miss
            words = ok_slot.reshape(len(lab), 15, 8).all(axis=2)
            for w in range(15):
                only = (lab == q.labels[u]) & ~words[:, w] & np.delete(words, w, axis=1).all(axis=1)
                if only.any():
                    fails_only_in_word.add(w)
    assert len(fails_only_in_word) >= 10


def test_sig_filter_is_sound():
    """Every image of every embedding passes its query vertex's filter."""
    g = F.HostGraph.chung_lu(4000, 20000, 2.5, 19).set_labels("zipf", 6, 1.2, 20)
    off, adj, lab = g.arrays()
    dsig = O.signature(off, adj, lab, True)
    for s in range(5):
        q = g.walk_query(5, 4, 8, seed=400 + s)
        qoff, qadj = q.csr()
        qsig = O.signature(qoff, qadj, q.labels, False).reshape(q.k, 16)
        cnt, rows = O.match(q.labels, q.adjm(), off, adj, lab, threads=4, cap=10 ** 6)
        for u in range(q.k):
            passed = np.zeros(g.n, bool)
            passed[O.sig_filter(dsig, qsig[u])] = True
            assert passed[rows[:, u]].all()


def test_triangles_equal_backtracking():
    for seed in range(3):
        g = F.HostGraph.rmat(12, 30000, seed=40 + seed)
        off, adj, lab = _unlabeled(g)
        z = np.zeros(3, np.int32)
        want = O.match(z, adjm_from_edges(3, [(0, 1), (1, 2), (0, 2)]), off, adj, lab, threads=4)[0]
        assert 6 * O.triangles(off, adj, threads=1 + seed) == want
    n = 9  # K9: C(9,3) triangles
    g = F.HostGraph.from_edges(n, [(u, v) for u in range(n) for v in range(u + 1, n)], [0] * n)
    off, adj, _ = g.arrays()
    assert O.triangles(off, adj) == 84


def test_match_budget_finished_equals_match_and_small_budget_reports():
    """A search that finishes within its step budget counts what match()
    counts; one that does not reports over_budget and no count."""
    g = F.HostGraph.chung_lu(3000, 15000, 2.3, 9).set_labels("uniform", 3, seed=10)
    off, adj, lab = g.arrays()
    for _, q in W.walk_queries(g, 12, 3, 7, 5):
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=4)
        got, stopped, over = O.match_budget(q.labels, q.adjm(), off, adj, lab, threads=4, max_steps=10 ** 12)
        assert (got, stopped, over) == (want, False, False)
    k = 5
    q = F.Query(k, np.zeros(k, np.int32), [(a, b) for a in range(k) for b in range(a + 1, k)])
    u = np.zeros(g.n, np.int32)
    got, stopped, over = O.match_budget(q.labels, q.adjm(), off, adj, u, threads=4, max_steps=200_000)
    assert over and got is None and not stopped
