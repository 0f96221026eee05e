"""Parity of the HIP engine (through the C ABI) with the oracle and the golden
fixtures.  Bit-exact: embedding counts (u64) and the sorted embedding sets."""
import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from conftest import adjm_from_edges, csr_from_edges, load_golden

pytestmark = pytest.mark.gpu


def _sorted_rows(a):
    a = np.asarray(a, dtype=np.uint32)
    if a.size == 0:
        return a.reshape(0, a.shape[1] if a.ndim == 2 else 0)
    return a[np.lexsort(a.T[::-1])]


def _check_set(dg, q, alg, off, adj, lab):
    want_cnt, want_rows = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, cap=10 ** 7)
    r = F.match(dg, q, alg, F.MODE_ENUMERATE, emb_cap=max(want_cnt, 1))
    assert r.count == want_cnt
    assert np.array_equal(_sorted_rows(r.embeddings), _sorted_rows(want_rows))
    rc = F.match(dg, q, alg, F.MODE_COUNT)
    assert rc.count == want_cnt
    return want_cnt


def test_vf2_fixtures_both_algorithms(gpu_ctx):
    cache = {}
    for c in load_golden("vf2_cases.json"):
        key = c["graph"]
        if key not in cache:
            g = F.HostGraph.from_edges(c["n"], c["edges"], c["labels"])
            cache[key] = gpu_ctx.upload(g, pcsr=True)
        dg = cache[key]
        q = F.Query(c["k"], c["qlabels"], c["qedges"])
        for alg in (F.ALG_MMATCH, F.ALG_GSI):
            r = F.match(dg, q, alg, F.MODE_ENUMERATE, emb_cap=max(c["count"], 1))
            assert r.count == c["count"], (c["graph"], c["query"], alg)
            assert _sorted_rows(r.embeddings).tolist() == c["rows"], (c["graph"], c["query"], alg)


def test_c1_diamond_full_enumeration(gpu_ctx):
    """10,000-vertex Chung-Lu (gamma 2.5, 50,000 edges, 8 uniform labels), the
    diamond query from a seeded random walk: full set, alg 1 and alg 2."""
    g = F.HostGraph.chung_lu(10000, 50000, 2.5, 1).set_labels("uniform", 8, seed=2)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True)
    q = g.walk_query(4, 5, 5, seed=3)
    for alg in (F.ALG_MMATCH, F.ALG_GSI):
        assert _check_set(dg, q, alg, off, adj, lab) >= 1


@pytest.mark.parametrize("seed", range(4))
def test_random_walk_queries_labeled(gpu_ctx, seed):
    g = F.HostGraph.chung_lu(20000, 120000, 2.3, 10 + seed).set_labels("zipf", 6, 1.2, 20 + seed)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True)
    for k in (3, 5, 6, 8):
        for dense in (False, True):
            lo, hi = ((3 * k + 1) // 2, 99) if dense else (k - 1, (3 * k - 1) // 2)
            try:
                q = g.walk_query(k, lo, hi, seed=1000 * seed + 10 * k + dense, bias=2 if dense else 0,
                                 max_attempts=20000)
            except F.FsigError:
                continue
            cnt, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, limit=2_000_000)
            if cnt == O.TOO_MANY:
                continue
            for alg in (F.ALG_MMATCH, F.ALG_GSI):
                _check_set(dg, q, alg, off, adj, lab)


def test_unlabeled_patterns_count(gpu_ctx):
    # the house runs on the smaller graph: its count on the larger one is
    # 1.5e10 and takes the oracle minutes
    pats = {"triangle": (3, [(0, 1), (1, 2), (0, 2)]), "cycle4": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
            "clique4": (4, [(a, b) for a in range(4) for b in range(a + 1, 4)]),
            "house": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (1, 4)])}
    for scale, m, names in ((12, 40000, ("triangle", "cycle4", "clique4")), (10, 7000, ("house",))):
        g = F.HostGraph.rmat(scale, m, seed=7)
        off, adj, _ = g.arrays()
        lab = np.zeros(g.n, np.int32)
        dg = gpu_ctx.upload(g, pcsr=True)
        for name in names:
            k, e = pats[name]
            q = F.Query(k, np.zeros(k, np.int32), e)
            want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)
            for alg in (F.ALG_MMATCH, F.ALG_GSI):
                assert F.match(dg, q, alg, F.MODE_COUNT).count == want, (name, alg)
        dg.free()


def _deg(q):
    return [int(bin(int(m)).count("1")) for m in q.adjm()]


@pytest.mark.parametrize("passes", [1, 2, 4])
def test_candidate_sets_equal_oracle_filter_and_refinement(gpu_ctx, passes):
    """Element by element: the label/degree filter's sets equal oracle_ldf and
    the refined sets equal oracle_refine after the same number of passes (the
    reference's filters only)."""
    g = F.HostGraph.chung_lu(8000, 40000, 2.5, 31).set_labels("uniform", 4, seed=32)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    gpu_ctx.set_tune(refine_passes=passes, nlf_filter=1)
    gpu_ctx.keep_candidates(True)
    try:
        shrunk = False
        for s in range(6):
            q = g.walk_query(6, 5, 12, seed=500 + s)
            F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT)
            sizes, sets = O.refine(q.labels, q.adjm(), off, adj, lab, passes, want_sets=True)
            deg = _deg(q)
            for u in range(q.k):
                ldf = gpu_ctx.last_candidate_set(u, 0)
                assert np.array_equal(ldf, O.ldf(off, lab, int(q.labels[u]), deg[u])), (s, u)
                got = gpu_ctx.last_candidate_set(u, 1)
                assert np.array_equal(got, np.nonzero(sets[u])[0]), (s, u)
                shrunk |= len(got) < len(ldf)
            assert np.array_equal(gpu_ctx.last_candidates(q.k), sizes)
        assert shrunk  # refinement removed candidates somewhere
    finally:
        gpu_ctx.keep_candidates(False)
        gpu_ctx.set_tune()


def test_candidate_sets_neighbour_label_filter_bounds(gpu_ctx):
    """With the neighbour-label filter on, each filtered set lies inside the
    label/degree set and holds every embedding's image."""
    g = F.HostGraph.chung_lu(8000, 40000, 2.5, 35).set_labels("uniform", 4, seed=36)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    gpu_ctx.keep_candidates(True)
    try:
        for s in range(4):
            q = g.walk_query(5, 4, 9, seed=600 + s)
            cnt, rows = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, cap=10 ** 7)
            assert F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count == cnt
            deg = _deg(q)
            for u in range(q.k):
                for stage in (0, 1):
                    got = gpu_ctx.last_candidate_set(u, stage)
                    assert np.isin(got, O.ldf(off, lab, int(q.labels[u]), deg[u])).all()
                    assert np.isin(rows[:, u], got).all()
    finally:
        gpu_ctx.keep_candidates(False)


def test_neighbour_label_filter_is_sound_and_prunes(gpu_ctx):
    # u (label 0) has two label-1 neighbours in different classes (a leaf and
    # a vertex with a further neighbour): refinement asks one neighbour in
    # each class, the neighbour-label filter two label-1 neighbours
    g = F.HostGraph.chung_lu(8000, 30000, 2.5, 33).set_labels("uniform", 3, seed=34)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    queries = [F.Query(4, [0, 1, 1, 2], [(0, 1), (0, 2), (2, 3)])]
    queries += [g.walk_query(6, 5, 12, seed=900 + s) for s in range(4)]
    pruned = False
    for q in queries:
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)
        sizes = {}
        for flag in (0, 1):
            gpu_ctx.set_tune(nlf_filter=flag)
            try:
                assert F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count == want
                sizes[flag] = gpu_ctx.last_candidates(q.k)
            finally:
                gpu_ctx.set_tune()
        assert np.all(sizes[0] <= sizes[1])
        pruned |= bool(np.any(sizes[0] < sizes[1]))
    assert pruned
    dg.free()


def test_gsi_candidate_sets_equal_oracle_signature_filter(gpu_ctx):
    """Element by element: alg 2's candidate sets equal oracle_sig_filter."""
    g = F.HostGraph.chung_lu(8000, 40000, 2.5, 41).set_labels("zipf", 5, 1.2, 42)
    dg = gpu_ctx.upload(g, pcsr=True)
    dsig = g.signature(True)
    gpu_ctx.keep_candidates(True)
    try:
        for s in range(4):
            q = g.walk_query(5, 4, 10, seed=700 + s)
            F.match(dg, q, F.ALG_GSI, F.MODE_COUNT)
            qsig = q.signature().reshape(q.k, 16)
            for u in range(q.k):
                assert np.array_equal(gpu_ctx.last_candidate_set(u, 1), O.sig_filter(dsig, qsig[u])), (s, u)
    finally:
        gpu_ctx.keep_candidates(False)


def test_reference_unit_ids_with_colliding_labels(gpu_ctx):
    """PCSR units numbered as the reference numbers them (graph/graph.h:
    107-122), where (1,0) and (0,0), (2,0) and (1,1) share a unit: alg 2's
    sets still equal the oracle's."""
    g = F.HostGraph.chung_lu(6000, 30000, 2.5, 43).set_labels("uniform", 3, seed=44)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True, reference_unit_ids=True)
    pairs = set()
    queries = [F.Query(3, [1, 0, 0], [(0, 1), (1, 2)]), F.Query(3, [2, 0, 1], [(0, 1), (1, 2), (0, 2)]),
               F.Query(4, [1, 1, 0, 2], [(0, 1), (1, 2), (2, 3), (0, 3)])]
    queries += [g.walk_query(5, 4, 8, seed=800 + s) for s in range(4)]
    for q in queries:
        for a, b in q.edges.tolist():
            pairs.add(tuple(sorted((int(q.labels[a]), int(q.labels[b])))))
        want = O.match_digest(q.labels, q.adjm(), off, adj, lab, threads=8)
        if want[0] <= 2_000_000:
            _check_set(dg, q, F.ALG_GSI, off, adj, lab)
        assert F.match_digest(dg, q, F.ALG_GSI) == want
        assert F.match(dg, q, F.ALG_GSI, F.MODE_COUNT).count == want[0]
    assert (0, 1) in pairs and (0, 0) in pairs and (0, 2) in pairs and (1, 1) in pairs


def test_digest_equals_oracle_and_enumerated_rows(gpu_ctx):
    """fsig_match_digest equals the oracle's digest of the same set and the
    digest of the rows fsig_match enumerates, for both algorithms."""
    g = F.HostGraph.chung_lu(20000, 120000, 2.3, 45).set_labels("zipf", 6, 1.2, 46)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True)
    done = 0
    for s, (k, lo, hi) in enumerate([(4, 3, 6), (5, 4, 8), (6, 7, 12), (7, 6, 9)] * 6):
        if done == 4:
            break
        try:
            q = g.walk_query(k, lo, hi, seed=900 + s, max_attempts=20000)
        except F.FsigError:
            continue
        if O.match(q.labels, q.adjm(), off, adj, lab, threads=8, limit=50_000_000)[0] == O.TOO_MANY:
            continue
        print(f"digest: k={k} seed={900 + s}", flush=True)
        want = O.match_digest(q.labels, q.adjm(), off, adj, lab, threads=8)
        assert want[0] > 0
        done += 1
        for alg in (F.ALG_MMATCH, F.ALG_GSI):
            assert F.match_digest(dg, q, alg) == want, (s, alg)
        if want[0] <= 2_000_000:
            r = F.match(dg, q, F.ALG_MMATCH, F.MODE_ENUMERATE, emb_cap=want[0])
            assert O.digest_numpy(r.embeddings) == want[1]
    assert done == 4


def test_edge_cases(gpu_ctx):
    g = F.HostGraph.chung_lu(3000, 12000, 2.5, 51).set_labels("uniform", 3, seed=52)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True)
    # single vertex: every vertex of the label
    q1 = F.Query(1, [1], np.zeros((0, 2)))
    for alg in (F.ALG_MMATCH, F.ALG_GSI):
        assert F.match(dg, q1, alg).count == int((lab == 1).sum())
    # label absent from the data graph
    qx = F.Query(2, [1, 7], [(0, 1)])
    assert F.match(dg, qx, F.ALG_MMATCH).count == 0
    # disconnected query (two edges)
    q2 = F.Query(4, [0, 1, 0, 1], [(0, 1), (2, 3)])
    for alg in (F.ALG_MMATCH, F.ALG_GSI):
        _check_set(dg, q2, alg, off, adj, lab)
    # a clique no vertex can host: empty candidate sets, count 0
    k = 12
    qk = F.Query(k, [0] * k, [(a, b) for a in range(k) for b in range(a + 1, k)])
    assert F.match(dg, qk, F.ALG_MMATCH).count == O.match(qk.labels, qk.adjm(), off, adj, lab, threads=8)[0]
    # malformed queries straight through the C ABI: self-loop, neighbour out
    # of range, asymmetric adjacency, too many vertices
    L = gpu_ctx._L
    cnt = F.C.c_uint64(7)
    labels = np.zeros(33, np.int32)
    for qoff, qadj, k in (([0, 1, 2], [0, 0], 2), ([0, 1, 2], [5, 0], 2), ([0, 1, 1], [1], 2),
                          (list(range(34)), list(range(1, 33)) + [0], 33)):
        qo = np.asarray(qoff, np.uint32)
        qa = np.asarray(qadj, np.uint32)
        rc = L.fsig_match(gpu_ctx._h, dg._h, k, F._ptr(labels, F._i32p), F._ptr(qo), F._ptr(qa),
                          F._ptr(None), F.ALG_MMATCH, F.MODE_COUNT, 0, F.C.byref(cnt), F._ptr(None), 0)
        assert rc in (F.FSIG_EINVAL, F.FSIG_EUNSUPPORTED), (qoff, qadj, rc)


def test_bounded_count(gpu_ctx):
    g = F.HostGraph.chung_lu(5000, 30000, 2.3, 63).set_labels("uniform", 2, seed=64)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True)
    q = F.Query(4, [0, 1, 0, 1], [(0, 1), (1, 2), (2, 3)])
    total, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)
    assert total > 1000
    for alg in (F.ALG_MMATCH, F.ALG_GSI):
        assert F.match(dg, q, alg, F.MODE_COUNT).count == total
        for bound in (1, 999, total - 1, total, total + 1):
            assert F.match(dg, q, alg, F.MODE_COUNT, max_embeddings=bound).count == min(total, bound), (alg, bound)


def test_bounded_enumeration(gpu_ctx):
    g = F.HostGraph.chung_lu(5000, 30000, 2.3, 61).set_labels("uniform", 2, seed=62)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    q = F.Query(3, [0, 1, 0], [(0, 1), (1, 2)])
    total = F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count
    bound = total // 3
    r = F.match(dg, q, F.ALG_MMATCH, F.MODE_ENUMERATE, max_embeddings=bound, emb_cap=bound)
    assert r.count == bound and len(r.embeddings) == bound
    rows = r.embeddings
    assert len({tuple(x) for x in rows.tolist()}) == bound  # distinct
    nb = [set(adj[off[v]:off[v + 1]].tolist()) for v in range(g.n)]
    for a, b, c in rows[:2000].tolist():  # each row is an embedding
        assert lab[a] == 0 and lab[b] == 1 and lab[c] == 0 and a != c and b in nb[a] and c in nb[b]


def test_small_arena_forces_windows():
    """A 24 MiB arena splits levels into many windows; the set must not change."""
    g = F.HostGraph.chung_lu(6000, 40000, 2.3, 71).set_labels("uniform", 2, seed=72)
    off, adj, lab = g.arrays()
    ctx = F.Context(0, 24 << 20)
    try:
        dg = ctx.upload(g)
        q = F.Query(4, [0, 1, 0, 1], [(0, 1), (1, 2), (2, 3)])
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)
        assert F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count == want
        big = F.Context(0, 1 << 30)
        try:
            bdg = big.upload(g)
            r_big = F.match(bdg, q, F.ALG_MMATCH, F.MODE_ENUMERATE, emb_cap=want)
        finally:
            pass
        r = F.match(dg, q, F.ALG_MMATCH, F.MODE_ENUMERATE, emb_cap=want)
        assert r.count == want
        assert np.array_equal(_sorted_rows(r.embeddings), _sorted_rows(r_big.embeddings))
        del bdg
        big.close()
    finally:
        del dg
        ctx.close()


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_lanes_per_row_do_not_change_results(gpu_ctx, lanes):
    g = F.HostGraph.chung_lu(8000, 50000, 2.3, 81).set_labels("uniform", 3, seed=82)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g, pcsr=True)
    q = g.walk_query(5, 5, 9, seed=83)
    want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)
    try:
        gpu_ctx.set_tune(lanes_per_row=lanes)
        for alg in (F.ALG_MMATCH, F.ALG_GSI):
            assert F.match(dg, q, alg, F.MODE_COUNT).count == want
    finally:
        gpu_ctx.set_tune()


def test_hub_candidates_unlabeled_closed_forms(gpu_ctx):
    # R-MAT with hubs far above the 512-entry row limit: candidate edges of
    # the hubs come from segments swept one wavefront each.  Monomorphism
    # counts against closed forms: triangle tr(A^3), 4-cycle tr(A^4) - 2 sum
    # d^2 + 2m
    import scipy.sparse as sp
    g = F.HostGraph.rmat(12, 150000, seed=27)
    assert g.max_degree > 4 * 512
    off, adj, _ = g.arrays()
    A = sp.csr_matrix((np.ones(len(adj)), adj.astype(np.int64), off.astype(np.int64)), shape=(g.n, g.n))
    A2 = A @ A
    d = np.diff(off).astype(np.float64)
    want = {"triangle": int(round(A2.multiply(A).sum())),
            "cycle4": int(round(A2.multiply(A2).sum() - 2 * (d * d).sum() + d.sum()))}
    dg = gpu_ctx.upload(g)
    qs = {"triangle": F.Query(3, [0, 0, 0], [(0, 1), (1, 2), (0, 2)]),
          "cycle4": F.Query(4, [0] * 4, [(0, 1), (1, 2), (2, 3), (3, 0)])}
    for name, q in qs.items():
        assert F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count == want[name], name
    counts, _ = F.match_batch(dg, list(qs.values()))
    assert counts.tolist() == [want["triangle"], want["cycle4"]]
    dg.free()
