"""Host side of the drop-in: reader, CSR, inverted index, PCSR, signatures,
MurmurHash2 and the generators, against the reference's own host code
(oracle/_ref/ref_dump, compiled from the reference sources) and against the
committed fixtures it produced (tests/golden/ref_tiny.json)."""
import os
import tempfile

import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from conftest import load_golden


def test_ref_fixtures_csr_inverted_signature_murmur():
    for case in load_golden("ref_tiny.json"):
        g = F.HostGraph.from_edges(case["n"], case["edges"], case["labels"])
        off, adj, lab = g.arrays()
        assert off.tolist() == case["OFF"]
        assert adj.tolist() == case["ADJ"]
        idx, lst = g.inverted_index()
        assert idx.tolist() == case["INVIDX"] and lst.tolist() == case["INVLIST"]
        assert g.signature(False).tolist() == case["SIGQ"]
        assert g.signature(True).tolist() == case["SIGD"]
        # the oracle's restatement agrees with the reference too
        assert O.signature(off, adj, lab, True).tolist() == case["SIGD"]
        hashes = [F.host_lib().fsig_murmur2(np.array([i], np.uint32).ctypes.data, 4, 17) for i in range(256)]
        assert hashes == case["MURMUR"]
        assert [O.murmur2_u32(i) for i in range(256)] == case["MURMUR"]


def test_ref_fixtures_pcsr_layout():
    for case in load_golden("ref_tiny.json"):
        g = F.HostGraph.from_edges(case["n"], case["edges"], case["labels"])
        p = g.pcsr(reference_unit_ids=True)
        assert p.num_units == len(case["units"])
        for u, ref in enumerate(case["units"]):
            o, e = p.unit(u)
            assert o.tolist() == ref["offset"]
            assert e.tolist() == ref["edges"]


def _pcsr_lookup(o, e, key):
    """Bucket walk of the PCSR layout (graph/graph.h:90-95)."""
    K = len(o) // 32
    if K == 0:
        return None
    grp = O.murmur2_u32(key) % K
    for _ in range(K + 1):
        words = o[32 * grp: 32 * grp + 32]
        for s in range(15):
            if words[2 * s] == 0xFFFFFFFF:
                return None
            if words[2 * s] == key:
                return e[words[2 * s + 1]: words[2 * s + 3]]
        grp = words[30]
        if grp == 0xFFFFFFFF:
            return None
    return None


@pytest.mark.parametrize("seed", [1, 2])
def test_pcsr_injective_units_hold_label_partitioned_lists(seed):
    g = F.HostGraph.chung_lu(3000, 15000, 2.3, seed).set_labels("zipf", 6, 1.2, seed)
    off, adj, lab = g.arrays()
    p = g.pcsr(reference_unit_ids=False)
    L = g.num_labels
    rng = np.random.default_rng(seed)
    for v in rng.choice(g.n, 200, replace=False):
        for l2 in range(L):
            hi, lo = max(lab[v], l2), min(lab[v], l2)
            o, e = p.unit(hi * (hi + 1) // 2 + lo)
            want = [int(x) for x in adj[off[v]:off[v + 1]] if lab[x] == l2]
            got = _pcsr_lookup(o, e, int(v))
            assert (list(got) if got is not None else []) == want


def test_generated_graph_matches_reference_on_file():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    g = F.HostGraph.chung_lu(5000, 20000, 2.5, 11).set_labels("zipf", 8, 1.2, 12)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "g.graph")
        g.write(path)
        r = O.ref_dump(path, True, 16)
        g2 = F.HostGraph.read(path, True)
    off, adj, lab = g.arrays()
    off2, adj2, lab2 = g2.arrays()
    assert np.array_equal(off, off2) and np.array_equal(adj, adj2) and np.array_equal(lab, lab2)
    assert np.array_equal(r["OFF"], off) and np.array_equal(r["ADJ"], adj)
    assert np.array_equal(r["SIGD"], g.signature(True))
    p = g.pcsr(reference_unit_ids=True)
    for u in range(p.num_units):
        o, e = p.unit(u)
        assert np.array_equal(o, r["units"][u]["offset"]) and np.array_equal(e, r["units"][u]["edges"])


def test_threaded_builders_match_reference_on_file(tmp_path):
    # 1.2M adjacency entries: the PCSR and signature builders run on threads
    # (above 2^20 entries); their output still equals the reference's own
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    g = F.HostGraph.chung_lu(60000, 600000, 2.3, 21).set_labels("zipf", 8, 1.2, 22)
    assert g.nnz > (1 << 20)
    path = str(tmp_path / "big.graph")
    g.write(path)
    r = O.ref_dump(path, True, 16)
    assert np.array_equal(r["SIGD"], g.signature(True))
    p = g.pcsr(reference_unit_ids=True)
    assert p.num_units == len(r["units"])
    for u in range(p.num_units):
        o, e = p.unit(u)
        assert np.array_equal(o, r["units"][u]["offset"]) and np.array_equal(e, r["units"][u]["edges"])


def _write(td, text):
    path = os.path.join(td, "x.graph")
    with open(path, "w") as f:
        f.write(text)
    return path


def test_reader_trailing_newline_adds_no_phantom_edge():
    with tempfile.TemporaryDirectory() as td:
        p = _write(td, "t 3 2\nv 0 1 1\nv 1 2 2\nv 2 1 1\ne 0 1\ne 1 2\n\n")
        g = F.HostGraph.read(p, True)
        assert g.nnz == 4 and g.n == 3
        assert g.arrays()[2].tolist() == [1, 2, 1]
        g0 = F.HostGraph.read(p, False)
        assert g0.arrays()[2].tolist() == [0, 0, 0]


def test_reader_errors():
    with tempfile.TemporaryDirectory() as td:
        with pytest.raises(F.FsigError) as e:
            F.HostGraph.read(os.path.join(td, "missing.graph"))
        assert e.value.code == -1
        with pytest.raises(F.FsigError) as e:
            F.HostGraph.read(_write(td, "x 1 2\n"))
        assert e.value.code == -2
        with pytest.raises(F.FsigError) as e:  # 2^64 + 1 must not wrap to vertex 1
            F.HostGraph.read(_write(td, "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 18446744073709551617"))
        assert e.value.code == -3
        with pytest.raises(F.FsigError) as e:  # empty file: no header
            F.HostGraph.read(_write(td, ""))
        assert e.value.code == -2
        with pytest.raises(F.FsigError) as e:  # a directory cannot be read as a graph
            F.HostGraph.read(td)
        assert e.value.code == -1
        with pytest.raises(F.FsigError) as e:  # id 1 has no v line
            F.HostGraph.read(_write(td, "t 3 1\nv 0 0 1\nv 2 0 1\ne 0 2"))
        assert e.value.code == -4
        with pytest.raises(F.FsigError) as e:  # endpoint without v line
            F.HostGraph.read(_write(td, "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 5"))
        assert e.value.code == -5
        with pytest.raises(F.FsigError) as e:  # a huge id: rejected before any array is sized from it
            F.HostGraph.read(_write(td, "t 2 1\nv 0 0 1\nv 4000000000 0 1\ne 0 4000000000"))
        assert e.value.code == -4
        # a duplicated v line does not hide a gap
        with pytest.raises(F.FsigError) as e:
            F.HostGraph.read(_write(td, "t 3 1\nv 0 0 1\nv 0 0 1\nv 2 0 1\ne 0 2"))
        assert e.value.code == -4


def test_generators_are_deterministic_and_sized():
    a = F.HostGraph.chung_lu(2000, 9000, 2.3, 5)
    b = F.HostGraph.chung_lu(2000, 9000, 2.3, 5)
    assert a.nnz == 2 * 9000 and np.array_equal(a.arrays()[1], b.arrays()[1])
    off, adj, _ = a.arrays()
    assert off[-1] - off[-2] > 0  # max id is not isolated
    r = F.HostGraph.rmat(12, 20000, seed=3)
    assert r.n == 4096 and r.nnz == 40000
    roff = r.arrays()[0]
    assert roff[-1] - roff[-2] > 0
    z = a.set_labels("zipf", 16, 1.2, 9).arrays()[2]
    freq = np.bincount(z, minlength=16)
    assert freq[0] > freq[1] > freq[3] > freq[15]


def test_walk_query_is_induced_and_has_a_witness():
    g = F.HostGraph.chung_lu(10000, 50000, 2.5, 1).set_labels("uniform", 8, seed=2)
    off, adj, lab = g.arrays()
    q = g.walk_query(4, 5, 5, seed=3)
    assert q.num_edges == 5  # 4 vertices + 5 edges = the diamond
    w = q.walk
    for i in range(4):
        assert lab[w[i]] == q.labels[i]
        for j in range(i + 1, 4):
            has = w[j] in set(adj[off[w[i]]:off[w[i] + 1]].tolist())
            assert has == ((i, j) in {tuple(e) for e in q.edges.tolist()})


def test_large_file_parsed_in_slices(tmp_path):
    # > 16 MiB: the reader parses slices in threads; the graph must equal a
    # generated one, and an error in the last slice must still be reported
    g = F.HostGraph.chung_lu(200000, 1500000, 2.3, 41).set_labels("zipf", 8, 1.2, 42)
    p = str(tmp_path / "big.graph")
    g.write(p)
    assert os.path.getsize(p) > (16 << 20)
    h = F.HostGraph.read(p)
    for a, b in zip(g.arrays(), h.arrays()):
        assert np.array_equal(a, b)
    with open(p, "a") as f:
        f.write("\ne 7 x")
    with pytest.raises(F.FsigError) as e:
        F.HostGraph.read(p)
    assert e.value.code == -3  # truncated or malformed record


def test_walk_bias_direction():
    # bias > 0 steps towards high-degree neighbours, bias < 0 towards low
    g = F.HostGraph.chung_lu(20000, 100000, 2.2, 61)
    off, _, _ = g.arrays()
    deg = np.diff(off.astype(np.int64))
    mean = {}
    for bias in (-4, 0, 4):
        ds = []
        for s in range(40):
            q = g.walk_query(6, 5, 15, seed=1000 + s, bias=bias, max_attempts=2000)
            ds.extend(deg[q.walk].tolist())
        mean[bias] = float(np.mean(ds))
    assert mean[-4] < mean[0] < mean[4]
    # the same seed and bias give the same query (starts are cached per graph)
    a = g.walk_query(6, 5, 15, seed=7, bias=-2)
    b = g.walk_query(6, 5, 15, seed=7, bias=-2)
    assert a.edges.tolist() == b.edges.tolist() and list(a.labels) == list(b.labels)


def test_pcsr_label_cap():
    g = F.HostGraph.chung_lu(6000, 20000, 2.3, 3).set_labels("uniform", 5000, seed=2)
    with pytest.raises(F.FsigError):
        g.pcsr()
    assert F.HostGraph.chung_lu(600, 2000, 2.3, 3).set_labels("uniform", 64, seed=2).pcsr().num_units == 64 * 65 // 2


@pytest.mark.parametrize("n", [0, 1, 1000, 300000])
def test_embedding_file_is_sorted_canonical_text(tmp_path, n):
    # fsig_write_embeddings: "#embedding: <count>" then every row, sorted
    # lexicographically, numbers separated by one space (300,000 rows take
    # the threaded sort and formatting)
    import ctypes as C
    k = 4
    rng = np.random.default_rng(n + 1)
    rows = rng.integers(0, 1 << 32, size=(n, k), dtype=np.uint64).astype(np.uint32)
    if n > 10:
        rows[: n // 10] = rows[0]                # equal rows
        rows[n // 10: n // 5, :2] = rows[1, :2]  # equal prefixes
    path = str(tmp_path / "emb.txt")
    rc = F.host_lib().fsig_write_embeddings(path.encode(), k, 12345, rows.ctypes.data_as(C.POINTER(C.c_uint32)), n)
    assert rc == 0
    want = "#embedding: 12345\n" + "".join(" ".join(str(int(x)) for x in r) + "\n"
                                          for r in sorted(map(tuple, rows.tolist())))
    assert open(path).read() == want


def test_threaded_pcsr_of_an_unlabeled_graph_holds_every_list():
    # one unit and more than 2^20 adjacency entries: the threaded sweep with a
    # single unit; every sampled key's list is its whole adjacency
    g = F.HostGraph.rmat(16, 600000, seed=5)
    assert g.nnz > (1 << 20)
    off, adj, _ = g.arrays()
    p = g.pcsr(reference_unit_ids=True)
    assert p.num_units == 1
    o, e = p.unit(0)
    rng = np.random.default_rng(3)
    for v in rng.choice(g.n, 300, replace=False):
        got = _pcsr_lookup(o, e, int(v))
        assert (list(got) if got is not None else []) == adj[off[v]:off[v + 1]].tolist()


def test_signature_with_large_label_values_equals_the_oracle():
    # labels at or past 2^20 take the path without a per-label slot table
    rng = np.random.default_rng(4)
    n = 3000
    edges = rng.integers(0, n, size=(15000, 2))
    edges = edges[edges[:, 0] != edges[:, 1]]
    labels = (rng.integers(0, 50, size=n) * 100003 + (1 << 20)).astype(np.int32)
    g = F.HostGraph.from_edges(n, edges, labels)
    off, adj, lab = g.arrays()
    for col in (True, False):
        assert np.array_equal(g.signature(col), O.signature(off, adj, lab, col))


@pytest.mark.parametrize("k", [1, 32])
def test_embedding_file_widths(tmp_path, k):
    import ctypes as C
    rng = np.random.default_rng(k)
    rows = rng.integers(0, 1000, size=(5000, k), dtype=np.uint64).astype(np.uint32)
    path = str(tmp_path / "emb.txt")
    rc = F.host_lib().fsig_write_embeddings(path.encode(), k, 5000, rows.ctypes.data_as(C.POINTER(C.c_uint32)), 5000)
    assert rc == 0
    want = "#embedding: 5000\n" + "".join(" ".join(map(str, r)) + "\n" for r in sorted(map(tuple, rows.tolist())))
    assert open(path).read() == want


def test_embedding_file_errors(tmp_path):
    import ctypes as C
    rows = np.zeros((4, 3), np.uint32)
    ptr = rows.ctypes.data_as(C.POINTER(C.c_uint32))
    # a directory cannot be opened for writing
    assert F.host_lib().fsig_write_embeddings(str(tmp_path).encode(), 3, 4, ptr, 4) == -1
    if os.path.exists("/dev/full"):  # every write to it fails
        assert F.host_lib().fsig_write_embeddings(b"/dev/full", 3, 4, ptr, 4) == -6


def test_graph_file_write_errors(tmp_path):
    g = F.HostGraph.chung_lu(2000, 8000, 2.5, 5).set_labels("uniform", 3, seed=6)
    with pytest.raises(F.FsigError) as e:
        g.write(str(tmp_path))  # a directory cannot be opened for writing
    assert e.value.code == -1
    if os.path.exists("/dev/full"):  # every write to it fails: the file would be truncated
        with pytest.raises(F.FsigError) as e:
            g.write("/dev/full")
        assert e.value.code == -6
    path = str(tmp_path / "ok.graph")
    g.write(path)
    assert F.HostGraph.read(path).n == g.n
