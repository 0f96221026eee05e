"""The batched engine (fsig_match_batch): every query's count equals its
plain fsig_match count and the oracle's, for a mix of sizes, densities,
disconnected queries, absent labels and empty candidate sets."""
import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from fsig_amd import workloads as W

pytestmark = pytest.mark.gpu


def _mixed_queries(g, n, seed):
    qs = [q for _, q in W.walk_queries(g, n, 3, 9, seed)]
    qs.append(F.Query(1, [2], np.zeros((0, 2))))                      # single vertex
    qs.append(F.Query(4, [0, 1, 0, 1], [(0, 1), (2, 3)]))             # disconnected
    qs.append(F.Query(3, [0, 99, 1], [(0, 1), (1, 2)]))               # label absent from the graph
    qs.append(F.Query(2, [0x7fffffff, 1], [(0, 1)]))                  # label far past the inverted index
    k = 9
    qs.append(F.Query(k, [0] * k, [(a, b) for a in range(k) for b in range(a + 1, k)]))  # no host
    qs.append(F.Query(2, [1, 1], [(0, 1)]))
    return qs


def _bounded(dg, qs, limit=50_000_000):
    """Queries whose plain count stays under `limit`, with those counts (the
    bounded count stops early on huge queries)."""
    out = []
    for q in qs:
        c = F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=limit + 1).count
        print("plain", q.k, q.num_edges, c, flush=True)
        if c <= limit:
            out.append((q, c))
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_equals_plain_and_oracle(gpu_ctx, seed):
    g = F.HostGraph.chung_lu(20000, 100000, 2.3, 90 + seed).set_labels("zipf", 8, 1.2, 95 + seed)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    qs = _bounded(dg, _mixed_queries(g, 40, seed))
    counts, join_ms = F.match_batch(dg, [q for q, _ in qs])
    assert join_ms >= 0
    for (q, plain), c in zip(qs, counts.tolist()):
        assert c == plain, (q.k, q.labels.tolist(), q.edges.tolist())
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, limit=20_000_000)
        if want != O.TOO_MANY:
            assert c == want


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("strategy", [0, 1, 2])
def test_batch_configurations_agree(gpu_ctx, lanes, strategy):
    g = F.HostGraph.chung_lu(15000, 90000, 2.3, 7).set_labels("uniform", 4, seed=8)
    dg = gpu_ctx.upload(g)
    qs, ref = zip(*_bounded(dg, [q for _, q in W.walk_queries(g, 24, 3, 8, 11)], 5_000_000))
    try:
        gpu_ctx.set_tune(lanes_per_row=lanes, strategy=strategy, lds_tile=128)
        counts, _ = F.match_batch(dg, qs)
        assert counts.tolist() == list(ref)
        got_lanes, _, engine = F.last_batch_tune(gpu_ctx)
        assert got_lanes == lanes
        assert engine == ("lockstep" if strategy == 0 else "persistent")
    finally:
        gpu_ctx.set_tune()


def test_persistent_graph_replays_after_updates(gpu_ctx):
    """The persistent walk keeps its instantiated graph per engine shape and
    updates it from each new capture: replays with other batches (other
    tables, queue sizes and counts), eager plain matches and
    synchronisations in between, each replay equal to the plain counts."""
    g = F.HostGraph.chung_lu(15000, 90000, 2.3, 17).set_labels("uniform", 4, seed=18)
    dg = gpu_ctx.upload(g)
    batches = [_bounded(dg, [q for _, q in W.walk_queries(g, n, 3, 8, seed)], 5_000_000)
               for n, seed in ((24, 21), (10, 22), (31, 23))]
    try:
        gpu_ctx.set_tune(lanes_per_row=16, strategy=1)
        for rep in range(2):
            for qs in batches:
                counts, _ = F.match_batch(dg, [q for q, _ in qs])
                assert counts.tolist() == [c for _, c in qs], rep
                assert F.last_batch_tune(gpu_ctx)[2] == "persistent"
                q0, c0 = qs[0]
                assert F.match(dg, q0, F.ALG_MMATCH, F.MODE_COUNT).count == c0  # eager work between replays
                gpu_ctx.synchronize()
    finally:
        gpu_ctx.set_tune()


def test_batch_autotuner_and_repeat(gpu_ctx):
    g = F.HostGraph.chung_lu(30000, 200000, 2.3, 12).set_labels("zipf", 16, 1.2, 13)
    dg = gpu_ctx.upload(g)
    qs, ref = zip(*_bounded(dg, [q for _, q in W.walk_queries(g, 64, 4, 12, 14)]))
    a, _ = F.match_batch(dg, qs)
    b, _ = F.match_batch(dg, qs)   # second call reuses the tuning decision
    assert a.tolist() == b.tolist()
    lanes, tile, engine = F.last_batch_tune(gpu_ctx)
    # lanes 0: the lockstep join with each query's own lane count
    assert lanes in (0, 4, 8, 16, 32, 64) and engine in ("lockstep", "persistent")
    assert a.tolist() == list(ref)
    # an empty batch returns at once and keeps the tuning decision
    e, ms = F.match_batch(dg, [])
    assert len(e) == 0 and ms == 0.0
    assert F.last_batch_tune(gpu_ctx) == (lanes, tile, engine)


def test_batch_lanes_per_query(gpu_ctx):
    g = F.HostGraph.chung_lu(30000, 200000, 2.3, 15).set_labels("zipf", 16, 1.2, 16)
    dg = gpu_ctx.upload(g)
    qs, ref = zip(*_bounded(dg, [q for _, q in W.walk_queries(g, 48, 3, 12, 17)]))
    try:
        gpu_ctx.set_tune(strategy=3)
        counts, _ = F.match_batch(dg, qs)
        assert F.last_batch_tune(gpu_ctx)[0] == 0
    finally:
        gpu_ctx.set_tune()
    assert counts.tolist() == list(ref)


def test_batch_hub_candidates(gpu_ctx):
    # R-MAT hubs (max degree 3201, far above the 512-entry row limit): their
    # candidate-edge lists come from segments swept one wavefront each;
    # counts must equal the plain path's
    g = F.HostGraph.rmat(13, 200000, seed=21).set_labels("uniform", 3, seed=22)
    assert g.max_degree > 4 * 512
    dg = gpu_ctx.upload(g)
    qs, ref = zip(*_bounded(dg, [q for _, q in W.walk_queries(g, 24, 3, 6, 23)], 20_000_000))
    off, adj, lab = g.arrays()
    for q, c in list(zip(qs, ref))[:6]:  # the plain path's hub lists against the oracle
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, limit=10 ** 8)
        if want != O.TOO_MANY:
            assert c == want
    for strategy in (3, 1):  # lockstep and persistent joins over the same lists
        try:
            gpu_ctx.set_tune(lanes_per_row=16 if strategy == 1 else 0, strategy=strategy)
            counts, _ = F.match_batch(dg, qs)
        finally:
            gpu_ctx.set_tune()
        assert counts.tolist() == list(ref), strategy
    dg.free()


@pytest.mark.parametrize("beam_log2", [0, 4])
def test_bounded_batch_equals_bounded_plain_and_oracle(gpu_ctx, beam_log2):
    """fsig_match_batch_bounded: min(count, bound) per query, equal to
    fsig_match's bounded count and to the oracle (stopped searches prove
    the count exceeds the bound).  beam_log2 = 4 keeps a 16-child beam of
    every bounded query in the lockstep join, so most of them end below their
    bound and have the subtrees of their cut rows counted by the depth-first
    driver; exact queries (bound 0) are never beamed."""
    g = F.HostGraph.chung_lu(20000, 120000, 2.3, 97).set_labels("zipf", 6, 1.2, 98)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    qs = [q for _, q in W.walk_queries(g, 40, 3, 12, 21)] + _mixed_queries(g, 0, 3)
    rng = np.random.default_rng(5)
    bounds = [int(b) for b in rng.choice([0, 1, 1000, 10 ** 5, 10 ** 6], size=len(qs))]
    for i, q in enumerate(qs):  # exact queries must stay small enough for the oracle
        if bounds[i] == 0 and F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=10 ** 7 + 1).count > 10 ** 7:
            bounds[i] = 10 ** 6
    try:
        gpu_ctx.set_tune(beam_log2=beam_log2)
        counts, _, st = F.match_batch(dg, qs, bounds=bounds, status=True)
    finally:
        gpu_ctx.set_tune()
    assert not st.any()
    for q, b, c in zip(qs, bounds, counts.tolist()):
        plain = F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=b).count
        assert c == plain, (q.k, q.edges.tolist(), b)
        want, _, stopped = O.match_partial(q.labels, q.adjm(), off, adj, lab, threads=8, limit=b)
        if b == 0 or not stopped:
            assert c == (min(want, b) if b else want), (q.k, b)
        else:
            assert c == b  # the oracle saw more than b embeddings


def test_batch_bitmaps_built_query_by_query(gpu_ctx):
    """On a graph whose batch of candidate bitmaps outgrows the last-level
    cache (4M vertices: 512 KB per target class, about 900 classes here),
    the candidate-edge build runs query by query with each query's bitmaps
    and vertex -> candidate index maps rebuilt in one region: counts equal
    fsig_match's, bounded and exact."""
    g = F.HostGraph.rmat(22, 12_000_000, seed=41).set_labels("uniform", 16, seed=42)
    ctx = F.Context(0, 32 << 30)  # 128 class masks of 16 MB alone outgrow the shared test context
    dg = ctx.upload(g)
    qs = [q for _, q in W.walk_queries(g, 128, 4, 10, 43)]
    bound = 10 ** 6
    plain = [F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=bound).count for q in qs]
    counts, _, st = F.match_batch(dg, qs, bounds=bound, status=True)
    assert not st.any()
    assert counts.tolist() == plain
    small = [i for i, c in enumerate(plain) if c < bound]  # exact counts below the bound
    counts2, _ = F.match_batch(dg, [qs[i] for i in small])
    assert counts2.tolist() == [plain[i] for i in small]
    dg.free()
    ctx.close()


def test_bounded_work_limit_sweep(gpu_ctx):
    """Under work limits from 1 to 2^30 rows, a bounded count (fsig_match with
    alg 1 and alg 2, and the bounded batch with a 16-row beam, whose resume
    counts the cut rows under the limit) is min(count, bound) or refused with
    FSIG_ELIMIT: the
    last window before the limit is clipped to the rows it leaves, so limits
    that fall inside a window run to the bound or refuse, never anything
    else.  The smallest limits refuse, the largest admit."""
    g = F.HostGraph.chung_lu(20000, 160000, 2.2, 99).set_labels("uniform", 2, seed=100)
    dg = gpu_ctx.upload(g, pcsr=True)
    q = F.Query(6, [0, 1, 0, 1, 0, 1], [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)])
    bound = 10 ** 7
    want = F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=bound).count
    assert want == bound  # the bound is reached: the run stops early
    plain, gsi, batch = [], [], []
    try:
        gpu_ctx.set_tune(beam_log2=4)
        for j in range(31):
            gpu_ctx.set_work_limit(1 << j)
            for alg, out in ((F.ALG_MMATCH, plain), (F.ALG_GSI, gsi)):
                try:
                    out.append(F.match(dg, q, alg, F.MODE_COUNT, max_embeddings=bound).count)
                except F.FsigError as e:
                    assert e.code == F.FSIG_ELIMIT
                    out.append(None)
            counts, _, st = F.match_batch(dg, [q], bounds=[bound], status=True)
            assert st[0] in (0, F.FSIG_ELIMIT)
            batch.append(int(counts[0]) if st[0] == 0 else None)
    finally:
        gpu_ctx.set_work_limit(0)
        gpu_ctx.set_tune()
    for got in (plain, gsi, batch):
        assert got[0] is None and got[-1] == want
        assert all(c in (None, want) for c in got)
    print("admitted from limit 2^%d (alg 1), 2^%d (alg 2), 2^%d (batch)"
          % (plain.index(want), gsi.index(want), batch.index(want)), flush=True)
    dg.free()


def test_bounded_batch_work_limit_refuses_per_query(gpu_ctx):
    """A bounded query over the context's work limit is refused alone
    (FSIG_ELIMIT in its status); the others complete."""
    g = F.HostGraph.chung_lu(20000, 160000, 2.2, 99).set_labels("uniform", 2, seed=100)
    dg = gpu_ctx.upload(g)
    heavy = F.Query(6, [0, 1, 0, 1, 0, 1], [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)])
    light = F.Query(2, [0, 1], [(0, 1)])
    want_light = F.match(dg, light, F.ALG_MMATCH, F.MODE_COUNT).count
    try:
        gpu_ctx.set_tune(beam_log2=4)
        gpu_ctx.set_work_limit(1000)
        counts, _, st = F.match_batch(dg, [heavy, light], bounds=[10 ** 12, 10 ** 12], status=True)
    finally:
        gpu_ctx.set_work_limit(0)
        gpu_ctx.set_tune()
    assert st.tolist() == [F.FSIG_ELIMIT, 0]
    assert counts.tolist() == [0, want_light]
