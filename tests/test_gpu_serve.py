"""The served loop (fsig_server_*): every query's count is the same when it
runs alone (fsig_match) and when it is submitted in a stream of other
queries, whatever batch and worker it lands in; label-pair statistics equal
a numpy count; a tiny budget forces one-query batches and the plain
fallback without changing any result."""
import threading
import time

import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from fsig_amd import workloads as W

pytestmark = pytest.mark.gpu


def _graph(seed):
    return F.HostGraph.chung_lu(20000, 100000, 2.3, 120 + seed).set_labels("zipf", 8, 1.2, 130 + seed)


def _queries(g, dg, n, seed):
    qs = [q for _, q in W.walk_queries(g, n, 3, 12, seed, dg=dg, max_count=2_000_000, max_rows=4_000_000)]
    qs.append(F.Query(1, [2], np.zeros((0, 2))))
    qs.append(F.Query(3, [0, 99, 1], [(0, 1), (1, 2)]))  # label absent: count 0
    qs.append(F.Query(2, [0x7fffffff, 1], [(0, 1)]))  # label far past the inverted index: count 0
    qs.append(F.Query(4, [0, 1, 0, 1], [(0, 1), (2, 3)]))  # disconnected
    return qs


def test_label_pairs_equal_numpy(gpu_ctx):
    g = _graph(0)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    got = F.label_pairs(dg)
    want = np.zeros((g.num_labels, g.num_labels), np.uint64)
    src = np.repeat(np.arange(g.n), np.diff(off.astype(np.int64)))
    np.add.at(want, (lab[src], lab[adj]), 1)
    assert np.array_equal(got, want)
    dg.free()


@pytest.mark.parametrize("workers,budget,max_batch", [(4, 1 << 30, 16), (2, 1 << 30, 64), (1, 1 << 30, 1)])
def test_served_equals_alone(gpu_ctx, workers, budget, max_batch):
    g = _graph(1)
    dg = gpu_ctx.upload(g)
    qs = _queries(g, dg, 40, seed=7)
    alone = [F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count for q in qs]
    print("alone", alone, flush=True)
    with F.Server(dg, workers=workers, hbm_budget=budget, max_batch=max_batch) as srv:
        # staggered arrivals from a second thread while results are collected
        tickets = [None] * len(qs)
        rng = np.random.default_rng(3)

        def feed():
            for i, q in enumerate(qs):
                tickets[i] = srv.submit(q)
                time.sleep(float(rng.exponential(0.002)))

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        got = [srv.wait(t, 60_000)[0] for t in tickets]
        st = srv.stats()
    print("stats", st, flush=True)
    assert got == alone
    assert st["completed"] == len(qs) and st["submitted"] == len(qs)
    assert st["max_batch_seen"] <= max_batch
    dg.free()


def test_served_tiny_budget_falls_back(gpu_ctx):
    # 4 MiB arenas: small batches, and a batch that runs out of arena is
    # counted query by query by the plain memory-blocked path
    g = _graph(2)
    off, adj, lab = g.arrays()
    dg = gpu_ctx.upload(g)
    qs = _queries(g, dg, 12, seed=9)
    alone = [F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count for q in qs]
    with F.Server(dg, workers=2, hbm_budget=8 << 20, max_batch=8) as srv:
        tickets = [srv.submit(q) for q in qs]
        got = [srv.wait(t, 60_000)[0] for t in tickets]
        st = srv.stats()
    print("stats", st, flush=True)
    assert got == alone
    assert st["completed"] == len(qs)
    for q, c in list(zip(qs, alone))[:4]:
        want, _ = O.match(q.labels, q.adjm(), off, adj, lab, threads=8)
        assert c == want
    dg.free()


def test_server_unknown_ticket_and_projection(gpu_ctx):
    g = _graph(3)
    dg = gpu_ctx.upload(g)
    with F.Server(dg, workers=1, hbm_budget=1 << 28, max_batch=4) as srv:
        with pytest.raises(F.FsigError) as e:
            srv.wait(12345, 0)
        assert e.value.code == F.FSIG_ENOTFOUND
        q = g.walk_query(5, 4, 10, seed=5)
        off, adj, lab = g.arrays()
        lc = np.bincount(lab, minlength=g.num_labels).astype(np.uint64)
        assert srv.projected(q) == F.project_bytes(g.n, lc, F.label_pairs(dg), q)
        t = srv.submit(q)
        assert srv.wait(t, 60_000)[0] == F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count
    dg.free()


def test_server_refuses_more_than_4096_labels(gpu_ctx):
    g = F.HostGraph.chung_lu(12000, 40000, 2.3, 8).set_labels("uniform", 5000, seed=9)
    dg = gpu_ctx.upload(g)
    with pytest.raises(F.FsigError) as e:
        F.Server(dg, workers=1, hbm_budget=1 << 28, max_batch=4)
    assert e.value.code == F.FSIG_EINVAL
    with pytest.raises(F.FsigError):
        F.label_pairs(dg)
    # the graph itself still matches
    q = g.walk_query(3, 2, 3, seed=4)
    assert F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count >= 1
    dg.free()


def test_served_bounded_equals_alone(gpu_ctx):
    """Bounded requests: the served count is fsig_match's min(count, bound)."""
    g = _graph(3)
    dg = gpu_ctx.upload(g)
    qs = [q for _, q in W.walk_queries(g, 30, 3, 12, 17)]
    bounds = [[0, 10, 10 ** 4, 10 ** 6][i % 4] for i in range(len(qs))]
    for i, q in enumerate(qs):
        if bounds[i] == 0 and F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=10 ** 8 + 1).count > 10 ** 8:
            bounds[i] = 10 ** 6
    alone = [F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=b).count for q, b in zip(qs, bounds)]
    with F.Server(dg, workers=3, hbm_budget=1 << 30, max_batch=8) as srv:
        tickets = [srv.submit(q, b) for q, b in zip(qs, bounds)]
        got = [srv.wait(t)[0] for t in tickets]
    assert got == alone
