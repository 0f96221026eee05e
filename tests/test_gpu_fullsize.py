"""Parity at the workloads' full sizes, through the C ABI.

  C2  dblp-shaped Chung-Lu (317,080 V / 1,049,866 E, 16 Zipf(1.2) labels), the
      8-vertex dense and sparse random-walk queries: alg 1 and alg 2 counts
      equal the oracle's, and so do the order-independent digests of the
      full embedding sets (every embedding enumerated on the device and on
      the host).
  C3  unlabeled R-MAT (0.57, 0.19, 0.19), 2^22 V / 35M E: the triangle count
      equals the oracle's (orientation count); the 4-clique count is the same
      through the fused last two levels and level by level (both with the
      symmetry conditions), and so is the 4-cycle count on the scale-20
      graph of the same generator and density; the first 10^8 embeddings of
      the triangle, 4-cycle, 4-clique and house are each checked row by row
      (labels, injectivity, every query edge) and for distinctness.
  C4  Chung-Lu 1M V / 10M E, 32 Zipf(1.2) labels, the 256 walk queries of the
      workload law counted to 1e8 under the 2e9-row work limit: the bounded
      batch answers every query the plain loop answers, with the same count;
      the oracle agrees on every query whose search ends within 5*10^8 tried
      candidates (the others are skipped and the number checked printed).
  C5  R-MAT 2^24 V / 200M E, 64 labels, the drawn 128-request stream: each
      served result equals the same query alone.  (The oracle finishes none
      of these requests within the step budget: its backtracking filters
      nothing on a 16M-vertex graph; the engine's counts are pinned at C2-C4.)

Each step prints what it runs before it runs it."""
import os
import time

import numpy as np
import pytest

import fsig_amd as F
import oracle as O
from fsig_amd import workloads as W

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)


@pytest.fixture(scope="module")
def big_ctx():
    ctx = F.Context(0, 96 << 30)
    yield ctx
    ctx.close()


def test_c2_full_sets_alg1_alg2_equal_oracle(big_ctx):
    g = W.c2_graph()
    off, adj, lab = g.arrays()
    dg = big_ctx.upload(g, pcsr=True)
    for name, q in W.c2_queries(g):
        say(f"C2 {name}: k={q.k} edges={q.num_edges}, oracle digest on {THREADS} threads")
        want = O.match_digest(q.labels, q.adjm(), off, adj, lab, threads=THREADS)
        say(f"C2 {name}: oracle {want[0]} embeddings")
        assert want[0] > 10 ** 9
        for alg in (F.ALG_MMATCH, F.ALG_GSI):
            say(f"C2 {name}: alg {alg} count and digest")
            assert F.match(dg, q, alg, F.MODE_COUNT).count == want[0], (name, alg)
            assert F.match_digest(dg, q, alg) == want, (name, alg)
    dg.free()


@pytest.fixture(scope="module")
def c3(big_ctx):
    say("C3 graph")
    g = W.c3_graph()
    dg = big_ctx.upload(g)
    yield g, dg
    dg.free()


def test_c3_triangle_count_equals_oracle(c3):
    g, dg = c3
    off, adj, _ = g.arrays()
    q = dict(W.c3_queries())["triangle"]
    say("C3 triangle: oracle orientation count")
    want = 6 * O.triangles(off, adj, threads=THREADS)
    say(f"C3 triangle: oracle {want}; device count")
    assert F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count == want


def _count_plan(ctx, dg, q, fused):
    ctx.set_tune(fused_tail=0 if fused else 1)
    try:
        return F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count
    finally:
        ctx.set_tune()


def test_c3_clique4_count_equal_across_plans(c3, big_ctx):
    g, dg = c3
    q = dict(W.c3_queries())["clique4"]
    say("C3 4-clique: fused last two levels")
    a = _count_plan(big_ctx, dg, q, True)
    say(f"C3 4-clique: {a}; level by level")
    b = _count_plan(big_ctx, dg, q, False)
    say(f"C3 4-clique: {b}")
    assert a == b and a > 0


def test_c3_scale20_cycle4_count_equal_across_plans(big_ctx):
    say("R-MAT scale 20 graph")
    g = W.c3_graph(20, 35_000_000 // 4)
    dg = big_ctx.upload(g)
    q = dict(W.c3_queries())["cycle4"]
    say("scale 20 4-cycle: fused last two levels")
    a = _count_plan(big_ctx, dg, q, True)
    say(f"scale 20 4-cycle: {a}; level by level")
    b = _count_plan(big_ctx, dg, q, False)
    say(f"scale 20 4-cycle: {b}")
    dg.free()
    assert a == b and a > 0


@pytest.mark.parametrize("name", ["triangle", "cycle4", "clique4", "house"])
def test_c3_first_1e8_embeddings_valid_and_distinct(c3, name):
    g, dg = c3
    off, adj, _ = g.arrays()
    lab = np.zeros(g.n, np.int32)
    q = dict(W.c3_queries())[name]
    cap = 10 ** 8
    say(f"C3 {name}: enumerate the first {cap} embeddings")
    r = F.match(dg, q, F.ALG_MMATCH, F.MODE_ENUMERATE, max_embeddings=cap, emb_cap=cap)
    rows = r.embeddings
    assert r.count == cap and len(rows) == cap
    say(f"C3 {name}: check {cap} rows on the host")
    bad, h = O.check_rows(q.labels, q.adjm(), off, adj, lab, rows, threads=THREADS, want_hashes=True)
    assert bad == 0
    del r
    h.sort()
    dup = np.nonzero(h[1:] == h[:-1])[0]
    if len(dup):  # equal hashes: the rows themselves must differ
        hs = O.row_hashes_numpy(rows)
        for i in dup[:100]:
            same = np.nonzero(hs == h[i])[0]
            assert len({tuple(rows[j]) for j in same}) == len(same), name
    say(f"C3 {name}: ok")


ORACLE_STEPS = 5 * 10 ** 8


def _alone(dg, qs, bound):
    out = []
    for q in qs:
        try:
            out.append(F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=bound).count)
        except F.FsigError as e:
            assert e.code == F.FSIG_ELIMIT
            out.append(None)
    return out


def _oracle_agrees(name, g, qs, got, bound):
    off, adj, lab = g.arrays()
    checked = 0
    for i, (q, c) in enumerate(zip(qs, got)):
        if i % 32 == 0:
            say(f"{name}: oracle on queries {i}..{min(i + 32, len(qs)) - 1}")
        if c is None:
            continue
        want, stopped, over = O.match_budget(q.labels, q.adjm(), off, adj, lab, threads=THREADS, limit=bound,
                                             max_steps=ORACLE_STEPS)
        if over:
            continue
        assert c == (bound if stopped else min(want, bound)), (name, q.k, q.edges.tolist())
        checked += 1
    say(f"{name}: the oracle finished {checked} of {len(qs)} queries within {ORACLE_STEPS} steps; all agree")


def test_c4_bounded_batch_equals_plain_and_oracle(big_ctx):
    say("C4 graph and queries")
    g = W.c4_graph()
    qs = W.c4_queries(g)
    p = W.C4_QUERIES
    dg = big_ctx.upload(g)
    big_ctx.set_work_limit(p["work_limit"])
    try:
        say("C4 plain loop")
        plain = _alone(dg, qs, p["bound"])
        say(f"C4 bounded batch ({sum(c is None for c in plain)} refused by the plain loop)")
        counts, _, st = F.match_batch(dg, qs, bounds=p["bound"], status=True)
    finally:
        big_ctx.set_work_limit(0)
    for q, a, c, s in zip(qs, plain, counts.tolist(), st.tolist()):
        if a is not None:
            assert s == 0 and c == a, (q.k, q.edges.tolist(), a, c, s)
    say("C4 oracle under a step budget")
    _oracle_agrees("C4", g, qs, plain, p["bound"])
    dg.free()


def test_c5_served_equals_alone(big_ctx):
    say("C5 graph and request stream")
    g = W.c5_graph()
    qs = [q for _, q in W.c5_stream(g)]
    p = W.C5_STREAM
    dg = big_ctx.upload(g)
    big_ctx.set_work_limit(p["work_limit"])
    try:
        say("C5 each request alone")
        alone = _alone(dg, qs, p["bound"])
    finally:
        big_ctx.set_work_limit(0)
    say("C5 the stream served (4 workers, batches of up to 64)")
    with F.Server(dg, workers=4, hbm_budget=64 << 30, max_batch=64) as srv:
        srv.set_work_limit(p["work_limit"])
        tickets = [srv.submit(q, p["bound"]) for q in qs]
        served = []
        for t in tickets:
            try:
                served.append(srv.wait(t)[0])
            except F.FsigError as e:
                assert e.code == F.FSIG_ELIMIT
                served.append(None)
    for q, a, b in zip(qs, alone, served):
        if a is not None:
            assert b == a, (q.k, q.edges.tolist(), a, b)
    say(f"C5: {sum(a is not None for a in alone)} of {len(qs)} requests answered alone, served equal")
    dg.free()
