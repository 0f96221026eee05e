"""Seeded synthetic workloads (graph shapes, labels and random-walk queries).

Every input of the benchmarks and size-dependent tests comes from here, so a
seed names a workload exactly:

  C1  Chung-Lu gamma 2.5, 10,000 vertices, 50,000 edges, 8 uniform labels;
      the 4-vertex diamond (4-cycle + chord) from a seeded random walk.
  C2  dblp-shaped Chung-Lu gamma 2.3, 317,080 vertices, 1,049,866 edges,
      16 labels with Zipf(1.2) frequencies; one 8-vertex dense query
      (average degree >= 3, i.e. >= 12 edges) and one 8-vertex sparse query
      (< 12 edges), both induced subgraphs of seeded random walks.
  C3  unlabeled R-MAT (0.57, 0.19, 0.19), 2^22 vertices, 35M edges; triangle,
      4-cycle, 4-clique and house.
  C4  Chung-Lu 1M vertices, 10M edges, 32 Zipf(1.2) labels; the first 256
      random-walk queries with k uniform in 4..16, half dense and half sparse,
      each counted to a bound of 1e8 embeddings.
  C5  R-MAT 2^24 vertices, 200M edges, 64 uniform labels; a Poisson stream of
      random-walk queries with k = round(exp(U(ln 4, ln 32))), each counted
      to a bound of 1e8 embeddings.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from . import HostGraph, Query, FsigError

C2_GRAPH = dict(n=317080, m=1049866, gamma=2.3, labels=16, zipf=1.2, graph_seed=2, label_seed=3)
# Query seeds fixed once: the first seeds (from 100) whose walks give the
# wanted edge counts; counts are in the tens of millions to billions.
C2_QUERIES = {"dense": dict(k=8, min_e=12, max_e=28, seed=100, bias=0),
              "sparse": dict(k=8, min_e=7, max_e=11, seed=100, bias=0)}


def c1() -> Tuple[HostGraph, Query]:
    g = HostGraph.chung_lu(10000, 50000, 2.5, 1).set_labels("uniform", 8, seed=2)
    return g, g.walk_query(4, 5, 5, seed=3)


def c2_graph() -> HostGraph:
    p = C2_GRAPH
    return HostGraph.chung_lu(p["n"], p["m"], p["gamma"], p["graph_seed"]).set_labels(
        "zipf", p["labels"], p["zipf"], p["label_seed"])


def c2_queries(g: HostGraph) -> List[Tuple[str, Query]]:
    out = []
    for name, q in C2_QUERIES.items():
        out.append((name, g.walk_query(q["k"], q["min_e"], q["max_e"], seed=q["seed"], bias=q["bias"],
                                       max_attempts=200000)))
    return out


C3_PATTERNS = {
    "triangle": (3, [(0, 1), (1, 2), (0, 2)]),
    "cycle4": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
    "clique4": (4, [(a, b) for a in range(4) for b in range(a + 1, 4)]),
    "house": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (1, 4)]),
}


def c3_graph(scale: int = 22, m: int = 35_000_000, seed: int = 4) -> HostGraph:
    return HostGraph.rmat(scale, m, 0.57, 0.19, 0.19, seed)


def c3_queries() -> List[Tuple[str, Query]]:
    return [(n, Query(k, np.zeros(k, np.int32), e)) for n, (k, e) in C3_PATTERNS.items()]


def c4_graph() -> HostGraph:
    return HostGraph.chung_lu(1_000_000, 10_000_000, 2.3, 5).set_labels("zipf", 32, 1.2, 6)


# C4 and C5 keep every drawn walk: no query is redrawn for its size.  Most
# walk-induced queries on these graphs have astronomically many embeddings
# (more than 1e9 for about 4 in 5 of the C4 draws), so each query is counted
# with a bound: its result is min(count, bound), exact below the bound
# (fsig_match's max_embeddings).  A query whose search produces more partial
# embeddings than the work limit is refused (FSIG_ELIMIT) and reported.
C4_QUERIES = dict(count=256, kmin=4, kmax=16, seed=4, bound=10 ** 8, work_limit=2 * 10 ** 9)


def c4_queries(g: HostGraph, info: dict = None) -> List[Query]:
    """The first 256 walk queries of the C4 law: k uniform in 4..16, dense and
    sparse alternating."""
    p = C4_QUERIES
    return [q for _, q in walk_queries(g, p["count"], p["kmin"], p["kmax"], p["seed"], info=info)]


def walk_queries(g: HostGraph, count: int, kmin: int, kmax: int, seed: int, log_uniform: bool = False,
                 max_edges: int = 10 ** 6, dg=None, max_count: int = 0,
                 max_rows: int = 0, progress=None, max_attempts: int = 2000,
                 alternate_drawn: bool = False, info: dict = None) -> List[Tuple[str, Query]]:
    """count queries, k uniform (or log-uniform) in [kmin, kmax], alternating
    dense (average degree >= 3) and sparse; a draw whose walks find no query
    of the wanted density within max_attempts is skipped deterministically
    (counted in info["no_walk"]).  With a device graph `dg`, a query is also
    rejected when it has more than max_count embeddings or its plain join
    would build more than max_rows intermediate rows (the engine's bounded
    count and work limit decide, so the check stops early on huge queries);
    the workloads never pass dg, tests use it to keep oracle checks small.
    With alternate_drawn, dense and sparse alternate over the drawn walks
    instead of the accepted queries."""
    rng = np.random.default_rng(seed)
    out = []
    s = seed * 100003
    draws = no_walk = 0
    while len(out) < count:
        draws += 1
        if progress is not None and draws % 25 == 0:
            progress(f"{draws} walks drawn, {len(out)} of {count} queries accepted")
        if log_uniform:
            k = int(round(math.exp(rng.uniform(math.log(kmin), math.log(kmax)))))
        else:
            k = int(rng.integers(kmin, kmax + 1))
        # dense and sparse alternate over accepted queries (an even mix), or
        # over drawn walks
        dense = (draws if alternate_drawn else len(out)) % 2 == 0
        lo, hi = ((3 * k + 1) // 2, max_edges) if dense else (k - 1, (3 * k - 1) // 2)
        s += 1
        try:
            q = g.walk_query(k, lo, hi, seed=s, bias=2 if dense else 0, max_attempts=max_attempts)
        except FsigError:
            no_walk += 1
            continue
        if dg is not None and (max_count or max_rows):
            from . import match, ALG_MMATCH, MODE_COUNT, FSIG_ELIMIT
            dg.ctx.set_work_limit(max_rows)
            try:
                c = match(dg, q, ALG_MMATCH, MODE_COUNT, max_embeddings=max_count + 1 if max_count else 0).count
            except FsigError as e:
                if e.code != FSIG_ELIMIT:
                    raise
                continue
            finally:
                dg.ctx.set_work_limit(0)
            if max_count and c > max_count:
                continue
        out.append(("dense" if dense else "sparse", q))
        if progress is not None and len(out) % 32 == 0:
            progress(f"{len(out)} of {count} queries drawn")
    if info is not None:
        info.update(draws=draws, no_walk=no_walk)
    return out


def c5_graph(scale: int = 24, m: int = 200_000_000, seed: int = 8) -> HostGraph:
    return HostGraph.rmat(scale, m, 0.57, 0.19, 0.19, seed).set_labels("uniform", 64, seed=9)


# C5 request stream: k = round(exp(U(ln 4, ln 32))), dense and sparse walks
# alternating per draw (200 attempts per walk), every drawn query kept,
# counted with a bound and refused over the work limit as C4's; Poisson
# arrivals at `qps` (exponential gaps from their own seed).
C5_STREAM = dict(count=128, kmin=4, kmax=32, seed=11, arrival_seed=12, qps=1000.0, bound=10 ** 8,
                 work_limit=2 * 10 ** 9, max_attempts=200)


def c5_stream(g: HostGraph, count: int = 0, qps: float = 0.0, info: dict = None) -> List[Tuple[float, Query]]:
    """(arrival time in s from the stream start, query) pairs."""
    p = dict(C5_STREAM)
    if count:
        p["count"] = count
    qps = qps or p["qps"]
    qs = [q for _, q in walk_queries(g, p["count"], p["kmin"], p["kmax"], p["seed"], log_uniform=True,
                                     max_attempts=p["max_attempts"], alternate_drawn=True, info=info)]
    gaps = np.random.default_rng(p["arrival_seed"]).exponential(1.0 / qps, size=len(qs))
    t = np.cumsum(gaps) - gaps[0]
    return list(zip(t.tolist(), qs))
