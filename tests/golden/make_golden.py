"""Writes the golden fixtures under tests/golden/ (run in the build container).

* vf2_cases.json: small labeled / unlabeled data graphs and queries with the
  embedding sets networkx 3.4 computes with VF2 subgraph monomorphisms
  (``GraphMatcher(G, Q, node_match=label).subgraph_monomorphisms_iter()``),
  i.e. every injective, label- and edge-preserving map.  Independent of the
  oracle and of the engine.
* ref_tiny.json: what the reference's own host code (oracle/_ref/ref_dump,
  built from the reference sources) produces for a few tiny graph files:
  CSR, inverted index, PCSR units (reference unit numbering), signatures and
  MurmurHash2 known answers.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import random
import sys
import tempfile

import networkx as nx
from networkx.algorithms import isomorphism as iso

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def rand_graph(rng, n, p, nlabels):
    edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    labels = [rng.randrange(nlabels) for _ in range(n)]
    return edges, labels


def vf2_embeddings(n, edges, labels, k, qedges, qlabels):
    G = nx.Graph()
    G.add_nodes_from((v, {"l": labels[v]}) for v in range(n))
    G.add_edges_from(edges)
    Q = nx.Graph()
    Q.add_nodes_from((u, {"l": qlabels[u]}) for u in range(k))
    Q.add_edges_from(qedges)
    gm = iso.GraphMatcher(G, Q, node_match=lambda a, b: a["l"] == b["l"])
    rows = []
    for m in gm.subgraph_monomorphisms_iter():  # data vertex -> query vertex
        row = [0] * k
        for dv, qv in m.items():
            row[qv] = dv
        rows.append(row)
    rows.sort()
    return rows


QUERIES = {
    "edge": (2, [(0, 1)]),
    "path3": (3, [(0, 1), (1, 2)]),
    "triangle": (3, [(0, 1), (1, 2), (0, 2)]),
    "star3": (4, [(0, 1), (0, 2), (0, 3)]),
    "cycle4": (4, [(0, 1), (1, 2), (2, 3), (3, 0)]),
    "diamond": (4, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]),
    "clique4": (4, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]),
    "house": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (1, 4)]),
    "twoedges": (4, [(0, 1), (2, 3)]),  # disconnected
}


def make_vf2():
    rng = random.Random(20261202)
    cases = []
    for gi in range(10):
        n = rng.randrange(8, 16)
        p = rng.choice([0.25, 0.35, 0.5])
        nl = rng.choice([1, 2, 3])
        edges, labels = rand_graph(rng, n, p, nl)
        for qname, (k, qe) in QUERIES.items():
            qlabels = [rng.randrange(nl) for _ in range(k)]
            rows = vf2_embeddings(n, edges, labels, k, qe, qlabels)
            if len(rows) > 5000:
                continue
            cases.append({"graph": gi, "n": n, "edges": edges, "labels": labels, "query": qname, "k": k,
                          "qedges": qe, "qlabels": qlabels, "count": len(rows), "rows": rows})
    with open(os.path.join(HERE, "vf2_cases.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))
    print("vf2_cases.json:", len(cases), "cases")


def write_graph_file(path, n, edges, labels, trailing_newline=False):
    lines = [f"t {n} {len(edges)}"]
    deg = [0] * n
    for u, v in edges:
        deg[u] += 1
        deg[v] += 1
    lines += [f"v {v} {labels[v]} {deg[v]}" for v in range(n)]
    lines += [f"e {u} {v}" for u, v in edges]
    with open(path, "w") as f:
        f.write("\n".join(lines) + ("\n" if trailing_newline else ""))


def make_ref():
    import oracle
    if not oracle.ref_available():
        print("oracle/_ref/ref_dump not built; skipping ref_tiny.json")
        return
    rng = random.Random(7)
    out = []
    with tempfile.TemporaryDirectory() as td:
        for gi in range(4):
            n = rng.randrange(20, 60)
            nl = [1, 3, 5, 8][gi]
            edges, labels = rand_graph(rng, n, 0.15, nl)
            # keep the max id non-isolated, as the reference reader needs
            if not any(n - 1 in e for e in edges):
                edges.append((0, n - 1))
            path = os.path.join(td, f"g{gi}.graph")
            write_graph_file(path, n, edges, labels)
            d = oracle.ref_dump(path, True, 256)
            out.append({
                "n": n, "edges": edges, "labels": labels,
                "OFF": d["OFF"].tolist(), "ADJ": d["ADJ"].tolist(), "INVIDX": d["INVIDX"].tolist(),
                "INVLIST": d["INVLIST"].tolist(), "SIGQ": d["SIGQ"].tolist(), "SIGD": d["SIGD"].tolist(),
                "units": [{"offset": u["offset"].tolist(), "edges": u["edges"].tolist()} for u in d["units"]],
                "MURMUR": d["MURMUR"].tolist(),
            })
    with open(os.path.join(HERE, "ref_tiny.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("ref_tiny.json:", len(out), "graphs")


if __name__ == "__main__":
    make_vf2()
    make_ref()
