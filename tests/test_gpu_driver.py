"""The FSIG driver (`FSIG <data> <query> <alg> <labeled> [out]`, the
reference's command line) through the engine: the printed count and the
canonical embedding file equal the oracle, for alg 1 and alg 2, and for a
data graph read from its binary CSR cache."""
import os
import re
import subprocess

import numpy as np
import pytest

import fsig_amd as F
import oracle as O

pytestmark = pytest.mark.gpu

FSIG = os.path.join(os.path.dirname(F.HIP_LIB_PATH), "bin", "FSIG")


def _run(args):
    p = subprocess.run([FSIG] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    m = re.search(r"#embeddings?: (\d+)", p.stdout)
    assert m, p.stdout
    return int(m.group(1))


def _rows(path):
    with open(path) as f:
        head = f.readline()
        rows = [tuple(int(x) for x in line.split()) for line in f if line.strip()]
    return int(head.split(":")[1]), rows


def test_driver_counts_and_embedding_file(tmp_path):
    g = F.HostGraph.chung_lu(3000, 12000, 2.5, 71).set_labels("uniform", 3, seed=72)
    off, adj, lab = g.arrays()
    data = str(tmp_path / "data.graph")
    g.write(data)
    q = g.walk_query(4, 4, 6, seed=73)
    query = str(tmp_path / "query.graph")
    q.host_graph().write(query)
    want, want_rows = O.match(q.labels, q.adjm(), off, adj, lab, threads=8, cap=10 ** 6)
    want_rows = sorted(map(tuple, np.asarray(want_rows).tolist()))
    cache = str(tmp_path / "data.csr.bin")
    g.save_bin(cache)
    for alg, src in ((1, data), (2, data), (1, cache)):
        out = str(tmp_path / f"emb_{alg}_{os.path.basename(src)}.txt")
        assert _run([src, query, alg, 1, out]) == want, (alg, src)
        cnt, rows = _rows(out)
        assert cnt == want and rows == want_rows, (alg, src)
