import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def csr_from_edges(n, edges, labels):
    """Plain numpy CSR (ascending, deduplicated) for fixtures."""
    adj = [set() for _ in range(n)]
    for u, v in edges:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    off = np.zeros(n + 1, np.uint32)
    flat = []
    for v in range(n):
        nb = sorted(adj[v])
        flat.extend(nb)
        off[v + 1] = len(flat)
    return off, np.asarray(flat, dtype=np.uint32), np.asarray(labels, dtype=np.int32)


def adjm_from_edges(k, edges):
    m = np.zeros(k, np.uint32)
    for a, b in edges:
        m[a] |= np.uint32(1 << b)
        m[b] |= np.uint32(1 << a)
    return m


@pytest.fixture(scope="session")
def gpu_ctx():
    import fsig_amd as F
    ctx = F.Context(0, 1 << 32)
    yield ctx
    ctx.close()
