"""The seeded generators produce the same graphs from build to build: the
benchmarks' configurations and their seeded random-walk queries depend on it.
SHA-256 of (offsets, adjacency, labels) for R-MAT below and above the size
where duplicates are removed over key shards on threads (4M edges), and for a
labeled Chung-Lu graph."""
import hashlib

import pytest

import fsig_amd as F


def _digest(g):
    h = hashlib.sha256()
    for a in g.arrays():
        h.update(a.tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("scale,m,seed,want", [
    (12, 20000, 3, "078037c3bd77c0f270b1b0c524ff985448cdde059540cf42ec5db95d6304ec2d"),
    (16, 600000, 5, "4baf48114f284633d2497b359a619f011ff13b3b15f8c246cb2ed19561362f80"),
    (20, 5000000, 7, "b53b7d2d498b1a5c1be24a3535965a036601b755a8f2bcac4eb9c29df67bfece"),
])
def test_rmat_graphs_are_pinned(scale, m, seed, want):
    assert _digest(F.HostGraph.rmat(scale, m, seed=seed)) == want


def test_chung_lu_graph_is_pinned():
    g = F.HostGraph.chung_lu(100000, 500000, 2.3, 9).set_labels("zipf", 16, 1.2, 10)
    assert _digest(g) == "4d6fd7f86d2a5e299478ef0c3d79c8d5516d16a3a7fa6bce08119cba9166399b"
