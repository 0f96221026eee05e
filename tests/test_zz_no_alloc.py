"""fsig_match_noalloc (the entry point that allocates nothing: its scratch is
a static device array) against fsig_match, called through the C ABI: equal
counts and embedding sets on two calls in a row, the second on a smaller
input, for both algorithms and both modes; a graph too large for the
scratch is rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest

import fsig_amd as F

pytestmark = pytest.mark.gpu


def _call(ctx, fn, dg, q, alg, mode, cap):
    off, adj = q.csr()
    sig = q.signature() if alg == F.ALG_GSI else None
    cnt = C.c_uint64(0)
    out = np.zeros((max(cap, 1), q.k), np.uint32)
    rc = fn(ctx._h, dg._h, q.k, F._ptr(q.labels, F._i32p), F._ptr(off), F._ptr(adj), F._ptr(sig), alg, mode, 0,
            C.byref(cnt), F._ptr(out) if cap else F._ptr(None), cap)
    ctx.check(rc)
    rows = out[: min(cnt.value, cap)] if cap else None
    return cnt.value, rows


def _sorted(rows):
    return rows[np.lexsort(rows.T[::-1])] if len(rows) else rows


def test_noalloc_equals_match_two_calls_then_smaller():
    ctx = F.Context(0, 1 << 30)
    try:
        L = ctx._L
        big = F.HostGraph.chung_lu(10000, 50000, 2.5, 1).set_labels("uniform", 8, seed=2)
        small = F.HostGraph.chung_lu(3000, 12000, 2.5, 5).set_labels("uniform", 3, seed=6)
        cases = [(big, big.walk_query(4, 5, 5, seed=3)), (small, small.walk_query(4, 3, 5, seed=7))]
        for g, q in cases:
            dg = ctx.upload(g, pcsr=True)
            for alg in (F.ALG_MMATCH, F.ALG_GSI):
                want, want_rows = _call(ctx, L.fsig_match, dg, q, alg, F.MODE_ENUMERATE, 1 << 20)
                assert 0 < want < 1 << 20
                got, rows = _call(ctx, L.fsig_match_noalloc, dg, q, alg, F.MODE_ENUMERATE, 1 << 20)
                assert got == want
                assert np.array_equal(_sorted(rows), _sorted(want_rows))
                assert _call(ctx, L.fsig_match_noalloc, dg, q, alg, F.MODE_COUNT, 0)[0] == want
            dg.free()
    finally:
        ctx.close()


def test_noalloc_rejects_a_graph_beyond_the_scratch():
    ctx = F.Context(0, 1 << 30)
    try:
        n = 23_000_000  # 12 bytes a vertex exceed the 256 MiB scratch
        g = F.HostGraph.from_edges(n, [(0, 1)], np.zeros(n, np.int32))
        dg = ctx.upload(g)
        q = F.Query(2, [0, 0], [(0, 1)])
        off, adj = q.csr()
        cnt = C.c_uint64(0)
        rc = ctx._L.fsig_match_noalloc(ctx._h, dg._h, 2, F._ptr(q.labels, F._i32p), F._ptr(off), F._ptr(adj),
                                       F._ptr(None), F.ALG_MMATCH, F.MODE_COUNT, 0, C.byref(cnt), F._ptr(None), 0)
        assert rc == F.FSIG_EINVAL
        assert b"scratch" in ctx._L.fsig_last_error(ctx._h)
        assert F.match(dg, q).count == 2  # fsig_match takes it
        dg.free()
    finally:
        ctx.close()
