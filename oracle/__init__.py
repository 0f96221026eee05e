"""Oracle: CPU restatement of the reference's matching semantics.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the thing
measured or shipped.  The product (fsig_amd) never imports it.

``liboracle.so`` is oracle.c (see its header for the reference lines each
function restates and how each is pinned).  ``_ref/ref_dump`` is the
reference's own host code compiled from its sources (oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional, Tuple

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")
REF_DUMP = os.path.join(ORACLE_DIR, "_ref", "ref_dump")

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.oracle_murmur2.restype = C.c_uint32
        L.oracle_murmur2.argtypes = [C.c_void_p, C.c_int, C.c_uint32]
        L.oracle_signature.argtypes = [C.c_uint32, _u32p, _u32p, _i32p, C.c_int, _u32p]
        L.oracle_ldf.restype = C.c_uint32
        L.oracle_ldf.argtypes = [C.c_uint32, _u32p, _i32p, C.c_int32, C.c_uint32, _u32p]
        L.oracle_sig_filter.restype = C.c_uint32
        L.oracle_sig_filter.argtypes = [C.c_uint32, _u32p, _u32p, _u32p]
        L.oracle_classes.restype = C.c_uint32
        L.oracle_classes.argtypes = [C.c_uint32, _i32p, _u32p, _u32p]
        L.oracle_refine.argtypes = [C.c_uint32, _i32p, _u32p, C.c_uint32, _u32p, _u32p, _i32p, C.c_uint32,
                                    _u32p, _u8p]
        L.oracle_match.restype = C.c_uint64
        L.oracle_match.argtypes = [C.c_uint32, _i32p, _u32p, C.c_uint32, _u32p, _u32p, _i32p, C.c_uint32,
                                   _u32p, C.c_uint64, _u64p]
        L.oracle_match_limit.restype = C.c_uint64
        L.oracle_match_limit.argtypes = L.oracle_match.argtypes + [C.c_uint64, C.POINTER(C.c_int)]
        L.oracle_match_budget.restype = C.c_uint64
        L.oracle_match_budget.argtypes = [C.c_uint32, _i32p, _u32p, C.c_uint32, _u32p, _u32p, _i32p, C.c_uint32,
                                          C.c_uint64, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.oracle_match_digest.restype = C.c_uint64
        L.oracle_match_digest.argtypes = [C.c_uint32, _i32p, _u32p, C.c_uint32, _u32p, _u32p, _i32p, C.c_uint32,
                                          _u64p]
        L.oracle_row_hash.restype = C.c_uint64
        L.oracle_row_hash.argtypes = [C.c_uint32, _u32p]
        L.oracle_triangles.restype = C.c_uint64
        L.oracle_triangles.argtypes = [C.c_uint32, _u32p, _u32p, C.c_uint32]
        L.oracle_check_rows.restype = C.c_uint64
        L.oracle_check_rows.argtypes = [C.c_uint32, _i32p, _u32p, C.c_uint32, _u32p, _u32p, _i32p, _u32p,
                                        C.c_uint64, C.c_uint32, _u64p]
        _lib = L
    return _lib


def _p(a, t=_u32p):
    return C.cast(None, t) if a is None else a.ctypes.data_as(t)


def _csr(off, adj, labels):
    return (np.ascontiguousarray(off, dtype=np.uint32), np.ascontiguousarray(adj, dtype=np.uint32),
            np.ascontiguousarray(labels, dtype=np.int32))


def murmur2_u32(key: int, seed: int = 17) -> int:
    k = np.array([key], dtype=np.uint32)
    return int(lib().oracle_murmur2(k.ctypes.data, 4, seed))


def signature(off, adj, labels, column_oriented: bool) -> np.ndarray:
    off, adj, labels = _csr(off, adj, labels)
    n = len(labels)
    out = np.empty(16 * n, np.uint32)
    lib().oracle_signature(n, _p(off), _p(adj), _p(labels, _i32p), int(column_oriented), _p(out))
    return out


def ldf(off, labels, label: int, mindeg: int) -> np.ndarray:
    off = np.ascontiguousarray(off, dtype=np.uint32)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    out = np.empty(len(labels), np.uint32)
    c = lib().oracle_ldf(len(labels), _p(off), _p(labels, _i32p), label, mindeg, _p(out))
    return out[:c].copy()


def sig_filter(dsig_col, qsig_row16) -> np.ndarray:
    dsig_col = np.ascontiguousarray(dsig_col, dtype=np.uint32)
    q = np.ascontiguousarray(qsig_row16, dtype=np.uint32)
    n = len(dsig_col) // 16
    out = np.empty(n, np.uint32)
    c = lib().oracle_sig_filter(n, _p(dsig_col), _p(q), _p(out))
    return out[:c].copy()


def classes(qlabels, qadjm) -> np.ndarray:
    ql = np.ascontiguousarray(qlabels, dtype=np.int32)
    qa = np.ascontiguousarray(qadjm, dtype=np.uint32)
    out = np.empty(len(ql), np.uint32)
    lib().oracle_classes(len(ql), _p(ql, _i32p), _p(qa), _p(out))
    return out


def refine(qlabels, qadjm, off, adj, labels, passes: int, want_sets: bool = False):
    """Candidate-set sizes per query vertex (and membership flags [k][n])."""
    off, adj, labels = _csr(off, adj, labels)
    ql = np.ascontiguousarray(qlabels, dtype=np.int32)
    qa = np.ascontiguousarray(qadjm, dtype=np.uint32)
    k, n = len(ql), len(labels)
    sizes = np.empty(k, np.uint32)
    sets = np.empty(k * n, np.uint8) if want_sets else None
    lib().oracle_refine(k, _p(ql, _i32p), _p(qa), n, _p(off), _p(adj), _p(labels, _i32p), passes, _p(sizes),
                        _p(sets, _u8p))
    return (sizes, sets.reshape(k, n)) if want_sets else sizes


TOO_MANY = (1 << 64) - 1


def match(qlabels, qadjm, off, adj, labels, threads: int = 1, cap: int = 0,
          limit: int = 0) -> Tuple[int, Optional[np.ndarray]]:
    """(count, rows) of the injective label- and edge-preserving maps; rows
    (query-vertex order) only when cap > 0, at most cap of them.  With
    limit > 0 the search stops after about `limit` embeddings and the count
    is TOO_MANY (match_partial returns the partial count)."""
    cnt, rows, stopped = match_partial(qlabels, qadjm, off, adj, labels, threads, cap, limit)
    return (TOO_MANY if stopped else cnt), rows


def match_partial(qlabels, qadjm, off, adj, labels, threads: int = 1, cap: int = 0, limit: int = 0):
    """(count, rows, stopped): like match(), but a stopped search returns the
    embeddings it had counted when it stopped."""
    off, adj, labels = _csr(off, adj, labels)
    ql = np.ascontiguousarray(qlabels, dtype=np.int32)
    qa = np.ascontiguousarray(qadjm, dtype=np.uint32)
    k = len(ql)
    rows = np.empty((cap, k), np.uint32) if cap else None
    written = C.c_uint64(0)
    stopped = C.c_int(0)
    cnt = lib().oracle_match_limit(k, _p(ql, _i32p), _p(qa), len(labels), _p(off), _p(adj), _p(labels, _i32p),
                                   threads, _p(rows), cap, C.byref(written), limit, C.byref(stopped))
    return int(cnt), (rows[: written.value] if cap else None), bool(stopped.value)


def match_budget(qlabels, qadjm, off, adj, labels, threads: int = 1, limit: int = 0, max_steps: int = 0):
    """(count, stopped, over_budget): the count of match(), or None when
    the search tried more than about max_steps candidates (over_budget).
    stopped: more than `limit` embeddings were seen (count partial)."""
    off, adj, labels = _csr(off, adj, labels)
    ql = np.ascontiguousarray(qlabels, dtype=np.int32)
    qa = np.ascontiguousarray(qadjm, dtype=np.uint32)
    stopped, over = C.c_int(0), C.c_int(0)
    cnt = lib().oracle_match_budget(len(ql), _p(ql, _i32p), _p(qa), len(labels), _p(off), _p(adj), _p(labels, _i32p),
                                    threads, limit, max_steps, C.byref(stopped), C.byref(over))
    return (None if over.value else int(cnt)), bool(stopped.value), bool(over.value)


def match_digest(qlabels, qadjm, off, adj, labels, threads: int = 1) -> Tuple[int, Tuple[int, int]]:
    """(count, digest) of the full embedding set: the digest of
    fsig_match_digest, computed while enumerating on the host."""
    off, adj, labels = _csr(off, adj, labels)
    ql = np.ascontiguousarray(qlabels, dtype=np.int32)
    qa = np.ascontiguousarray(qadjm, dtype=np.uint32)
    dig = np.zeros(2, np.uint64)
    cnt = lib().oracle_match_digest(len(ql), _p(ql, _i32p), _p(qa), len(labels), _p(off), _p(adj),
                                    _p(labels, _i32p), threads, _p(dig, _u64p))
    return int(cnt), (int(dig[0]), int(dig[1]))


_M64 = (1 << 64) - 1


def _mix64_np(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    return z ^ (z >> np.uint64(31))


def row_hashes_numpy(rows) -> np.ndarray:
    """The digest's row hash of each row (query-vertex order), in numpy: an
    independent restatement of oracle_row_hash."""
    rows = np.asarray(rows, dtype=np.uint64)
    n, k = rows.shape
    with np.errstate(over="ignore"):
        x = np.full(n, (0x9e3779b97f4a7c15 * (k + 1)) & _M64, dtype=np.uint64)
        for i in range(k):
            x = _mix64_np(x + rows[:, i] + np.uint64(i << 32))
    return x


def digest_numpy(rows) -> Tuple[int, int]:
    """Set digest of explicit rows (numpy restatement)."""
    x = row_hashes_numpy(rows)
    with np.errstate(over="ignore"):
        d1 = _mix64_np(x ^ np.uint64(0xd6e8feb86659fd93))
    return int(x.sum(dtype=np.uint64)), int(d1.sum(dtype=np.uint64))


def triangles(off, adj, threads: int = 1) -> int:
    """Number of triangles of an undirected CSR graph (each counted once);
    the unlabeled triangle query has 6 embeddings per triangle."""
    off = np.ascontiguousarray(off, dtype=np.uint32)
    adj = np.ascontiguousarray(adj, dtype=np.uint32)
    return int(lib().oracle_triangles(len(off) - 1, _p(off), _p(adj), threads))


def check_rows(qlabels, qadjm, off, adj, labels, rows, threads: int = 1, want_hashes: bool = False):
    """(number of rows that are not embeddings, row hashes or None)."""
    off, adj, labels = _csr(off, adj, labels)
    ql = np.ascontiguousarray(qlabels, dtype=np.int32)
    qa = np.ascontiguousarray(qadjm, dtype=np.uint32)
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    nrows = rows.shape[0]
    hashes = np.empty(nrows, np.uint64) if want_hashes else None
    bad = lib().oracle_check_rows(len(ql), _p(ql, _i32p), _p(qa), len(labels), _p(off), _p(adj), _p(labels, _i32p),
                                  _p(rows), nrows, threads, _p(hashes, _u64p))
    return int(bad), hashes


# ---------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_DUMP)


def ref_ingest_seconds(path: str, labeled: bool = True, timeout: float = 600.0) -> Dict[str, float]:
    """Wall seconds of the reference's own ingest stages on a file (its
    reader, PCSR transform, CSR transform and data signature), as timed by
    ref_dump itself: {"read": s, "pcsr": s, "csr": s, "signature": s}."""
    out = subprocess.run([REF_DUMP, path, "1" if labeled else "0", "time"], check=True, capture_output=True,
                         text=True, timeout=timeout).stdout
    return {f[1]: float(f[2]) for f in (line.split() for line in out.splitlines()) if len(f) == 3 and f[0] == "TIME"}


def ref_dump(path: str, labeled: bool = True, murmur_count: int = 64) -> Dict[str, object]:
    """Runs the reference's own reader/CSR/PCSR/signature code on a file."""
    out = subprocess.run([REF_DUMP, path, "1" if labeled else "0", str(murmur_count)], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    res: Dict[str, object] = {}
    i = 0

    def nums(line):
        return np.array(line.split(), dtype=np.uint64).astype(np.uint32) if line.strip() else np.empty(0, np.uint32)

    units = []
    while i < len(out):
        line = out[i]
        if not line:
            i += 1
            continue
        tag, *rest = line.split()
        if tag == "CSR":
            res["n"], res["m"], res["L"] = int(rest[0]), int(rest[1]), int(rest[2])
            i += 1
        elif tag in ("OFF", "ADJ", "LAB", "INVIDX", "INVLIST", "SIGQ", "SIGD", "UOFF", "UEDGE"):
            arr = nums(out[i + 1])
            if tag == "UOFF":
                units[-1]["offset"] = arr
            elif tag == "UEDGE":
                units[-1]["edges"] = arr
            else:
                res[tag] = arr
            i += 2
        elif tag == "PCSR":
            i += 1
        elif tag == "UNIT":
            units.append({"keys": int(rest[0]), "nedges": int(rest[1])})
            i += 1
        elif tag == "MURMUR":
            res["MURMUR"] = nums(out[i + 1])
            i += 2
        else:
            i += 1
    res["units"] = units
    return res
