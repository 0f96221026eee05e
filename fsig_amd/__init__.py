"""fsig_amd: MI355X (gfx950) subgraph-matching engine behind FSIG's matcher
interface.

The engine is native: ``libfsig_hip.so`` (HIP kernels + the C ABI of
``include/fsig.h``) and ``libfsig_host.so`` (graph IO, CSR/PCSR/signature
builders, seeded generators; ``include/fsig_host.h``).  This module binds both
with ctypes for tests and benchmarks, and mirrors the reference's matcher
classes (``GPUSubgraphMatcher``, ``MMatch``, ``GSI``; match/subgraph_match.h:
39-81, match/MMatch.h:155-176, match/GSI.h:67-91) so that a test reads like a
call of the reference's ``match()``.

Nothing here falls back to a CPU path: if ``libfsig_hip.so`` is missing or no
GPU is present, the GPU entry points raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(PKG_DIR, "libfsig_hip.so")
HOST_LIB_PATH = os.path.join(PKG_DIR, "libfsig_host.so")
FSIG_BIN = os.path.join(PKG_DIR, "bin", "FSIG")

FSIG_OK, FSIG_EINVAL, FSIG_ENOMEM, FSIG_EHIP, FSIG_EUNSUPPORTED, FSIG_ENOTFOUND, FSIG_EAGAIN = 0, -1, -2, -3, -4, -5, -6
FSIG_ELIMIT = -7
ALG_MMATCH, ALG_GSI = 1, 2
MODE_COUNT, MODE_ENUMERATE = 0, 1
K_FILTER, K_REFINE, K_CEDGE, K_JOIN_COUNT, K_JOIN_WRITE, K_SCAN, K_BATCH, K_JOIN_TAIL = range(8)
KERNEL_FAMILIES = ["filter", "refine", "cedge", "join_count", "join_write", "scan", "batch", "join_tail"]

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p


class FsigError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fsig error {code}: {msg}")
        self.code = code


def _ptr(a: Optional[np.ndarray], t=_u32p):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------
# library loading
_host = None
_hip = None


def host_lib():
    """libfsig_host.so (CPU only)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise FileNotFoundError(f"{HOST_LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(HOST_LIB_PATH)
        L.fsig_hg_read.restype = _vp
        L.fsig_hg_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
        L.fsig_hg_from_csr.restype = _vp
        L.fsig_hg_from_csr.argtypes = [C.c_uint32, _u32p, _u32p, _i32p]
        L.fsig_hg_from_edges.restype = _vp
        L.fsig_hg_from_edges.argtypes = [C.c_uint32, C.c_uint64, _u32p, _i32p]
        L.fsig_hg_chung_lu.restype = _vp
        L.fsig_hg_chung_lu.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_uint64]
        L.fsig_hg_rmat.restype = _vp
        L.fsig_hg_rmat.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64]
        L.fsig_hg_set_labels.argtypes = [_vp, C.c_int, C.c_int32, C.c_double, C.c_uint64]
        L.fsig_hg_write.argtypes = [_vp, C.c_char_p]
        L.fsig_hg_info.argtypes = [_vp, _u32p, _u64p, _i32p, _u32p]
        L.fsig_hg_export.argtypes = [_vp, _u32p, _u32p, _i32p]
        L.fsig_hg_inverted_index.argtypes = [_vp, _u32p, _u32p]
        L.fsig_hg_signature.argtypes = [_vp, C.c_int, C.c_int, _u32p]
        L.fsig_hg_free.argtypes = [_vp]
        L.fsig_hg_walk_query.argtypes = [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int,
                                         C.c_uint32, _i32p, _u32p, _u32p]
        L.fsig_pcsr_build.restype = _vp
        L.fsig_pcsr_build.argtypes = [_vp, C.c_int]
        L.fsig_pcsr_num_units.restype = C.c_uint32
        L.fsig_pcsr_num_units.argtypes = [_vp]
        L.fsig_pcsr_unit_info.argtypes = [_vp, C.c_uint32, _u32p, _u32p]
        L.fsig_pcsr_unit_offset.restype = _u32p
        L.fsig_pcsr_unit_offset.argtypes = [_vp, C.c_uint32]
        L.fsig_pcsr_unit_edges.restype = _u32p
        L.fsig_pcsr_unit_edges.argtypes = [_vp, C.c_uint32]
        L.fsig_pcsr_free.argtypes = [_vp]
        L.fsig_murmur2.restype = C.c_uint32
        L.fsig_murmur2.argtypes = [C.c_void_p, C.c_int, C.c_uint32]
        L.fsig_write_embeddings.argtypes = [C.c_char_p, C.c_uint32, C.c_uint64, _u32p, C.c_uint64]
        L.fsig_hg_save_bin.argtypes = [_vp, C.c_char_p]
        L.fsig_hg_load_bin.restype = _vp
        L.fsig_hg_load_bin.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.fsig_pcsr_save_bin.argtypes = [_vp, _vp, C.c_char_p]
        L.fsig_pcsr_load_bin.restype = _vp
        L.fsig_pcsr_load_bin.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.fsig_pcsr_graph_info.argtypes = [_vp, _u32p, _u32p, _i32p]
        L.fsig_pcsr_labels.restype = _i32p
        L.fsig_pcsr_labels.argtypes = [_vp]
        _host = L
    return _host


HIP_SYMBOLS = [
    "fsig_ctx_create", "fsig_ctx_destroy", "fsig_last_error", "fsig_ctx_synchronize", "fsig_graph_upload_csr",
    "fsig_graph_upload_pcsr", "fsig_graph_free", "fsig_match", "fsig_set_profiling",
    "fsig_stats_reset", "fsig_kernel_stats", "fsig_last_candidates", "fsig_last_peak_bytes",
    "fsig_set_tune", "fsig_get_tune", "fsig_stream_copy", "fsig_match_batch", "fsig_last_batch_tune",
    "fsig_set_work_limit", "fsig_server_create", "fsig_server_submit", "fsig_server_wait",
    "fsig_server_get_stats", "fsig_server_projected", "fsig_server_last_error", "fsig_server_destroy",
    "fsig_project_bytes", "fsig_graph_label_pairs", "fsig_match_digest", "fsig_set_keep_candidates",
    "fsig_last_candidate_set", "fsig_match_batch_bounded", "fsig_server_submit_bounded",
    "fsig_server_set_work_limit", "fsig_match_noalloc",
]


class Tune(C.Structure):
    _fields_ = [("lanes_per_row", C.c_uint32), ("strategy", C.c_uint32), ("lds_tile", C.c_uint32),
                ("refine_passes", C.c_uint32), ("fused_tail", C.c_uint32), ("nlf_filter", C.c_uint32),
                ("beam_log2", C.c_uint32), ("symmetry", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("fallbacks", C.c_uint64), ("max_batch_seen", C.c_uint64), ("refused", C.c_uint64),
                ("max_concurrent", C.c_uint32),
                ("pad", C.c_uint32), ("busy_ms", C.c_double), ("wall_ms", C.c_double)]


def hip_lib():
    """libfsig_hip.so.  Loading does not touch the GPU."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise FileNotFoundError(f"{HIP_LIB_PATH} is not built; the HIP engine is required (no CPU fallback)")
        L = C.CDLL(HIP_LIB_PATH)
        L.fsig_ctx_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(_vp)]
        L.fsig_ctx_destroy.argtypes = [_vp]
        L.fsig_ctx_synchronize.argtypes = [_vp]
        L.fsig_last_error.restype = C.c_char_p
        L.fsig_last_error.argtypes = [_vp]
        L.fsig_graph_upload_csr.argtypes = [_vp, C.c_uint32, _u32p, _u32p, _i32p, C.c_uint32, _u32p, _u32p,
                                            C.POINTER(_vp)]
        L.fsig_graph_upload_pcsr.argtypes = [_vp, _vp, C.c_uint32, _u32p, C.POINTER(_u32p), _u32p,
                                             C.POINTER(_u32p), _u32p, C.c_int]
        L.fsig_graph_free.argtypes = [_vp, _vp]
        L.fsig_match.argtypes = [_vp, _vp, C.c_uint32, _i32p, _u32p, _u32p, _u32p, C.c_int, C.c_int,
                                 C.c_uint64, _u64p, _u32p, C.c_uint64]
        L.fsig_set_profiling.argtypes = [_vp, C.c_int]
        L.fsig_stats_reset.argtypes = [_vp]
        L.fsig_kernel_stats.argtypes = [_vp, C.c_int, _u64p, C.POINTER(C.c_double), _u64p, _u64p]
        L.fsig_last_candidates.argtypes = [_vp, C.c_uint32, _u32p]
        L.fsig_last_peak_bytes.restype = C.c_uint64
        L.fsig_last_peak_bytes.argtypes = [_vp]
        L.fsig_set_tune.argtypes = [_vp, C.POINTER(Tune)]
        L.fsig_get_tune.argtypes = [_vp, C.POINTER(Tune)]
        L.fsig_stream_copy.argtypes = [_vp, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
        L.fsig_set_work_limit.argtypes = [_vp, C.c_uint64]
        L.fsig_match_digest.argtypes = [_vp, _vp, C.c_uint32, _i32p, _u32p, _u32p, _u32p, C.c_int, _u64p, _u64p]
        L.fsig_set_keep_candidates.argtypes = [_vp, C.c_int]
        L.fsig_match_noalloc.argtypes = L.fsig_match.argtypes
        L.fsig_last_candidate_set.argtypes = [_vp, C.c_uint32, C.c_int, _u32p, C.c_uint32, _u32p]
        _optional_bindings(L)
        _hip = L
    return _hip


def _optional_bindings(L):
    """Batched and served entry points (see include/fsig.h)."""
    L.fsig_match_batch.argtypes = [_vp, _vp, C.c_uint32, _u32p, C.POINTER(_i32p), C.POINTER(_u32p),
                                   C.POINTER(_u32p), _u64p, C.POINTER(C.c_double)]
    L.fsig_match_batch_bounded.argtypes = [_vp, _vp, C.c_uint32, _u32p, C.POINTER(_i32p), C.POINTER(_u32p),
                                           C.POINTER(_u32p), _u64p, _u64p, _i32p, C.POINTER(C.c_double)]
    L.fsig_last_batch_tune.argtypes = [_vp, _u32p, _u32p, _u32p]
    L.fsig_server_create.argtypes = [C.c_int, _vp, C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(Tune),
                                     C.POINTER(_vp)]
    L.fsig_server_submit.argtypes = [_vp, C.c_uint32, _i32p, _u32p, _u32p, _u64p]
    L.fsig_server_submit_bounded.argtypes = [_vp, C.c_uint32, _i32p, _u32p, _u32p, C.c_uint64, _u64p]
    L.fsig_server_set_work_limit.argtypes = [_vp, C.c_uint64]
    L.fsig_server_wait.argtypes = [_vp, C.c_uint64, C.c_int64, _u64p, C.POINTER(C.c_double)]
    L.fsig_server_get_stats.argtypes = [_vp, C.POINTER(ServerStats)]
    L.fsig_server_projected.argtypes = [_vp, C.c_uint32, _i32p, _u32p, _u32p, _u64p]
    L.fsig_server_last_error.restype = C.c_char_p
    L.fsig_server_last_error.argtypes = [_vp]
    L.fsig_server_destroy.argtypes = [_vp]
    L.fsig_server_destroy.restype = None
    L.fsig_project_bytes.argtypes = [C.c_uint32, C.c_uint32, _u64p, _u64p, C.c_uint32, _i32p, _u32p, _u32p,
                                     _u64p]
    L.fsig_graph_label_pairs.argtypes = [_vp, _vp, _u64p]


# ---------------------------------------------------------------------------
# host graphs and queries
class HostGraph:
    """A CSR data graph held by libfsig_host.so."""

    def __init__(self, handle):
        if not handle:
            raise FsigError(FSIG_EINVAL, "host graph construction failed")
        self._h = handle
        self._L = host_lib()
        n, nnz, nl, md = C.c_uint32(), C.c_uint64(), C.c_int32(), C.c_uint32()
        self._L.fsig_hg_info(self._h, C.byref(n), C.byref(nnz), C.byref(nl), C.byref(md))
        self.n, self.nnz, self.num_labels, self.max_degree = n.value, nnz.value, nl.value, md.value
        self._arrays = None

    def __del__(self):
        try:
            if self._h:
                self._L.fsig_hg_free(self._h)
        except Exception:
            pass

    # constructors
    @classmethod
    def read(cls, path: str, labeled: bool = True) -> "HostGraph":
        st = C.c_int(0)
        h = host_lib().fsig_hg_read(path.encode(), int(labeled), C.byref(st))
        if not h:
            raise FsigError(st.value, f"cannot read {path}")
        return cls(h)

    @classmethod
    def load_bin(cls, path: str) -> "HostGraph":
        """A graph from its binary CSR cache (fsig_hg_load_bin)."""
        st = C.c_int(0)
        h = host_lib().fsig_hg_load_bin(path.encode(), C.byref(st))
        if not h:
            raise FsigError(st.value, f"cannot load the CSR cache {path}")
        return cls(h)

    def save_bin(self, path: str) -> None:
        if self._L.fsig_hg_save_bin(self._h, path.encode()):
            raise FsigError(FSIG_EINVAL, f"cannot write {path}")

    @classmethod
    def from_edges(cls, n: int, edges, labels=None) -> "HostGraph":
        e = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
        lab = None if labels is None else np.ascontiguousarray(np.asarray(labels, dtype=np.int32))
        return cls(host_lib().fsig_hg_from_edges(n, len(e), _ptr(e), _ptr(lab, _i32p)))

    @classmethod
    def chung_lu(cls, n: int, m: int, gamma: float, seed: int) -> "HostGraph":
        return cls(host_lib().fsig_hg_chung_lu(n, m, gamma, seed))

    @classmethod
    def rmat(cls, scale: int, m: int, a=0.57, b=0.19, c=0.19, seed: int = 1) -> "HostGraph":
        return cls(host_lib().fsig_hg_rmat(scale, m, a, b, c, seed))

    def set_labels(self, kind: str, num_labels: int, s: float = 1.2, seed: int = 1) -> "HostGraph":
        self._L.fsig_hg_set_labels(self._h, 0 if kind == "uniform" else 1, num_labels, s, seed)
        self.num_labels = num_labels
        self._arrays = None
        return self

    def write(self, path: str):
        rc = self._L.fsig_hg_write(self._h, path.encode())
        if rc:
            raise FsigError(rc, f"cannot write {path}")

    # arrays
    def arrays(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        if self._arrays is None:
            off = np.empty(self.n + 1, np.uint32)
            adj = np.empty(self.nnz, np.uint32)
            lab = np.empty(self.n, np.int32)
            self._L.fsig_hg_export(self._h, _ptr(off), _ptr(adj), _ptr(lab, _i32p))
            self._arrays = (off, adj, lab)
        return self._arrays

    def inverted_index(self) -> Tuple[np.ndarray, np.ndarray]:
        idx = np.empty(self.num_labels + 1, np.uint32)
        lst = np.empty(self.n, np.uint32)
        self._L.fsig_hg_inverted_index(self._h, _ptr(idx), _ptr(lst))
        return idx, lst

    def signature(self, column_oriented: bool = True, x86_shift: bool = True) -> np.ndarray:
        out = np.empty(16 * self.n, np.uint32)
        self._L.fsig_hg_signature(self._h, int(column_oriented), int(x86_shift), _ptr(out))
        return out

    def pcsr(self, reference_unit_ids: bool = False) -> "HostPCSR":
        h = self._L.fsig_pcsr_build(self._h, int(reference_unit_ids))
        if not h:
            raise FsigError(-1, f"PCSR is built for at most 4096 labels (graph has {self.num_labels})")
        return HostPCSR(h, reference_unit_ids)

    def walk_query(self, k: int, min_edges: int = 0, max_edges: int = 10 ** 6, seed: int = 1, bias: int = 0,
                   max_attempts: int = 100000) -> "Query":
        max_e = min(max_edges, k * (k - 1) // 2)
        lab = np.empty(k, np.int32)
        ed = np.empty(2 * max(max_e, 1), np.uint32)
        walk = np.empty(k, np.uint32)
        ne = self._L.fsig_hg_walk_query(self._h, k, min_edges, max_e, seed, bias, max_attempts,
                                        _ptr(lab, _i32p), _ptr(ed), _ptr(walk))
        if ne < 0:
            raise FsigError(FSIG_EINVAL, f"no walk query with k={k}, edges in [{min_edges},{max_e}]")
        q = Query(k, lab, ed[: 2 * ne].reshape(-1, 2))
        q.walk = walk
        q.walk_seed = seed
        return q


class HostPCSR:
    def __init__(self, handle, reference_unit_ids: bool):
        self._h = handle
        self._L = host_lib()
        self.reference_unit_ids = reference_unit_ids
        self.num_units = self._L.fsig_pcsr_num_units(handle)

    def __del__(self):
        try:
            self._L.fsig_pcsr_free(self._h)
        except Exception:
            pass

    @classmethod
    def load_bin(cls, path: str, reference_unit_ids: bool = False) -> "HostPCSR":
        """Units (and the graph header and labels) from a PCSR cache in the
        reference's layout; reference_unit_ids states how its units are
        numbered."""
        st = C.c_int(0)
        h = host_lib().fsig_pcsr_load_bin(path.encode(), C.byref(st))
        if not h:
            raise FsigError(st.value, f"cannot load the PCSR cache {path}")
        return cls(h, reference_unit_ids)

    def save_bin(self, path: str, graph: "HostGraph") -> None:
        if self._L.fsig_pcsr_save_bin(self._h, graph._h, path.encode()):
            raise FsigError(FSIG_EINVAL, f"cannot write {path}")

    def graph_info(self) -> Tuple[int, int, int, np.ndarray]:
        """(vertices, undirected edges, labels, label array) of a loaded cache."""
        n, m, nl = C.c_uint32(), C.c_uint32(), C.c_int32()
        self._L.fsig_pcsr_graph_info(self._h, C.byref(n), C.byref(m), C.byref(nl))
        lab = np.ctypeslib.as_array(self._L.fsig_pcsr_labels(self._h), shape=(n.value,)).copy() \
            if n.value else np.empty(0, np.int32)
        return n.value, m.value, nl.value, lab

    def unit(self, u: int) -> Tuple[np.ndarray, np.ndarray]:
        keys, ne = C.c_uint32(), C.c_uint32()
        self._L.fsig_pcsr_unit_info(self._h, u, C.byref(keys), C.byref(ne))
        off = np.ctypeslib.as_array(self._L.fsig_pcsr_unit_offset(self._h, u), shape=(32 * keys.value,)) \
            if keys.value else np.empty(0, np.uint32)
        ed = np.ctypeslib.as_array(self._L.fsig_pcsr_unit_edges(self._h, u), shape=(ne.value,)) \
            if ne.value else np.empty(0, np.uint32)
        return off.copy(), ed.copy()


@dataclass
class Query:
    """A query graph: k vertices, labels, undirected edge list (u, v)."""
    k: int
    labels: np.ndarray
    edges: np.ndarray
    walk: Optional[np.ndarray] = None

    def __post_init__(self):
        self.labels = np.ascontiguousarray(np.asarray(self.labels, dtype=np.int32))
        self.edges = np.asarray(self.edges, dtype=np.uint32).reshape(-1, 2)

    def adjm(self) -> np.ndarray:
        m = np.zeros(self.k, np.uint32)
        for a, b in self.edges:
            if a != b:
                m[a] |= np.uint32(1 << int(b))
                m[b] |= np.uint32(1 << int(a))
        return m

    def csr(self) -> Tuple[np.ndarray, np.ndarray]:
        m = self.adjm()
        off = np.zeros(self.k + 1, np.uint32)
        adj = []
        for u in range(self.k):
            nb = [w for w in range(self.k) if (int(m[u]) >> w) & 1]
            adj.extend(nb)
            off[u + 1] = len(adj)
        return off, np.asarray(adj, dtype=np.uint32)

    def host_graph(self) -> HostGraph:
        return HostGraph.from_edges(self.k, self.edges, self.labels)

    def signature(self) -> np.ndarray:
        """Row-major [k][16] signature (graph_format_transformer.cc:217-281)."""
        return self.host_graph().signature(column_oriented=False)

    @property
    def num_edges(self) -> int:
        return int(len({(min(a, b), max(a, b)) for a, b in self.edges.tolist() if a != b}))


# ---------------------------------------------------------------------------
# device side
class Context:
    """fsig_ctx: device, stream and HBM arena (match/subgraph_match.h:51-80)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._L = hip_lib()
        h = _vp()
        rc = self._L.fsig_ctx_create(device, hbm_budget, C.byref(h))
        if rc:
            raise FsigError(rc, f"cannot create a context on device {device}")
        self._h = h
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._L.fsig_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int):
        if rc:
            raise FsigError(rc, self._L.fsig_last_error(self._h).decode())

    def synchronize(self):
        self.check(self._L.fsig_ctx_synchronize(self._h))

    def upload(self, g: HostGraph, pcsr: bool = False, reference_unit_ids: bool = False) -> "DeviceGraph":
        return DeviceGraph(self, g, pcsr, reference_unit_ids)

    def set_profiling(self, on: bool):
        self.check(self._L.fsig_set_profiling(self._h, int(on)))

    def reset_stats(self):
        self.check(self._L.fsig_stats_reset(self._h))

    def kernel_stats(self, family: int) -> dict:
        la, by, un = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ms = C.c_double()
        self.check(self._L.fsig_kernel_stats(self._h, family, C.byref(la), C.byref(ms), C.byref(by), C.byref(un)))
        return {"launches": la.value, "ms": ms.value, "alg_bytes": by.value, "units": un.value}

    def set_tune(self, lanes_per_row=0, strategy=0, lds_tile=0, refine_passes=0, fused_tail=0, nlf_filter=0,
                 beam_log2=0, symmetry=0):
        t = Tune(lanes_per_row, strategy, lds_tile, refine_passes, fused_tail, nlf_filter, beam_log2, symmetry)
        self.check(self._L.fsig_set_tune(self._h, C.byref(t)))

    def get_tune(self) -> dict:
        t = Tune()
        self.check(self._L.fsig_get_tune(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tune._fields_}

    def last_candidates(self, k: int) -> np.ndarray:
        out = np.zeros(k, np.uint32)
        self.check(self._L.fsig_last_candidates(self._h, k, _ptr(out)))
        return out

    def keep_candidates(self, on: bool = True):
        """Keep host copies of every match's candidate sets (filter and
        refinement stages) for last_candidate_set."""
        self.check(self._L.fsig_set_keep_candidates(self._h, int(on)))

    def last_candidate_set(self, u: int, stage: int = 1) -> np.ndarray:
        """Ascending candidate vertices of query vertex u in the last match
        (stage 0: after the filter, 1: after refinement)."""
        size = C.c_uint32()
        self.check(self._L.fsig_last_candidate_set(self._h, u, stage, _ptr(None), 0, C.byref(size)))
        out = np.empty(size.value, np.uint32)
        self.check(self._L.fsig_last_candidate_set(self._h, u, stage, _ptr(out), size.value, C.byref(size)))
        return out

    def set_work_limit(self, max_rows: int):
        self.check(self._L.fsig_set_work_limit(self._h, max_rows))

    def stream_copy(self, nbytes: int, iters: int = 20) -> float:
        g = C.c_double()
        self.check(self._L.fsig_stream_copy(self._h, nbytes, iters, C.byref(g)))
        return g.value

    def last_peak_bytes(self) -> int:
        return int(self._L.fsig_last_peak_bytes(self._h))


class DeviceGraph:
    """A data graph resident in HBM (CSR, and PCSR + signatures when asked)."""

    def __init__(self, ctx: Context, g: HostGraph, pcsr: bool, reference_unit_ids: bool):
        self.ctx, self.host = ctx, g
        off, adj, lab = g.arrays()
        idx, lst = g.inverted_index()
        h = _vp()
        ctx.check(ctx._L.fsig_graph_upload_csr(ctx._h, g.n, _ptr(off), _ptr(adj), _ptr(lab, _i32p),
                                               g.num_labels, _ptr(idx), _ptr(lst), C.byref(h)))
        self._h = h
        self.has_pcsr = False
        if pcsr:
            self.attach_pcsr(reference_unit_ids)

    def attach_pcsr(self, reference_unit_ids: bool = False):
        p = self.host.pcsr(reference_unit_ids)
        U = p.num_units
        keys = np.zeros(U, np.uint32)
        ned = np.zeros(U, np.uint32)
        offs, edges = [], []
        for u in range(U):
            o, e = p.unit(u)
            keys[u], ned[u] = len(o) // 32, len(e)
            offs.append(np.ascontiguousarray(o, dtype=np.uint32))
            edges.append(np.ascontiguousarray(e, dtype=np.uint32))
        op = (_u32p * U)(*[_ptr(o) for o in offs])
        ep = (_u32p * U)(*[_ptr(e) for e in edges])
        sig = self.host.signature(column_oriented=True)
        self.ctx.check(self.ctx._L.fsig_graph_upload_pcsr(self.ctx._h, self._h, U, _ptr(keys), op, _ptr(ned),
                                                          ep, _ptr(sig), int(reference_unit_ids)))
        self.has_pcsr = True

    def free(self):
        if getattr(self, "_h", None):
            self.ctx._L.fsig_graph_free(self.ctx._h, self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class MatchResult:
    count: int
    embeddings: Optional[np.ndarray] = None  # rows x k, query-vertex order
    candidates: Optional[np.ndarray] = None


def match(dg: DeviceGraph, q: Query, alg: int = ALG_MMATCH, mode: int = MODE_COUNT, max_embeddings: int = 0,
          emb_cap: int = 0) -> MatchResult:
    """One fsig_match call."""
    ctx = dg.ctx
    off, adj = q.csr()
    sig = q.signature() if alg == ALG_GSI else None
    cnt = C.c_uint64(0)
    out = np.empty((emb_cap, q.k), np.uint32) if emb_cap else None
    ctx.check(ctx._L.fsig_match(ctx._h, dg._h, q.k, _ptr(q.labels, _i32p), _ptr(off), _ptr(adj), _ptr(sig),
                                alg, mode, max_embeddings, C.byref(cnt), _ptr(out), emb_cap))
    res = MatchResult(cnt.value)
    if out is not None:
        res.embeddings = out[: min(cnt.value, emb_cap)]
    res.candidates = ctx.last_candidates(q.k)
    return res


def match_digest(dg: DeviceGraph, q: Query, alg: int = ALG_MMATCH) -> Tuple[int, Tuple[int, int]]:
    """fsig_match_digest: (count, order-independent digest of the full
    embedding set), every embedding enumerated on the device."""
    ctx = dg.ctx
    off, adj = q.csr()
    sig = q.signature() if alg == ALG_GSI else None
    cnt = C.c_uint64(0)
    dig = np.zeros(2, np.uint64)
    ctx.check(ctx._L.fsig_match_digest(ctx._h, dg._h, q.k, _ptr(q.labels, _i32p), _ptr(off), _ptr(adj), _ptr(sig),
                                       alg, C.byref(cnt), _ptr(dig, _u64p)))
    return cnt.value, (int(dig[0]), int(dig[1]))


def match_batch(dg: "DeviceGraph", queries: Sequence[Query], bounds=None, status: bool = False):
    """fsig_match_batch (bounds None) or fsig_match_batch_bounded: counts of
    every query (alg 1 semantics; with bounds[i] != 0, min(count,
    bounds[i])) and the join's device time in ms; with status, also each
    query's status (FSIG_ELIMIT: refused by the context's work limit)."""
    ctx = dg.ctx
    nq = len(queries)
    ks = np.array([q.k for q in queries], np.uint32)
    csrs = [q.csr() for q in queries]
    labs = [np.ascontiguousarray(q.labels, dtype=np.int32) for q in queries]
    lp = (_i32p * nq)(*[_ptr(l, _i32p) for l in labs])
    op = (_u32p * nq)(*[_ptr(o) for o, _ in csrs])
    ap = (_u32p * nq)(*[_ptr(a if len(a) else np.zeros(1, np.uint32)) for _, a in csrs])
    counts = np.zeros(nq, np.uint64)
    jm = C.c_double(0)
    st = np.zeros(nq, np.int32)
    if bounds is None:
        ctx.check(ctx._L.fsig_match_batch(ctx._h, dg._h, nq, _ptr(ks), lp, op, ap, _ptr(counts, _u64p),
                                          C.byref(jm)))
    else:
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(bounds, dtype=np.uint64), (nq,)))
        ctx.check(ctx._L.fsig_match_batch_bounded(ctx._h, dg._h, nq, _ptr(ks), lp, op, ap, _ptr(b, _u64p),
                                                  _ptr(counts, _u64p), _ptr(st, _i32p), C.byref(jm)))
        if not status and np.any(st != 0):
            raise FsigError(int(st[st != 0][0]), "a batch query exceeded the work limit")
    return (counts, jm.value, st) if status else (counts, jm.value)


BATCH_ENGINES = {1: "lockstep", 2: "persistent"}


def _query_arrays(q: "Query"):
    off, adj = q.csr()
    lab = np.ascontiguousarray(q.labels, dtype=np.int32)
    adj = adj if len(adj) else np.zeros(1, np.uint32)
    return lab, off, adj


def project_bytes(n: int, label_count: np.ndarray, pair_edges: np.ndarray, q: "Query") -> int:
    """fsig_project_bytes: the arena bytes the served loop admits `q` with on a
    graph of n vertices with these label and label-pair adjacency counts."""
    L = hip_lib()
    lc = np.ascontiguousarray(label_count, dtype=np.uint64)
    pe = np.ascontiguousarray(pair_edges, dtype=np.uint64).reshape(-1)
    if pe.size != lc.size * lc.size:
        raise FsigError(FSIG_EINVAL, "pair_edges must hold num_labels^2 counts")
    lab, off, adj = _query_arrays(q)
    out = C.c_uint64()
    rc = L.fsig_project_bytes(n, lc.size, _ptr(lc, _u64p), _ptr(pe, _u64p), q.k, _ptr(lab, _i32p), _ptr(off),
                              _ptr(adj), C.byref(out))
    if rc:
        raise FsigError(rc, "fsig_project_bytes failed")
    return out.value


def label_pairs(dg: "DeviceGraph") -> np.ndarray:
    """[L][L] adjacency-entry counts between labels of a resident graph."""
    ctx = dg.ctx
    L = dg.host.num_labels
    out = np.zeros(L * L, np.uint64)
    ctx.check(ctx._L.fsig_graph_label_pairs(ctx._h, dg._h, _ptr(out, _u64p)))
    return out.reshape(L, L)


class Server:
    """fsig_server: continuous batching of count queries over a resident
    graph (workers with their own stream and arena slice; admission by
    projected intermediate-table bytes)."""

    def __init__(self, dg: "DeviceGraph", workers: int = 4, hbm_budget: int = 0, max_batch: int = 64,
                 tune: Tune = None):
        self._L = hip_lib()
        self._h = None
        self.dg = dg  # keeps the graph alive
        h = _vp()
        rc = self._L.fsig_server_create(dg.ctx.device, dg._h, workers, hbm_budget, max_batch,
                                        C.byref(tune) if tune is not None else None, C.byref(h))
        if rc:
            raise FsigError(rc, "fsig_server_create failed")
        self._h = h

    def set_work_limit(self, max_rows: int):
        rc = self._L.fsig_server_set_work_limit(self._h, max_rows)
        if rc:
            raise FsigError(rc, "fsig_server_set_work_limit failed")

    def submit(self, q: "Query", bound: int = 0) -> int:
        """Queues a count query (bound != 0: min(count, bound)); returns its ticket."""
        lab, off, adj = _query_arrays(q)
        t = C.c_uint64()
        rc = self._L.fsig_server_submit_bounded(self._h, q.k, _ptr(lab, _i32p), _ptr(off), _ptr(adj), bound,
                                                C.byref(t))
        if rc:
            raise FsigError(rc, self._L.fsig_server_last_error(self._h).decode())
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> Tuple[int, float]:
        """(count, latency ms) of a ticket; FsigError(FSIG_EAGAIN) on timeout."""
        cnt, lat = C.c_uint64(), C.c_double()
        rc = self._L.fsig_server_wait(self._h, ticket, timeout_ms, C.byref(cnt), C.byref(lat))
        if rc:
            raise FsigError(rc, self._L.fsig_server_last_error(self._h).decode() or "server wait failed")
        return cnt.value, lat.value

    def projected(self, q: "Query") -> int:
        lab, off, adj = _query_arrays(q)
        out = C.c_uint64()
        rc = self._L.fsig_server_projected(self._h, q.k, _ptr(lab, _i32p), _ptr(off), _ptr(adj), C.byref(out))
        if rc:
            raise FsigError(rc, "fsig_server_projected failed")
        return out.value

    def stats(self) -> dict:
        st = ServerStats()
        rc = self._L.fsig_server_get_stats(self._h, C.byref(st))
        if rc:
            raise FsigError(rc, "fsig_server_get_stats failed")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_ if f != "pad"}

    def close(self):
        if self._h:
            self._L.fsig_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def last_batch_tune(ctx: "Context") -> Tuple[int, int, str]:
    """(lanes per row/group, LDS tile words, engine) of the last batch."""
    g, t, e = C.c_uint32(), C.c_uint32(), C.c_uint32()
    ctx.check(ctx._L.fsig_last_batch_tune(ctx._h, C.byref(g), C.byref(t), C.byref(e)))
    return g.value, t.value, BATCH_ENGINES.get(e.value, "none")


class GPUSubgraphMatcher:
    """Mirror of the reference's abstract matcher (match/subgraph_match.h:39-81):
    holds the query and data graph, ``match()`` runs on the device."""
    alg = ALG_MMATCH

    def __init__(self, query_graph: Query, data_graph: HostGraph, dev_id: int = 0, ctx: Context = None,
                 device_graph: DeviceGraph = None):
        self.query_graph_ = query_graph
        self.data_graph_ = data_graph
        self.dev_id_ = dev_id
        self.ctx = ctx
        self.dg = device_graph
        self.num_embeddings_ = 0
        self.embeddings_ = None

    def init_GPU(self, dev_id: int, hbm_budget: int = 1 << 31):
        if self.ctx is None:
            self.ctx = Context(dev_id, hbm_budget)

    def match(self, mode: int = MODE_ENUMERATE, max_embeddings: int = 0, emb_cap: int = 0) -> MatchResult:
        self.init_GPU(self.dev_id_)
        if self.dg is None:
            self.dg = self.ctx.upload(self.data_graph_, pcsr=self.alg == ALG_GSI)
        r = match(self.dg, self.query_graph_, self.alg, mode, max_embeddings, emb_cap)
        self.num_embeddings_ = r.count
        self.embeddings_ = r.embeddings
        return r


class MMatch(GPUSubgraphMatcher):
    """Alg 1 (match/MMatch.h:75-230)."""
    alg = ALG_MMATCH


class GSI(GPUSubgraphMatcher):
    """Alg 2 (match/GSI.h:47-101)."""
    alg = ALG_GSI
