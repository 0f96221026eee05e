"""Binary CSR and PCSR caches: round trips, the PCSR byte layout of the
reference's commented STORE block, and rejection of damaged files."""
import os
import struct

import numpy as np
import pytest

import fsig_amd as F


def _graph():
    return F.HostGraph.chung_lu(500, 2000, 2.5, 7).set_labels("zipf", 4, 1.2, 8)


def test_csr_round_trip(tmp_path):
    g = _graph()
    p = str(tmp_path / "g.csr.bin")
    g.save_bin(p)
    h = F.HostGraph.load_bin(p)
    for a, b in zip(g.arrays(), h.arrays()):
        assert np.array_equal(a, b)
    assert (h.n, h.nnz, h.num_labels) == (g.n, g.nnz, g.num_labels)
    # 3 header words, labels, offsets, adjacency
    assert os.path.getsize(p) == 4 * (3 + g.n + g.n + 1 + g.nnz)


@pytest.mark.parametrize("ref_ids", [False, True])
def test_pcsr_round_trip_and_layout(tmp_path, ref_ids):
    g = _graph()
    pc = g.pcsr(reference_unit_ids=ref_ids)
    p = str(tmp_path / "g.pcsr.bin")
    pc.save_bin(p, g)
    back = F.HostPCSR.load_bin(p, ref_ids)
    assert back.num_units == pc.num_units
    for u in range(pc.num_units):
        a, b = pc.unit(u), back.unit(u)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    n, m, nl, lab = back.graph_info()
    assert (n, m, nl) == (g.n, g.nnz // 2, g.num_labels)
    assert np.array_equal(lab, g.arrays()[2])
    # the layout field by field: Vertex, Edge, Label header, labels,
    # num_units, then per unit num_keys, num_edges, offsets, edges
    raw = open(p, "rb").read()
    n0, m0, l0 = struct.unpack_from("<IIi", raw, 0)
    assert (n0, m0, l0) == (g.n, g.nnz // 2, g.num_labels)
    at = 12 + 4 * g.n
    (nu,) = struct.unpack_from("<I", raw, at)
    assert nu == pc.num_units == g.num_labels * (g.num_labels + 1) // 2
    at += 4
    for u in range(nu):
        keys, ne = struct.unpack_from("<II", raw, at)
        off, ed = pc.unit(u)
        assert keys * 32 == len(off) and ne == len(ed)
        at += 8
        assert np.array_equal(np.frombuffer(raw, np.uint32, 32 * keys, at), off)
        at += 4 * 32 * keys
        assert np.array_equal(np.frombuffer(raw, np.uint32, ne, at), ed)
        at += 4 * ne
    assert at == len(raw)


def test_damaged_caches_are_rejected(tmp_path):
    g = _graph()
    p = str(tmp_path / "g.csr.bin")
    g.save_bin(p)
    raw = open(p, "rb").read()
    for bad in (raw[:-4], raw + b"\0\0\0\0", raw[:10]):
        q = str(tmp_path / "bad.bin")
        open(q, "wb").write(bad)
        with pytest.raises(F.FsigError):
            F.HostGraph.load_bin(q)
    with pytest.raises(F.FsigError):
        F.HostGraph.load_bin(str(tmp_path / "missing.bin"))
    pc = g.pcsr()
    pp = str(tmp_path / "g.pcsr.bin")
    pc.save_bin(pp, g)
    open(pp, "ab").write(b"\1\0\0\0")
    with pytest.raises(F.FsigError):
        F.HostPCSR.load_bin(pp)


def test_huge_counts_in_a_header_are_rejected_before_allocation(tmp_path):
    g = _graph()
    p = str(tmp_path / "g.csr.bin")
    g.save_bin(p)
    raw = bytearray(open(p, "rb").read())
    for field in (0, 4):                        # vertex count, edge count
        bad = bytearray(raw)
        bad[field:field + 4] = struct.pack("<I", 0xFFFFFFF0)
        q = str(tmp_path / "huge.bin")
        open(q, "wb").write(bytes(bad))
        with pytest.raises(F.FsigError):
            F.HostGraph.load_bin(q)
    pc = g.pcsr()
    pp = str(tmp_path / "g.pcsr.bin")
    pc.save_bin(pp, g)
    raw = bytearray(open(pp, "rb").read())
    n = struct.unpack_from("<I", raw, 0)[0]
    first_unit = 12 + 4 * n + 4                  # header, labels, unit count
    for off in (first_unit, first_unit + 4):    # key count, edge count of unit 0
        bad = bytearray(raw)
        bad[off:off + 4] = struct.pack("<I", 0x7FFFFFFF)
        q = str(tmp_path / "huge.pcsr.bin")
        open(q, "wb").write(bytes(bad))
        with pytest.raises(F.FsigError):
            F.HostPCSR.load_bin(q)
