"""The FSIG driver's argument and input checks, which all run before it
touches the GPU: a bad command line or an unreadable graph ends with a
message and a nonzero status (the reference leaves its transformer
uninitialised on a bad algorithm, main/run.cc:101-111, and exits 0 on
errors)."""
import os
import subprocess

import pytest

import fsig_amd as F

FSIG = os.path.join(os.path.dirname(F.HIP_LIB_PATH), "bin", "FSIG")


def _run(args):
    if not os.path.exists(FSIG):
        pytest.skip("bin/FSIG not built")
    return subprocess.run([FSIG] + [str(a) for a in args], capture_output=True, text=True, timeout=60)


def _graph(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_too_few_arguments():
    p = _run(["only-one"])
    assert p.returncode != 0 and "usage" in p.stdout


@pytest.mark.parametrize("alg", [0, 3, -1])
def test_bad_algorithm(tmp_path, alg):
    d = _graph(tmp_path, "d.graph", "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1")
    p = _run([d, d, alg])
    assert p.returncode != 0 and "algorithm type" in p.stdout


def test_binary_cache_serves_alg_1_only(tmp_path):
    g = F.HostGraph.chung_lu(200, 600, 2.5, 3)
    b = str(tmp_path / "g.csr.bin")
    g.save_bin(b)
    q = _graph(tmp_path, "q.graph", "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1")
    p = _run([b, q, 2])
    assert p.returncode != 0 and "algorithm 1 only" in p.stdout


@pytest.mark.parametrize("which", ["data", "query"])
def test_unreadable_graphs(tmp_path, which):
    good = _graph(tmp_path, "ok.graph", "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1")
    missing = str(tmp_path / "missing.graph")
    gap = _graph(tmp_path, "gap.graph", "t 3 1\nv 0 0 1\nv 2 0 1\ne 0 2")
    for bad, msg in ((missing, "cannot open"), (gap, "no v line")):
        args = [bad, good] if which == "data" else [good, bad]
        p = _run(args + [1])
        assert p.returncode != 0 and msg in p.stdout, p.stdout


def test_damaged_binary_cache(tmp_path):
    b = _graph(tmp_path, "bad.csr.bin", "not a cache")
    q = _graph(tmp_path, "q.graph", "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1")
    p = _run([b, q, 1])
    assert p.returncode != 0 and "Error" in p.stdout
