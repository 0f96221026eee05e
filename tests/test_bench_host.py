"""Host-side checks of bench.py: the one-GPU refusal and the CPU baseline's
fields and counts (the baseline is the oracle, timed on the host)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_baseline_fields_and_counts():
    import bench
    import fsig_amd as F
    import oracle as O
    g = F.HostGraph.chung_lu(3000, 12000, 2.3, 5)
    g.set_labels("uniform", 4, seed=1)
    qs = [("dense", g.walk_query(4, 5, 6, seed=3)), ("sparse", g.walk_query(5, 4, 4, seed=4))]
    cb, checks = bench.cpu_baseline(g, qs, seconds_target=0.5)
    for key in ("value", "unit", "cores", "kind", "sample", "cpu_model", "single_thread_value"):
        assert key in cb
    assert cb["kind"] == "port" and cb["unit"] == "embeddings/s" and cb["cores"] >= 1
    assert cb["value"] > 0 and cb["single_thread_value"] > 0
    assert "dense" in checks
    off, adj, lab = g.arrays()
    for name, q in qs:
        if name in checks:
            assert checks[name] == O.match(q.labels, q.adjm(), off, adj, lab)[0]


def test_ingest_workload_line(monkeypatch, capsys):
    # the ingest workload on small graphs: one JSON line with the contract's
    # fields, every stage timed, the reference's stages as the baseline when
    # oracle/_ref is built
    import argparse
    import bench
    import fsig_amd as F
    import oracle as O
    from fsig_amd import workloads as W
    monkeypatch.setattr(W, "c2_graph", lambda: F.HostGraph.chung_lu(3000, 12000, 2.3, 1).set_labels("zipf", 4, 1.2, 2))
    monkeypatch.setattr(W, "c4_graph", lambda: F.HostGraph.chung_lu(5000, 20000, 2.3, 3).set_labels("zipf", 8, 1.2, 4))
    bench.bench_ingest(argparse.Namespace(steps=1, no_cpu=False, ingest_c5=False))
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config"):
        assert key in line
    assert line["unit"] == "MB/s" and line["value"] > 0
    for name in ("c2", "c4"):
        secs = line["stages"][name]["seconds"]
        assert set(secs) == {"read_csr", "pcsr", "signature", "bin_save", "bin_load"}
    if O.ref_available():
        assert line["cpu_baseline"]["kind"] == "reference"
        assert set(line["stages"]["reference_c2"]["seconds"]) == {"read", "pcsr", "csr", "signature"}


def test_roofline_traffic_matches_the_committed_profiles():
    """The traffic bench.py reports for the dominant kernel comes from the
    committed FETCH_SIZE / WRITE_SIZE summary: the fused-tail family is
    there with positive bytes, and every kernel instance the summary names
    appears in the committed C2 alg-1 kernel trace (the two passes profiled
    the same kernels)."""
    import bench
    rec = json.load(open(bench.COUNTERS))
    tail = bench.KERNEL_OF_FAMILY["join_tail"]
    assert bench.measured_traffic(tail) > 0
    cal = rec["calibration"]
    assert abs(cal["fetch_factor"] - 2.0) < 0.01 and abs(cal["write_factor"] - 1.0) < 0.01
    trace = open(os.path.join(ROOT, "profiles", "c2_alg1_kernel_trace.txt")).read()
    for fam, k in rec["kernels"].items():
        assert k["dispatches"] > 0 and k["fetch_bytes"] >= 0 and k["write_bytes"] >= 0, fam
        for inst in k["instances"]:
            assert inst in trace, (fam, inst)
    assert bench.measured_traffic("no such kernel") is None
