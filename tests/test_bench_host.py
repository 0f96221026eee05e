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
