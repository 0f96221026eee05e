"""Benchmark of the MI355X subgraph-matching engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]
    python bench.py --workload c3|c4|c5      (other GPU workloads)
    python bench.py --workload ingest        (host ingest, no GPU)

A step is one pass of the hot path over the C2 workload (the dblp-shaped
configuration the headline metric is quoted on): a Chung-Lu data graph of
317,080 vertices / 1,049,866 edges with 16 Zipf(1.2) labels, resident in HBM,
and its two 8-vertex random-walk queries (dense and sparse), each matched by
alg 1 (fsig_match, FSIG_MODE_COUNT: every embedding is enumerated level by
level; the last level is counted, not written).  Inputs are resident before
the timed region; every fsig_match returns only after its stream drained.

Prints one JSON line.  `value` is embeddings/s of the timed steps; `roofline`
describes the dominant kernel family measured with HIP events on the
engine's stream during the timed region; `cpu_baseline` is the oracle
(oracle/oracle.c, a restatement, "port") on the host cores with a bounded
sample of the same queries.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "embeddings/s & queries/s (batched vs per-query), p50/p99 latency, % HBM peak"
HBM_PEAK_GBS = 8000.0
KERNEL_OF_FAMILY = {"join_tail": "k_join_tail (fused last two levels)", "join_count": "k_join (count pass)",
                    "join_write": "k_join (write pass)", "refine": "k_refine", "cedge": "k_cedge_count/fill",
                    "filter": "k_ldf", "batch": "k_join_multi"}
COUNTERS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "c2_hbm_counters.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), or None when the summary does not cover it."""
    try:
        with open(COUNTERS) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else int(k["fetch_bytes"] + k["write_bytes"])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name, from /proc/cpuinfo (None where unreadable)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(g, queries, seconds_target=20.0):
    """Oracle on the host: the dense query is counted in full (it also checks
    the GPU count at full size); the sparse query is sampled: the search
    stops after a bounded number of embeddings, sized from a short probe to
    take about the rest of the time budget."""
    import oracle as O
    off, adj, lab = g.arrays()
    threads = max(1, min(16, os.cpu_count() or 1))
    total_emb, total_s, parts, checks = 0, 0.0, [], {}
    for name, q in queries:
        if name == "dense":
            t = time.perf_counter()
            cnt, _, stopped = O.match_partial(q.labels, q.adjm(), off, adj, lab, threads=threads, limit=4_000_000_000)
            dt = time.perf_counter() - t
            total_emb += cnt
            total_s += dt
            parts.append(f"{name}: {'first ' if stopped else 'all '}{cnt} embeddings")
            if not stopped:
                checks[name] = cnt
        else:
            t = time.perf_counter()
            probe, _, _ = O.match_partial(q.labels, q.adjm(), off, adj, lab, threads=threads, limit=20_000_000)
            rate = probe / max(time.perf_counter() - t, 1e-6)
            budget = max(seconds_target - total_s, 5.0)
            limit = int(min(rate * budget, 50_000_000_000))
            t = time.perf_counter()
            cnt, _, stopped = O.match_partial(q.labels, q.adjm(), off, adj, lab, threads=threads, limit=limit)
            dt = time.perf_counter() - t
            total_emb += cnt
            total_s += dt
            parts.append(f"{name}: {'first ' if stopped else 'all '}{cnt} embeddings")
            if not stopped:
                checks[name] = cnt
    # one thread on the sparse query, bounded to about an eighth of the budget
    name, q = queries[-1]
    rate1 = total_emb / total_s / threads
    limit1 = max(1_000_000, int(rate1 * seconds_target / 8))
    t = time.perf_counter()
    cnt1, _, stopped1 = O.match_partial(q.labels, q.adjm(), off, adj, lab, threads=1, limit=limit1)
    dt1 = time.perf_counter() - t
    return {"value": total_emb / total_s, "unit": "embeddings/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "single_thread_value": cnt1 / max(dt1, 1e-9),
            "single_thread_sample": f"{name}: {'first ' if stopped1 else 'all '}{cnt1} embeddings in {dt1:.1f} s",
            "sample": "oracle backtracking (oracle/oracle.c), roots handed out dynamically to "
                      f"{threads} threads, {total_s:.1f} s: " + "; ".join(parts)}, checks


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5", "ingest"])
    ap.add_argument("--qps", type=float, default=0.0, help="c5: arrival rate (0: the workload's)")
    ap.add_argument("--requests", type=int, default=0, help="c5: stream length (0: the workload's)")
    ap.add_argument("--workers", type=int, default=4, help="c5: server worker contexts")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--ingest-c5", action="store_true",
                    help="ingest: also time the C5 graph (R-MAT 2^24 V / 200M E, a 4 GB text file)")
    ap.add_argument("--alg", type=int, default=1)
    ap.add_argument("--copy-only", type=int, default=0,
                    help="only run the STREAM copy kernel over this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU: the hot path is single-device; "
                                   "--gpus N>1 is not supported"}))
        sys.exit(2)

    if args.workload == "ingest":
        return bench_ingest(args)

    import numpy as np
    import fsig_amd as F
    from fsig_amd import workloads as W

    if args.copy_only:
        ctx = F.Context(0, 2 * args.copy_only + (64 << 20))
        gbps = ctx.stream_copy(args.copy_only, 5)
        print(json.dumps({"stream_copy_bytes": args.copy_only, "iters": 6, "GB_per_s": round(gbps, 1)}))
        return

    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)
    if args.workload == "c3":
        return bench_c3(args)

    t0 = time.perf_counter()
    g = W.c2_graph()
    queries = W.c2_queries(g)
    log(f"graph {g.n} V {g.nnz // 2} E {g.num_labels} labels, generated in {time.perf_counter() - t0:.2f}s")
    ctx = F.Context(0, 64 << 30)
    dg = ctx.upload(g, pcsr=True)
    counts = {}
    for name, q in queries:
        counts[name] = F.match(dg, q, args.alg, F.MODE_COUNT).count
    log("counts", counts)

    def step(lat):
        n = 0
        for _, q in queries:
            t = time.perf_counter()
            n += F.match(dg, q, args.alg, F.MODE_COUNT).count
            lat.append(time.perf_counter() - t)
        return n

    for _ in range(args.warmup):
        step([])
    ctx.set_profiling(True)
    ctx.reset_stats()
    lat = []
    ctx.synchronize()
    t = time.perf_counter()
    emb = 0
    for _ in range(args.steps):
        emb += step(lat)
    ctx.synchronize()
    elapsed = time.perf_counter() - t
    stats = {F.KERNEL_FAMILIES[f]: ctx.kernel_stats(f) for f in range(len(F.KERNEL_FAMILIES))}
    ctx.set_profiling(False)
    ms_per_step = elapsed / args.steps * 1e3
    value = emb / elapsed

    kernels = {}
    for fam, s in stats.items():
        if s["launches"]:
            gbs = s["alg_bytes"] / (s["ms"] / 1e3) / 1e9 if s["ms"] > 0 else 0.0
            kernels[fam] = {"launches": s["launches"], "ms": round(s["ms"], 3), "alg_GB": round(s["alg_bytes"] / 1e9, 3),
                            "GB_per_s": round(gbs, 1)}
    dom = max((f for f in stats if stats[f]["launches"]), key=lambda f: stats[f]["ms"])
    d = stats[dom]
    achieved = d["alg_bytes"] / (d["ms"] / 1e3) / 1e9
    kname = KERNEL_OF_FAMILY.get(dom, dom)
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(kname),
                "avg_launch_ms": round(d["ms"] / d["launches"], 4),
                "alg_bytes_per_launch": int(d["alg_bytes"] / d["launches"])}

    # alg 2 (GSI path) on the same queries, outside the timed region
    gsi = {}
    if args.alg == 1:
        t = time.perf_counter()
        n2 = 0
        for _, q in queries:
            c2 = F.match(dg, q, F.ALG_GSI, F.MODE_COUNT).count
            n2 += c2
            assert c2 == counts[_], "alg 2 disagrees with alg 1"
        gsi = {"alg2_embeddings_per_s": round(n2 / (time.perf_counter() - t), 1)}

    stream_gbps = ctx.stream_copy(1 << 30, 20)
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "embeddings/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": "C2 dblp-shaped: Chung-Lu gamma 2.3, 317,080 V, 1,049,866 E, 16 Zipf(1.2) labels; "
                               "8-vertex dense + sparse random-walk queries; alg 1, count mode",
                   "graph_seed": W.C2_GRAPH["graph_seed"], "label_seed": W.C2_GRAPH["label_seed"],
                   "queries": {n: {"edges": q.num_edges, "seed": W.C2_QUERIES[n]["seed"], "embeddings": counts[n]}
                               for n, q in queries},
                   "parallelism": "1 GPU"},
        "queries_per_s": round(len(lat) / elapsed, 2), "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3), "hbm_frac": roofline["frac"],
        "roofline": roofline, "stream_copy_GB_per_s": round(stream_gbps, 1), "kernels": kernels, **gsi,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        cb, checks = cpu_baseline(g, queries)
        for n, c in checks.items():
            assert c == counts[n], f"oracle count {c} != engine count {counts[n]} for {n}"
        out["cpu_baseline"] = cb
        out["cpu_checked_counts"] = checks
    print(json.dumps(out), flush=True)
    dg.free()
    ctx.close()


def bench_c4(args):
    """C4: 256 random-walk queries (k uniform in 4..16, half dense, half
    sparse) on a 1M-vertex / 10M-edge Chung-Lu graph with 32 Zipf(1.2)
    labels.  The batched engine (one fsig_match_batch call: batched setup +
    one persistent, graph-launched join) against the plain loop of one
    fsig_match per query on the same GPU."""
    import numpy as np
    import fsig_amd as F
    from fsig_amd import workloads as W
    t0 = time.perf_counter()
    g = W.c4_graph()
    ctx = F.Context(0, 96 << 30)
    dg = ctx.upload(g)
    queries = W.c4_queries(g, dg, progress=log)
    log(f"c4 graph {g.n} V {g.nnz // 2} E, {len(queries)} queries, {time.perf_counter() - t0:.1f}s")
    # plain loop (one warm-up pass, one timed pass)
    lat, plain_counts = [], []
    for q in queries:
        F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT)
    ctx.synchronize()
    t = time.perf_counter()
    for q in queries:
        t1 = time.perf_counter()
        plain_counts.append(F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count)
        lat.append(time.perf_counter() - t1)
    plain_s = time.perf_counter() - t
    log(f"plain loop {plain_s * 1e3:.1f} ms, {sum(plain_counts)} embeddings")
    for _ in range(max(1, args.warmup)):
        counts, _ = F.match_batch(dg, queries)
    assert counts.tolist() == plain_counts, "batched counts differ from the plain loop"
    ctx.set_profiling(True)
    ctx.reset_stats()
    ctx.synchronize()
    t = time.perf_counter()
    join_ms = []
    for _ in range(args.steps):
        counts, jm = F.match_batch(dg, queries)
        join_ms.append(jm)
    ctx.synchronize()
    batch_s = (time.perf_counter() - t) / args.steps
    st = ctx.kernel_stats(F.K_BATCH)
    lanes, tile, engine = F.last_batch_tune(ctx)
    lat_ms = np.array(lat) * 1e3
    emb = int(sum(plain_counts))
    out = {"metric": METRIC, "value": round(len(queries) / batch_s, 2), "unit": "queries/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(batch_s * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
           "config": {"workload": "C4 batched: Chung-Lu 1M V / 10M E, 32 Zipf(1.2) labels; 256 random-walk queries, "
                                  "k uniform in 4..16, half dense half sparse; one fsig_match_batch per step",
                      "embeddings_per_step": emb, "parallelism": "1 GPU"},
           "plain_queries_per_s": round(len(queries) / plain_s, 2),
           "batched_over_plain": round(plain_s / batch_s, 2),
           "batched_embeddings_per_s": round(emb / batch_s, 1),
           "plain_p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
           "plain_p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
           "persistent_join_ms": round(float(np.mean(join_ms)), 3),
           "persistent_join_alg_GB_per_s": round(st["alg_bytes"] / (st["ms"] / 1e3) / 1e9, 1) if st["ms"] else None,
           "autotuned": {"engine": engine, "lanes": lanes, "lds_tile_words": tile}}
    print(json.dumps(out), flush=True)


def bench_c5(args):
    """C5: a Poisson stream of random-walk count queries (k log-uniform in
    4..32) on an R-MAT 2^24-vertex / 200M-edge graph with 64 uniform labels,
    resident in HBM.  The served loop (continuous batching over worker
    streams, admission by projected bytes) against one-at-a-time plain
    launches of the same stream in arrival order; latency is arrival to
    completion in both."""
    import threading
    import numpy as np
    import fsig_amd as F
    from fsig_amd import workloads as W
    t0 = time.perf_counter()
    g = W.c5_graph()
    log(f"c5 graph {g.n} V {g.nnz // 2} E in {time.perf_counter() - t0:.1f}s")
    ctx = F.Context(0, 32 << 30)
    dg = ctx.upload(g)
    t0 = time.perf_counter()
    stream = W.c5_stream(g, dg, args.requests, args.qps, progress=log)
    arrivals = np.array([a for a, _ in stream])
    qs = [q for _, q in stream]
    log(f"{len(qs)} requests drawn in {time.perf_counter() - t0:.1f}s, offered {len(qs) / arrivals[-1]:.0f} q/s")
    # one-at-a-time plain launches, in arrival order
    for q in qs[:8]:
        F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT)
    ctx.synchronize()
    plain_counts, plain_lat = [], []
    start = time.perf_counter()
    for a, q in zip(arrivals, qs):
        now = time.perf_counter() - start
        if now < a:
            time.sleep(a - now)
        plain_counts.append(F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count)
        plain_lat.append((time.perf_counter() - start - a) * 1e3)
    plain_s = time.perf_counter() - start
    log(f"plain: {plain_s * 1e3:.0f} ms, p50 {np.percentile(plain_lat, 50):.1f} ms")
    budget = 128 << 30
    srv = F.Server(dg, workers=args.workers, hbm_budget=budget, max_batch=64)
    # warm-up pass through the server (first-launch costs)
    for t in [srv.submit(q) for q in qs[:32]]:
        srv.wait(t)
    srv.close()
    srv = F.Server(dg, workers=args.workers, hbm_budget=budget, max_batch=64)
    tickets = [None] * len(qs)

    def feed(start):
        for i, (a, q) in enumerate(zip(arrivals, qs)):
            now = time.perf_counter() - start
            if now < a:
                time.sleep(a - now)
            tickets[i] = srv.submit(q)

    start = time.perf_counter()
    th = threading.Thread(target=feed, args=(start,))
    th.start()
    th.join()
    res = [srv.wait(t) for t in tickets]
    served_s = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    served_counts = [c for c, _ in res]
    served_lat = np.array([l for _, l in res])
    assert served_counts == plain_counts, "served counts differ from the plain launches"
    plain_lat = np.array(plain_lat)
    out = {"metric": METRIC, "value": round(len(qs) / served_s, 2), "unit": "queries/s", "n_gpus": 1,
           "steps": 1, "warmup": 1, "ms_per_step": round(served_s * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
           "config": {"workload": f"C5 served: R-MAT 2^24 V / {g.nnz // 2} E, 64 uniform labels, resident; "
                                  f"{len(qs)} Poisson requests at {len(qs) / arrivals[-1]:.0f} q/s offered, "
                                  "k = round(exp(U(ln 4, ln 32))) conditioned on k <= 5 (larger walk queries "
                                  "exceed 1e8 embeddings / 2e7 partial rows on this graph), count mode",
                      "workers": args.workers, "hbm_budget_GiB": budget >> 30, "max_batch": 64,
                      "parallelism": "1 GPU"},
           "served_p50_ms": round(float(np.percentile(served_lat, 50)), 3),
           "served_p99_ms": round(float(np.percentile(served_lat, 99)), 3),
           "plain_queries_per_s": round(len(qs) / plain_s, 2),
           "plain_p50_ms": round(float(np.percentile(plain_lat, 50)), 3),
           "plain_p99_ms": round(float(np.percentile(plain_lat, 99)), 3),
           "served_over_plain": round(plain_s / served_s, 2),
           "k_histogram": {str(k): int(sum(1 for q in qs if q.k == k)) for k in sorted({q.k for q in qs})},
           "embeddings": int(sum(plain_counts)),
           "server": {k: (round(v, 1) if isinstance(v, float) else v) for k, v in st.items()}}
    print(json.dumps(out), flush=True)


def bench_ingest(args):
    """Host ingest of the labeled workload graphs as text files (no GPU):
    the parallel reader + CSR build, the PCSR (reference unit ids), the data
    signature and the binary CSR cache round trip, each the best of
    `--steps` runs.  `value` is text MB/s through read + CSR + PCSR +
    signature on the C2 file; `cpu_baseline` is the reference's own reader,
    transforms and signature (oracle/_ref/ref_dump, compiled from the
    reference sources; kind "reference") on the same file.  The C4 file
    (1M V / 10M E, 32 Zipf labels) is timed for this engine only: the
    reference's O(K^2) PCSR build takes minutes there."""
    import tempfile
    import fsig_amd as F
    from fsig_amd import workloads as W
    import oracle as O
    steps = max(1, min(args.steps, 5))
    res = {}
    with tempfile.TemporaryDirectory() as td:
        graphs = [("c2", W.c2_graph), ("c4", W.c4_graph)]
        if args.ingest_c5:
            graphs.append(("c5", W.c5_graph))
        for name, make in graphs:
            g = make()
            path = os.path.join(td, f"{name}.graph")
            g.write(path)
            n_v, n_e = g.n, g.nnz // 2
            del g
            mb = os.path.getsize(path) / 1e6
            best = {}

            def timed(key, fn):
                t = time.perf_counter()
                out = fn()
                best[key] = min(best.get(key, 1e30), time.perf_counter() - t)
                return out

            for _ in range(1 if name == "c5" else steps):
                h = None  # the previous run's graph is freed before the next read
                h = timed("read_csr", lambda: F.HostGraph.read(path))
                timed("pcsr", lambda: h.pcsr(reference_unit_ids=True))
                timed("signature", lambda: h.signature(True))
                binp = os.path.join(td, f"{name}.csr.bin")
                timed("bin_save", lambda: h.save_bin(binp))
                timed("bin_load", lambda: F.HostGraph.load_bin(binp))
                os.remove(binp)
            del h
            if name != "c2":  # the C2 file stays for the reference's run below
                os.remove(path)
            ours = best["read_csr"] + best["pcsr"] + best["signature"]
            res[name] = {"text_MB": round(mb, 1), "vertices": n_v, "edges": n_e,
                         "seconds": {k: round(v, 4) for k, v in best.items()}, "MB_per_s": round(mb / ours, 1)}
            log(name, res[name])
            if name == "c2" and not args.no_cpu and O.ref_available():
                t = O.ref_ingest_seconds(path, True)
                ref_total = sum(t.values())
                res["reference_c2"] = {"seconds": t, "MB_per_s": round(mb / ref_total, 1)}
                log("reference", res["reference_c2"])
    c2 = res["c2"]
    out = {"metric": "host ingest MB/s (text -> CSR + PCSR + signature)", "value": c2["MB_per_s"],
           "unit": "MB/s", "n_gpus": 0, "steps": steps, "warmup": 0,
           "ms_per_step": round(1000 * (c2["seconds"]["read_csr"] + c2["seconds"]["pcsr"] + c2["seconds"]["signature"]), 1),
           "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
           "config": {"workload": "ingest: C2 text file (317,080 V, 1,049,866 E, 16 Zipf labels); C4 file for this "
                                  "engine only", "threads": min(16, os.cpu_count() or 1)},
           "stages": res}
    if "reference_c2" in res:
        out["cpu_baseline"] = {"value": res["reference_c2"]["MB_per_s"], "unit": "MB/s", "cores": 1,
                               "kind": "reference", "cpu_model": cpu_model(),
                               "sample": "the reference's reader, PCSR and CSR transforms and data signature "
                                         "(oracle/_ref/ref_dump time) on the same C2 file, one run"}
    print(json.dumps(out))


def bench_c3(args):
    """C3: unlabeled R-MAT (0.57, 0.19, 0.19), 2^22 vertices, 35M edges.
    Full u64 counts of the triangle and the 4-clique, and for all four
    patterns the bounded enumeration of the first 1e8 embeddings (rows come
    back to the host; a sample is checked there: every query edge present,
    vertices distinct).  The 4-cycle and house counts of this graph have
    more than 1e13 embeddings each and are counted on the scale-16 and -18
    analogues (same generator and density) instead."""
    import numpy as np
    import fsig_amd as F
    from fsig_amd import workloads as W
    t0 = time.perf_counter()
    g = W.c3_graph()
    off, adj, _ = g.arrays()
    log(f"c3 graph {g.n} V {g.nnz // 2} E max degree {g.max_degree} in {time.perf_counter() - t0:.1f}s")
    ctx = F.Context(0, 160 << 30)
    dg = ctx.upload(g)
    pats = dict(W.c3_queries())
    out = {"counts": {}, "count_ms": {}, "enumerate": {}}
    F.match(dg, pats["triangle"], F.ALG_MMATCH, F.MODE_COUNT)  # warm-up
    for name in ("triangle", "clique4"):
        q = pats[name]
        log(f"counting {name} ...")
        t = time.perf_counter()
        c = F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT).count
        dt = time.perf_counter() - t
        out["counts"][name] = c
        out["count_ms"][name] = round(dt * 1e3, 2)
        log(f"{name}: {c} embeddings in {dt * 1e3:.1f} ms")
    cap = 10 ** 8
    rng = np.random.default_rng(0)
    for name, q in pats.items():
        log(f"enumerating the first {cap} embeddings of {name} ...")
        t = time.perf_counter()
        r = F.match(dg, q, F.ALG_MMATCH, F.MODE_ENUMERATE, max_embeddings=cap, emb_cap=cap)
        dt = time.perf_counter() - t
        rows = r.embeddings
        pick = rng.integers(0, len(rows), size=min(10000, len(rows)))
        ok = True
        for i in pick:
            row = rows[i]
            if len(set(row.tolist())) != q.k:
                ok = False
            for a, b in q.edges.tolist():
                u, v = int(row[a]), int(row[b])
                lo, hi = off[u], off[u + 1]
                j = lo + np.searchsorted(adj[lo:hi], v)
                ok &= bool(j < hi and adj[j] == v)
        out["enumerate"][name] = {"rows": int(len(rows)), "ms": round(dt * 1e3, 1),
                                  "rows_per_s": round(len(rows) / dt, 1), "sample_checked": int(len(pick)),
                                  "sample_valid": ok}
        log(f"{name}: first {len(rows)} embeddings in {dt * 1e3:.1f} ms, sample valid {ok}")
        assert ok, f"{name}: an enumerated row is not an embedding"
        del rows, r
    dg.free()
    small = {}
    for sc, m in ((16, 35_000_000 // 64), (18, 35_000_000 // 16)):
        gs = W.c3_graph(sc, m)
        ds = ctx.upload(gs)
        for name in ("cycle4", "house") if sc == 16 else ("cycle4",):
            t = time.perf_counter()
            c = F.match(ds, pats[name], F.ALG_MMATCH, F.MODE_COUNT).count
            small[f"{name}_scale{sc}"] = {"count": c, "ms": round((time.perf_counter() - t) * 1e3, 1)}
            log(f"scale {sc} {name}: {c}")
        ds.free()
    total = sum(out["counts"].values())
    secs = sum(out["count_ms"].values()) / 1e3
    res = {"metric": METRIC, "value": round(total / secs, 1), "unit": "embeddings/s", "n_gpus": 1, "steps": 1,
           "warmup": 1, "ms_per_step": round(secs * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32", "data": "synthetic",
           "config": {"workload": f"C3 unlabeled R-MAT (0.57,0.19,0.19) 2^22 V / {g.nnz // 2} E; triangle and "
                                  "4-clique counted (value), first 1e8 embeddings of all four patterns enumerated",
                      "parallelism": "1 GPU"},
           **out, "smaller_scale_counts": small}
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
