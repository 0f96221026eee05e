"""Benchmark of the MI355X subgraph-matching engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W]       (all three GPU workloads, one JSON line)
    python bench.py --workload c2|c4|c5|c3                    (one GPU workload)
    python bench.py --workload ingest                         (host ingest, no GPU)

The default run prints ONE JSON line whose headline is the served workload
(C5), the largest single-GPU configuration of the metric "embeddings/s &
queries/s (batched vs per-query), p50/p99 latency, % HBM peak":

  C5  served: an R-MAT graph of 2^24 vertices / 200M edges with 64 uniform
      labels, resident in HBM; a seeded Poisson stream of 128 random-walk
      count queries (k = round(exp(U(ln 4, ln 32)))), every drawn query kept,
      each counted to a bound of 1e8 embeddings (min(count, 1e8), exact
      below it), a query over the work limit (2e9 partial embeddings)
      refused and reported.  A step serves the whole stream through the
      continuous-batching server (fsig_server_*); `value` is requests/s over
      the K timed steps, with p50/p99 arrival-to-completion latency.  The
      same stream through one-at-a-time fsig_match calls in arrival order is
      the plain baseline (served_over_plain).
  C4  batched: a Chung-Lu graph of 1M vertices / 10M edges with 32 Zipf(1.2)
      labels; the first 256 random-walk queries (k uniform in 4..16, dense
      and sparse alternating), bounded and refused as C5's; one
      fsig_match_batch_bounded call per step against the loop of fsig_match
      calls (batched_over_plain), counts equal.
  C2  per query: the dblp-shaped Chung-Lu graph (317,080 V / 1,049,866 E, 16
      Zipf(1.2) labels) and its 8-vertex dense and sparse random-walk
      queries counted exactly by alg 1; embeddings/s, and the roofline of
      the dominant kernel family (HIP events on the engine's stream over the
      timed steps; algorithmic bytes counted on the device) with the HBM
      traffic rocprofv3 measured for it (profiles/).  `cpu_baseline` is the
      oracle (oracle/oracle.c, a restatement: "port") on the host's cores
      over a bounded sample of the sparse query: one warm-up, the median of
      5 runs, on all cores and on one.

Inputs are resident in HBM before every timed region; every call returns
after its stream drained; results are compared across engines before they
are reported.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "embeddings/s & queries/s (batched vs per-query), p50/p99 latency, % HBM peak"
HBM_PEAK_GBS = 8000.0
KERNEL_OF_FAMILY = {"join_tail": "k_join_tail (fused last two levels)", "join_count": "k_join (count pass)",
                    "join_write": "k_join (write pass)", "refine": "k_refine", "cedge": "k_seg_cedge + k_row_count/fill",
                    "filter": "k_ldf", "batch": "k_join_multi"}
COUNTERS = os.path.join(ROOT, "profiles", "c2_hbm_counters.json")


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
    print(time.strftime("%H:%M:%S"), *a, file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU as lscpu describes it: model name, then family / model /
    stepping, sockets x cores x threads and the hypervisor (a virtual
    machine may report a shortened model name; family and model still
    identify the part); /proc/cpuinfo's model name where lscpu does not
    answer; None where neither does."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        f = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                f[k.strip()] = v.strip()
        if "Model name" in f:
            parts = [f["Model name"]]
            if "CPU family" in f:
                parts.append(f"family {f['CPU family']} model {f.get('Model', '?')} stepping {f.get('Stepping', '?')}")
            if "Socket(s)" in f:
                parts.append(f"{f['Socket(s)']} socket(s) x {f.get('Core(s) per socket', '?')} cores x "
                             f"{f.get('Thread(s) per core', '?')} threads, {f.get('CPU(s)', '?')} CPUs")
            if "Hypervisor vendor" in f:
                parts.append(f"{f['Hypervisor vendor']} guest")
            return "; ".join(parts)
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def pct(a, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(a, dtype=float), q)), 3) if len(a) else None


# ---------------------------------------------------------------------------
# C2: per-query embeddings/s, roofline, CPU baseline
def cpu_baseline(g, queries, seconds_target=4.0):
    """The oracle on the host over a bounded sample of the sparse query (the
    search stops after `limit` embeddings, sized by a probe to take about
    seconds_target on all cores): one warm-up, then the median of 5 runs, on
    all cores and on one thread (a limit 1/16 as large).  The dense query is
    also counted in full once, as a check of the device count."""
    import numpy as np
    import oracle as O
    off, adj, lab = g.arrays()
    threads = max(1, min(16, os.cpu_count() or 1))
    q = dict(queries)["sparse"]

    def run(limit, th):
        t = time.perf_counter()
        cnt, _, stopped = O.match_partial(q.labels, q.adjm(), off, adj, lab, threads=th, limit=limit)
        return cnt, time.perf_counter() - t, stopped

    cnt, dt, _ = run(20_000_000, threads)
    limit = int(max(1e6, min(cnt / max(dt, 1e-6) * seconds_target, 5e10)))
    rates = []
    for i in range(6):
        cnt, dt, stopped = run(limit, threads)
        if i:
            rates.append(cnt / dt)
    limit1 = max(1_000_000, limit // 16)
    rates1 = []
    for i in range(6):
        cnt1, dt1, _ = run(limit1, 1)
        if i:
            rates1.append(cnt1 / dt1)
    checks = {}
    qd = dict(queries)["dense"]
    dense, _ = O.match(qd.labels, qd.adjm(), off, adj, lab, threads=threads)
    checks["dense"] = dense
    return {"value": round(float(np.median(rates)), 1), "unit": "embeddings/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "single_thread_value": round(float(np.median(rates1)), 1),
            "runs": [round(r, 1) for r in rates], "single_thread_runs": [round(r, 1) for r in rates1],
            "sample": f"oracle backtracking (oracle/oracle.c) on the C2 sparse query, the search stopped after "
                      f"{limit} embeddings ({limit1} on one thread); roots handed out dynamically to "
                      f"{threads} threads; one warm-up, then the median of 5 runs"}, checks


def run_c2(args, with_cpu=True):
    import numpy as np
    import fsig_amd as F
    from fsig_amd import workloads as W
    t0 = time.perf_counter()
    g = W.c2_graph()
    queries = W.c2_queries(g)
    log(f"C2 graph {g.n} V {g.nnz // 2} E {g.num_labels} labels in {time.perf_counter() - t0:.2f}s")
    ctx = F.Context(0, 64 << 30)
    dg = ctx.upload(g, pcsr=True)
    counts = {name: F.match(dg, q, args.alg, F.MODE_COUNT).count for name, q in queries}
    log("C2 counts", counts)

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
    roofline = {"bound": "hbm", "kernel": kname, "workload": "C2", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(kname), "avg_launch_ms": round(d["ms"] / d["launches"], 4),
                "alg_bytes_per_launch": int(d["alg_bytes"] / d["launches"])}
    gsi = {}
    if args.alg == 1:  # alg 2 (GSI path) on the same queries, outside the timed region
        t = time.perf_counter()
        n2 = 0
        for name, q in queries:
            c2 = F.match(dg, q, F.ALG_GSI, F.MODE_COUNT).count
            assert c2 == counts[name], "alg 2 disagrees with alg 1"
            n2 += c2
        gsi = {"alg2_embeddings_per_s": round(n2 / (time.perf_counter() - t), 1)}
    stream_gbps = ctx.stream_copy(1 << 30, 20)
    lat_ms = np.array(lat) * 1e3
    out = {"embeddings_per_s": round(emb / elapsed, 1), "ms_per_step": round(elapsed / args.steps * 1e3, 3),
           "steps": args.steps, "queries_per_s": round(len(lat) / elapsed, 2), "p50_ms": pct(lat_ms, 50),
           "p99_ms": pct(lat_ms, 99), "workload": "C2 dblp-shaped: Chung-Lu gamma 2.3, 317,080 V, 1,049,866 E, "
                                                 "16 Zipf(1.2) labels; 8-vertex dense + sparse random-walk queries; "
                                                 "alg 1, count mode (exact)",
           "queries": {n: {"edges": q.num_edges, "seed": W.C2_QUERIES[n]["seed"], "embeddings": counts[n]}
                       for n, q in queries},
           "kernels": kernels, "stream_copy_GB_per_s": round(stream_gbps, 1), **gsi}
    cb = None
    if with_cpu:
        log("C2 cpu baseline ...")
        cb, checks = cpu_baseline(g, queries)
        for n, c in checks.items():
            assert c == counts[n], f"oracle count {c} != engine count {counts[n]} for {n}"
        out["cpu_checked_counts"] = checks
    dg.free()
    ctx.close()
    return out, roofline, cb


# ---------------------------------------------------------------------------
# C4: one bounded batch against the plain loop
def run_c4(args, steps):
    import numpy as np
    import fsig_amd as F
    from fsig_amd import workloads as W
    p = W.C4_QUERIES
    t0 = time.perf_counter()
    g = W.c4_graph()
    ctx = F.Context(0, 96 << 30)
    dg = ctx.upload(g)
    info = {}
    qs = W.c4_queries(g, info)
    log(f"C4 graph {g.n} V {g.nnz // 2} E, {len(qs)} queries ({info['no_walk']} draws without a walk) in "
        f"{time.perf_counter() - t0:.1f}s")
    B, WL = p["bound"], p["work_limit"]
    ctx.set_work_limit(WL)

    def plain_pass():
        res, lat = [], []
        for q in qs:
            t = time.perf_counter()
            try:
                res.append(F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=B).count)
            except F.FsigError as e:
                if e.code != F.FSIG_ELIMIT:
                    raise
                res.append(None)
            lat.append(time.perf_counter() - t)
        return res, lat

    plain_pass()  # warm-up
    ctx.synchronize()
    plain, lat = plain_pass()
    adm = [i for i, c in enumerate(plain) if c is not None]
    plain_s = sum(lat[i] for i in adm)
    aq = [qs[i] for i in adm]
    want = [plain[i] for i in adm]
    log(f"C4 plain: {len(adm)} admitted in {plain_s * 1e3:.0f} ms, {len(qs) - len(adm)} refused")
    counts, _, st = F.match_batch(dg, aq, bounds=B, status=True)  # warm-up, and the check
    assert not st.any(), "a query the plain engine admitted was refused in the batch"
    assert counts.tolist() == want, "batched counts differ from the plain loop"
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        counts, _ = F.match_batch(dg, aq, bounds=B)
    ctx.synchronize()
    batch_s = (time.perf_counter() - t) / steps
    assert counts.tolist() == want
    ctx.set_work_limit(0)
    out = {"workload": "C4 batched: Chung-Lu 1M V / 10M E, 32 Zipf(1.2) labels; the first 256 random-walk "
                       "queries, k uniform in 4..16, dense and sparse alternating, each counted to a bound of 1e8; "
                       "one fsig_match_batch_bounded call per step",
           "batched_queries_per_s": round(len(aq) / batch_s, 2), "plain_queries_per_s": round(len(aq) / plain_s, 2),
           "batched_over_plain": round(plain_s / batch_s, 3), "ms_per_batch": round(batch_s * 1e3, 3),
           "steps": steps, "queries": len(qs), "admitted": len(adm), "refused_by_work_limit": len(qs) - len(adm),
           "exact_below_bound": int(sum(1 for c in want if c < B)), "at_bound": int(sum(1 for c in want if c == B)),
           "bound": B, "work_limit": WL, "draws_without_walk": info["no_walk"],
           "embeddings_per_step": int(sum(want)), "plain_p50_ms": pct(np.array(lat) * 1e3, 50),
           "plain_p99_ms": pct(np.array(lat) * 1e3, 99)}
    dg.free()
    ctx.close()
    return out


# ---------------------------------------------------------------------------
# C5: the served stream (headline)
def run_c5(args, steps, warmup):
    import threading
    import numpy as np
    import fsig_amd as F
    from fsig_amd import workloads as W
    p = W.C5_STREAM
    t0 = time.perf_counter()
    g = W.c5_graph()
    log(f"C5 graph {g.n} V {g.nnz // 2} E in {time.perf_counter() - t0:.1f}s")
    ctx = F.Context(0, 32 << 30)
    dg = ctx.upload(g)
    info = {}
    stream = W.c5_stream(g, args.requests, args.qps, info=info)
    arrivals = np.array([a for a, _ in stream])
    qs = [q for _, q in stream]
    B, WL = p["bound"], p["work_limit"]
    log(f"C5 {len(qs)} requests ({info['no_walk']} draws without a walk), offered "
        f"{len(qs) / max(arrivals[-1], 1e-9):.0f} q/s")
    # one-at-a-time plain launches in arrival order (after a warm-up pass)
    ctx.set_work_limit(WL)

    def plain_one(q):
        try:
            return F.match(dg, q, F.ALG_MMATCH, F.MODE_COUNT, max_embeddings=B).count
        except F.FsigError as e:
            if e.code != F.FSIG_ELIMIT:
                raise
            return None

    for q in qs:
        plain_one(q)
    ctx.synchronize()
    plain, plain_lat = [], []
    start = time.perf_counter()
    for a, q in zip(arrivals, qs):
        now = time.perf_counter() - start
        if now < a:
            time.sleep(a - now)
        plain.append(plain_one(q))
        plain_lat.append((time.perf_counter() - start - a) * 1e3)
    plain_s = time.perf_counter() - start
    ctx.set_work_limit(0)
    log(f"C5 plain: {plain_s * 1e3:.0f} ms, p50 {np.percentile(plain_lat, 50):.1f} ms")
    srv = F.Server(dg, workers=args.workers, hbm_budget=128 << 30, max_batch=args.max_batch)
    srv.set_work_limit(WL)

    def serve_once():
        tickets = [None] * len(qs)

        def feed(t_start):
            for i, (a, q) in enumerate(zip(arrivals, qs)):
                now = time.perf_counter() - t_start
                if now < a:
                    time.sleep(a - now)
                tickets[i] = srv.submit(q, B)

        t_start = time.perf_counter()
        th = threading.Thread(target=feed, args=(t_start,))
        th.start()
        th.join()
        res = []
        for t in tickets:
            try:
                res.append(srv.wait(t))
            except F.FsigError as e:
                if e.code != F.FSIG_ELIMIT:
                    raise
                res.append((None, None))
        return res, time.perf_counter() - t_start

    for _ in range(warmup):
        serve_once()
    st0 = srv.stats()
    served_s, lat_all, served = 0.0, [], None
    for _ in range(steps):
        res, dt = serve_once()
        served_s += dt
        lat_all += [l for c, l in res if c is not None]
        got = [c for c, _ in res]
        for a, b in zip(got, plain):  # requests both engines answered agree
            assert a is None or b is None or a == b, "served count differs from the plain launch"
        served = got
    st = srv.stats()
    srv.close()
    nreq = len(qs) * steps
    out = {"workload": f"C5 served: R-MAT 2^24 V / {g.nnz // 2} E, 64 uniform labels, resident; "
                       f"{len(qs)} Poisson requests offered at {len(qs) / max(arrivals[-1], 1e-9):.0f} q/s, "
                       "k = round(exp(U(ln 4, ln 32))), every drawn query kept, each counted to a bound of 1e8; "
                       "one stream served per step",
           "served_queries_per_s": round(nreq / served_s, 2), "ms_per_step": round(served_s / steps * 1e3, 3),
           "served_p50_ms": pct(lat_all, 50), "served_p99_ms": pct(lat_all, 99),
           "plain_queries_per_s": round(len(qs) / plain_s, 2), "plain_p50_ms": pct(plain_lat, 50),
           "plain_p99_ms": pct(plain_lat, 99), "served_over_plain": round((plain_s * steps) / served_s, 3),
           "requests": len(qs), "refused_by_work_limit": {"plain": sum(1 for c in plain if c is None),
                                                         "served": sum(1 for c in served if c is None)},
           "exact_below_bound": sum(1 for c in plain if c is not None and c < B),
           "at_bound": sum(1 for c in plain if c == B), "bound": B, "work_limit": WL,
           "draws_without_walk": info["no_walk"],
           "k_histogram": {str(k): int(sum(1 for q in qs if q.k == k)) for k in sorted({q.k for q in qs})},
           "workers": args.workers, "hbm_budget_GiB": 128, "max_batch": args.max_batch,
           "server": {k: (round(v - st0[k], 1) if isinstance(v, float) else v - st0[k])
                      for k, v in st.items() if k not in ("max_concurrent", "max_batch_seen", "wall_ms")}}
    dg.free()
    ctx.close()
    return out


def line(value, unit, steps, warmup, ms_per_step, config, **rest):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic", "config": config, **rest}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all", choices=["all", "c2", "c3", "c4", "c5", "ingest"])
    ap.add_argument("--qps", type=float, default=0.0, help="c5: arrival rate (0: the workload's)")
    ap.add_argument("--requests", type=int, default=0, help="c5: stream length (0: the workload's)")
    ap.add_argument("--workers", type=int, default=6, help="c5: server worker contexts")
    ap.add_argument("--max-batch", type=int, default=32, help="c5: most requests a worker batches")
    ap.add_argument("--c4-steps", type=int, default=5, help="all: timed batches of C4")
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
    import fsig_amd as F
    if args.copy_only:
        ctx = F.Context(0, 2 * args.copy_only + (64 << 20))
        gbps = ctx.stream_copy(args.copy_only, 5)
        print(json.dumps({"stream_copy_bytes": args.copy_only, "iters": 6, "GB_per_s": round(gbps, 1)}))
        return
    if args.workload == "c3":
        return bench_c3(args)
    parallel = "1 GPU"
    if args.workload == "c2":
        c2, roof, cb = run_c2(args, with_cpu=not args.no_cpu)
        out = line(c2["embeddings_per_s"], "embeddings/s", args.steps, args.warmup, c2["ms_per_step"],
                   {"workload": c2["workload"], "parallelism": parallel}, roofline=roof, c2=c2)
        if cb:
            out["cpu_baseline"] = cb
        print(json.dumps(out), flush=True)
        return
    if args.workload == "c4":
        c4 = run_c4(args, args.steps)
        print(json.dumps(line(c4["batched_queries_per_s"], "queries/s", args.steps, 1, c4["ms_per_batch"],
                              {"workload": c4["workload"], "parallelism": parallel}, c4=c4)), flush=True)
        return
    if args.workload == "c5":
        c5 = run_c5(args, args.steps, args.warmup)
        print(json.dumps(line(c5["served_queries_per_s"], "queries/s", args.steps, args.warmup, c5["ms_per_step"],
                              {"workload": c5["workload"], "parallelism": parallel}, c5=c5)), flush=True)
        return
    # all: C2 (roofline, CPU baseline), C4 (batched vs plain), C5 (headline)
    c2, roof, cb = run_c2(args, with_cpu=not args.no_cpu)
    c4 = run_c4(args, max(1, min(args.c4_steps, args.steps)))
    c5 = run_c5(args, args.steps, args.warmup)
    out = line(c5["served_queries_per_s"], "queries/s", args.steps, args.warmup, c5["ms_per_step"],
               {"workload": c5["workload"], "parallelism": parallel},
               served_p50_ms=c5["served_p50_ms"], served_p99_ms=c5["served_p99_ms"],
               served_over_plain=c5["served_over_plain"], batched_over_plain=c4["batched_over_plain"],
               embeddings_per_s=c2["embeddings_per_s"], hbm_frac=roof["frac"], roofline=roof,
               c5=c5, c4=c4, c2=c2)
    if cb:
        out["cpu_baseline"] = cb
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
    Full u64 counts of the triangle, the 4-clique and the 4-cycle (one
    embedding per orbit of the pattern's automorphisms, times their number),
    and for all four patterns the bounded enumeration of the first 1e8
    embeddings (rows come back to the host; a sample is checked there: every
    query edge present, vertices distinct).  The house count grows about
    4.5x per scale step (28 s at scale 16, 127 s at 17) and is counted on
    the scale-16 analogue (same generator and density) instead."""
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
    for name in ("triangle", "clique4", "cycle4"):
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
    for sc, m in ((16, 35_000_000 // 64),):
        gs = W.c3_graph(sc, m)
        ds = ctx.upload(gs)
        for name in ("house",):
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
           "config": {"workload": f"C3 unlabeled R-MAT (0.57,0.19,0.19) 2^22 V / {g.nnz // 2} E; triangle, "
                                  "4-clique and 4-cycle counted (value), first 1e8 embeddings of all four patterns "
                                  "enumerated",
                      "parallelism": "1 GPU"},
           **out, "smaller_scale_counts": small}
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
