#!/usr/bin/env python3
"""Benchmark: temporal-motif queries on one MI355X.

Workload (default): 256 random motif queries (3-6 edges on <= 4 nodes, delta
log-uniform in [600, 604800] s, seed 4) on the resident 64M-edge power-law
graph (2.6M nodes, 7.5 years, seed 3).  Every query of the set runs in every
pass.

  plain   one query per launch, sequentially (tm_count), host wall latency each
  batch   the whole set in one tm_batch_count call per step; value = queries/s
          of the batched step (every count checked bit-exact against plain)
  serve   Poisson arrivals at several rates (80% count, 20% enumerate with a
          1M-row cap) through the server (continuous dispatch over two
          launch slots, one request per launch, oldest first) and through a
          plain FIFO loop (one query per launch, in arrival order, one at a
          time), both run for real on the same stream

roofline: the single-query counting kernel instantiation with the largest
total time in the plain pass; achieved = the bytes its own search reads per
launch (counted by the metered kernel, tm_count_bytes: gap words, edge
records, candidate words, every search probe) / its mean launch time (HIP
events on the library stream).  traffic = HBM bytes per launch from the
committed counter summary, when it covers this run's launches.

cpu_baseline: the C oracle (a multithreaded restatement of the reference's
search) on seeded root windows of every query, extrapolated per query, median
of 3 repetitions.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("#", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--graph", default="pl64m")
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline budget per repetition (0 = skip)")
    ap.add_argument("--json-out", default="")
    ap.add_argument("--plain-only", action="store_true", help="stop after the plain pass (counter runs)")
    ap.add_argument("--no-tuning", action="store_true", help="ignore the autotuned table (library defaults)")
    ap.add_argument("--serve-rates", default="50,100,200,500", help="Poisson arrival rates (requests/s)")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests per rate (0 = no serving run)")
    ap.add_argument("--serve-enum-frac", type=float, default=0.2, help="share of enumerate requests")
    ap.add_argument("--serve-enum-cap", type=int, default=1000000, help="row cap of enumerate requests")
    ap.add_argument("--budget-s", type=float, default=520.0,
                    help="wall-clock budget of the whole run: serving and the CPU baseline get what the plain "
                         "pass, the metered pass and the K+W batch steps leave")
    ap.add_argument("--serve-only", action="store_true",
                    help="plain pass, then the serving sweep with --serve-requests per rate, no time budget")
    ap.add_argument("--reduce-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="no GPU: reduce two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of --plain-only "
                         "--calibrate-copy into the per-launch HBM bytes of the roofline kernel (--json-out)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also copy this many bytes device-to-device (HBM counter scale for profiles)")
    ap.add_argument("--constraints", action="store_true",
                    help="instead: the constraint-feature ladder of the reference's additional-features "
                         "queries (no / v / t / all / f) through the frontend on --graph")
    args = ap.parse_args()
    if args.reduce_pmc:
        reduce_pmc(args.reduce_pmc[0], args.reduce_pmc[1], args.json_out)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X "
              "(--gpus %d requested; the multi-GPU path is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.constraints:
        constraint_ladder(args)
        return

    import numpy as np
    from everest_amd import abi, core
    from everest_amd.gen import graphs, queries

    if abi.lib().tm_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)

    t_start = time.time()
    t0 = time.time()
    U, V, T = graphs.make(args.graph)
    el = core.data.EdgeListLoader.fromArrays(args.graph, U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    log("graph %s: E=%d V=%d prepared in %.1fs" % (args.graph, hg.numEdges, hg.numVertices, time.time() - t0))
    t0 = time.time()
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    log("resident on device: %.2f GB, upload+index %.1fs" % (dg.info()["device_bytes"] / 1e9, time.time() - t0))

    qs = queries.batch_queries(args.queries, seed=args.seed)
    motifs = [(queries.as_motif(e), d) for e, d in qs]
    nq = len(motifs)

    # launch parameters per shape class from the autotuner's table for this
    # graph (python -m everest_amd.autotune), used by every pass below
    tuned = os.path.join(ROOT, "everest_amd", "tuned", "%s.json" % args.graph)
    tuning_src = "library defaults"
    if os.path.exists(tuned) and not args.no_tuning:
        from everest_amd import autotune
        table, meta = autotune.load(tuned)
        autotune.apply(dg, table)
        tuning_src = "autotuned table everest_amd/tuned/%s.json (%d classes)" % (args.graph, len(table))
    log("launch parameters: %s" % tuning_src)

    # ---- plain: one query per launch
    plain_ms, counts, kernel_ms, variant = [], [], [], []
    t = time.perf_counter()
    for i, (m, d) in enumerate(motifs):
        t1 = time.perf_counter()
        c = dg.count(m, d)
        plain_ms.append((time.perf_counter() - t1) * 1e3)
        kernel_ms.append(dg.last_kernel_ms())
        counts.append(c)
        st = dg.last_stats()
        nmap = 4 if max(max(a, b) for a, b, _ in m) + 1 <= 4 else 8
        # queries counted by the window-sweep kernel are not mining-kernel launches
        variant.append(("sweep",) if st["sweep"] else (nmap, st["wpe"]))
        if (i + 1) % 32 == 0:
            log("plain: %d/%d queries, %.1fs" % (i + 1, nq, time.perf_counter() - t))
    plain_total_s = sum(plain_ms) / 1e3
    plain_qps = nq / plain_total_s
    slow = sorted(range(nq), key=lambda i: -plain_ms[i])[:5]
    log("plain: %d queries in %.3fs (%.1f q/s); slowest %s" % (
        nq, plain_total_s, plain_qps, ", ".join("%d:%.0fms" % (i, plain_ms[i]) for i in slow)))

    if args.plain_only:
        log("plain pass only (--plain-only)")
        calibrate_copy(args.calibrate_copy)
        return
    if args.serve_only:
        rates = [float(x) for x in args.serve_rates.split(",") if x.strip()]
        sweep = [serve(dg, motifs, counts, rate, args, args.serve_requests) for rate in rates]
        line = json.dumps({"serving_sweep": sweep, "plain_pass_qps": round(plain_qps, 3),
                           "plain_mean_service_ms": round(plain_total_s * 1e3 / nq, 3),
                           "workload": "%d queries seed %d on %s" % (nq, args.seed, args.graph)})
        print(line, flush=True)
        if args.json_out:
            with open(args.json_out, "w") as f:
                f.write(line + "\n")
        return

    # ---- roofline of the dominant single-query counting kernel (largest total
    # time): the search's own bytes per launch (metered kernel, untimed) over
    # the mean launch duration of the timed kernel
    groups = {}
    for i in range(nq):
        groups.setdefault(variant[i], []).append(i)
    dom = max((k for k in groups if k[0] != "sweep"), key=lambda k: sum(kernel_ms[i] for i in groups[k]))
    members = groups[dom]
    t = time.perf_counter()
    own = {}
    for i in members:
        c, b = dg.count_bytes(motifs[i][0], motifs[i][1])
        assert c == counts[i], "metered kernel count differs"
        own[i] = b
    log("metered kernel: %d queries in %.1fs" % (len(members), time.perf_counter() - t))
    dom_n = len(members)
    dom_bytes = sum(own[i] for i in members) / dom_n
    dom_ms = sum(kernel_ms[i] for i in members) / dom_n
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    # the instantiation's name as the profiler prints it
    kname = "void tmk::mine_kernel<8, %d, false, 0, tmk::DynSpec, %d, false>(tm_kernel_args)" % dom
    log("dominant kernel %s: %d launches, %.4f GB and %.3f ms per launch -> %.1f GB/s" % (
        kname, dom_n, dom_bytes / 1e9, dom_ms, achieved))
    # HBM bytes per launch from the committed counter runs of the same plain
    # pass (bench.py --reduce-pmc), matched launch by launch in query order
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", "mine_kernel_pl64m_hbm.json")
    if os.path.exists(tf) and args.graph == "pl64m" and args.queries == 256 and args.seed == 4:
        with open(tf) as f:
            tj = json.load(f)
        per = tj.get("hbm_bytes_per_launch", [])
        if tj.get("kernel") == kname and len(per) == dom_n:
            traffic = round(sum(per) / dom_n, 1)
            traffic_src = "profiles/mine_kernel_pl64m_hbm.md (FETCH_SIZE x %.3f + WRITE_SIZE x %.3f)" % (
                tj["fetch_scale"], tj["write_scale"])

    # ---- batched: the whole set in one call per step
    for _ in range(args.warmup):
        got = dg.batch_count(motifs)
        assert got == counts, "batched counts differ from single-query counts"
    step_ms = []
    for s in range(args.steps):
        t = time.perf_counter()
        got = dg.batch_count(motifs)
        step_ms.append((time.perf_counter() - t) * 1e3)
        assert got == counts, "batched counts differ from single-query counts"
        log("step %d: %.2f ms" % (s, step_ms[-1]))
    ms_per_step = float(np.mean(step_ms))
    batch_qps = nq / (ms_per_step / 1e3)
    matches = float(sum(counts))

    # ---- serving sweep and CPU baseline in what the budget leaves: each rate
    # gets an equal share of 70% of it, sized by the plain pass's mean
    # service time (the plain loop and the server each serve the stream once)
    left = args.budget_s - (time.time() - t_start)
    log("%.0fs of the %.0fs budget left for serving and the CPU baseline" % (left, args.budget_s))
    serving = None
    rates = [float(x) for x in args.serve_rates.split(",") if x.strip()]
    if args.serve_requests > 0 and rates and left > 30:
        mean_s = plain_total_s / nq
        per_rate = 0.7 * left / len(rates)
        n_req = int(max(20, min(args.serve_requests, per_rate / (2.5 * mean_s))))
        serving = [serve(dg, motifs, counts, rate, args, n_req) for rate in rates]
    left = args.budget_s - (time.time() - t_start)
    cpu = None
    if args.cpu_seconds > 0 and left > 10:
        cpu = cpu_baseline(hg, arrs, motifs, dg, min(args.cpu_seconds, left / 4))

    res = {
        "metric": "queries/s + p99 latency (batched/served vs plain) on 1 GPU; bit-exact counts",
        "value": round(batch_qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32/u64",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": "config4-batched: all %d queries, M=3..6 on <=4 nodes, delta log-uniform [600,604800]s, "
                        "seed %d; graph %s E=%d V=%d" % (nq, args.seed, args.graph, hg.numEdges, hg.numVertices),
            "plain_qps": round(plain_qps, 3),
            "batched_over_plain": round(batch_qps / plain_qps, 3),
            "plain_latency_ms_p50": round(float(np.percentile(plain_ms, 50)), 3),
            "plain_latency_ms_p99": round(float(np.percentile(plain_ms, 99)), 3),
            "plain_slowest_ms": {str(i): round(plain_ms[i], 1) for i in slow},
            "batch_step_ms_p50": round(float(np.percentile(step_ms, 50)), 3),
            "batch_step_ms_p99": round(float(np.percentile(step_ms, 99)), 3),
            "matches_per_s": round(matches / (ms_per_step / 1e3), 1),
            "bit_exact_batch_vs_plain": True,
            "launch_parameters": tuning_src,
            "serving": serving,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": kname,
            "launches": dom_n,
            "algorithmic_bytes_per_launch": round(dom_bytes, 1),
            "algorithmic_bytes_model": "the search's own reads (metered kernel): 16 B per lane and gap rule per "
                                       "4-root block, 12 B per root edge, 4 B per candidate word and per search probe",
            "mean_launch_ms": round(dom_ms, 4),
        },
        "cpu_baseline": cpu,
    }
    line = json.dumps(res)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")


def serve(dg, motifs, counts, rate, args, n):
    """One serving run at `rate` requests/s: requests drawn from the query set
    (seeded), a share of them enumerate requests, arriving as a Poisson
    stream.  Served: a client thread submits to the server at the arrival
    times.  Plain: the same stream through a FIFO loop, one query per launch
    (tm_count / tm_enum), in arrival order.  Latency = arrival -> result on
    the host; queries/s = requests / (last completion - first arrival).
    Served counts and rows equal the plain loop's (tm_enum's rows at a cap
    are deterministic)."""
    import numpy as np
    from everest_amd import abi

    rng = np.random.default_rng(args.seed * 1000 + int(rate))
    pick = [int(x) for x in rng.integers(0, len(motifs), n)]
    enum = rng.random(n) < args.serve_enum_frac
    gaps = rng.exponential(1.0 / rate, n)
    arrive = np.cumsum(gaps) - gaps[0]
    cap = args.serve_enum_cap
    log("serving at %.0f/s: %d requests (%d enumerate, cap %d)" % (rate, n, int(enum.sum()), cap))

    # plain FIFO loop on the same stream
    plain_lat, plain_rows = np.zeros(n), {}
    t0 = time.perf_counter()
    for k in range(n):
        dt = arrive[k] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        m, d = motifs[pick[k]]
        if enum[k]:
            c, rows = dg.enum(m, d, cap=cap)
            plain_rows[k] = (c, rows)
        else:
            c = dg.count(m, d)
            assert c == counts[pick[k]]
        plain_lat[k] = (time.perf_counter() - t0 - arrive[k]) * 1e3
        if (k + 1) % 500 == 0:
            log("  plain loop: %d/%d requests, %.1fs" % (k + 1, n, time.perf_counter() - t0))
    plain_span = time.perf_counter() - t0

    # served: continuous dispatch over the server's launch slots
    srv = abi.Server(dg)
    tickets = [None] * n
    t0 = time.perf_counter()

    def client():
        for k in range(n):
            dt = arrive[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            m, d = motifs[pick[k]]
            tickets[k] = srv.submit(m, d, cap if enum[k] else -1)

    th = threading.Thread(target=client)
    th.start()
    lat, batch = np.zeros(n), np.zeros(n)
    last_done = 0.0
    for k in range(n):
        while tickets[k] is None:
            time.sleep(0.0002)
        r = srv.wait(tickets[k])
        assert r["status"] == 0, r
        lat[k], batch[k] = r["latency_ms"], r["batch"]
        last_done = max(last_done, arrive[k] + r["latency_ms"] / 1e3)
        m = motifs[pick[k]][0]
        if enum[k]:
            c, rows_want = plain_rows[k]
            rows = srv.take_rows(tickets[k], len(m), cap)
            assert r["count"] == c, "served enumerate count differs"
            assert rows.shape == rows_want.shape and np.array_equal(canon(rows), canon(rows_want)), \
                "served rows differ"
        else:
            assert r["count"] == counts[pick[k]], "served count differs"
            srv.release(tickets[k])
        if (k + 1) % 500 == 0:
            log("  served: %d/%d requests, %.1fs" % (k + 1, n, time.perf_counter() - t0))
    th.join()
    st = srv.stats()
    srv.close()
    res = {
        "offered_qps": rate, "requests": n, "enumerate_requests": int(enum.sum()), "enum_cap": cap,
        "served_qps": round(n / last_done, 3),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "plain_qps": round(n / plain_span, 3),
        "plain_latency_ms_p50": round(float(np.percentile(plain_lat, 50)), 3),
        "plain_latency_ms_p99": round(float(np.percentile(plain_lat, 99)), 3),
        "served_over_plain_qps": round((n / last_done) / (n / plain_span), 3),
        "launches": st["launches"], "largest_batch": st["max_batch"], "mean_batch": round(float(batch.mean()), 2),
        "bit_exact_served_vs_plain": True,
    }
    log("  served %.1f q/s p50 %.1f p99 %.1f ms | plain %.1f q/s p50 %.1f p99 %.1f ms | %d launches" % (
        res["served_qps"], res["latency_ms_p50"], res["latency_ms_p99"], res["plain_qps"],
        res["plain_latency_ms_p50"], res["plain_latency_ms_p99"], res["launches"]))
    return res


# The reference's additional-features ladder (system/queries/additional-features/WI-M4-{no,v,t,all,f}
# .yaml): the 4-cycle, then node-label predicates, temporal gaps, an anti-edge, a filter.  The
# filter names the matched non-anti edges (the reference's names the anti edge's slot, e2).
NODES4 = ["node0: f > 500000000", "node1: f < 500000000", "node2: f > 500000000", "node3: f < 500000000"]
LADDER = [
    ("no", "M4.txt", []),
    ("v", "M4.txt", NODES4),
    ("t", "M4.txt", NODES4 + ["temporal: [21600, 43200, 21600]"]),
    ("all", "M4-A.txt", NODES4 + ["antiEdge: [False, False, True, False, False]",
                                  "temporal: [21600, 43200, 43200, 21600]"]),
    ("f", "M4-A.txt", NODES4 + ["antiEdge: [False, False, True, False, False]",
                                "temporal: [21600, 43200, 43200, 21600]",
                                'filter: "(e0 + e1 + e3 + e4) > 1000000000"']),
]


def constraint_ladder(args):
    """Each query of LADDER at delta 86400 as the reference's frontend runs
    it (system/tmine/frontend.py:52-96: YAML -> KernelGen -> plugin ->
    GPUWorkerLibManager -> SingleGPUExecution; the constrained kernel for
    v / t / all / f), on the synthetic graph loaded from arrays with its
    seeded labels; the ladder runs twice and the second pass is reported (the
    first uploads the graph)."""
    import tempfile
    from everest_amd import core
    from everest_amd.gen import graphs
    from everest_amd.tmine import KernelGen, Query, utils
    root = tempfile.mkdtemp(prefix="everest_cons_")
    t0 = time.time()
    U, V, T = graphs.make(args.graph)
    el = core.data.EdgeListLoader.fromArrays(args.graph, U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    log("graph %s: E=%d prepared in %.1fs" % (args.graph, hg.numEdges, time.time() - t0))
    dm = core.data.MakeGraphDataManager(1)
    lm = core.GPUWorkerLibManager()
    lib = os.path.join(root, "libDYN.so")
    res = []
    for rep in range(2):
        for name, pat, lines in LADDER:
            q = os.path.join(root, "%s.yaml" % name)
            with open(q, "w") as f:
                f.write("pattern: %s\ngraph: %s.txt\nconstraints:\n  delta: 86400\n%s\noptions:\n  gpu: 1\n" % (
                    pat, args.graph, "".join("  %s\n" % x for x in lines)))
            query = Query.Query(q)
            src = os.path.join(root, "%s.hip" % name)
            with open(src, "w") as f:
                f.write(KernelGen.KernelGen(query.pattern(), query).generate())
            utils.compilePlugin(src, lib)
            lm.openLib(lib)
            w = lm.construct("DYN", 0)
            ex = core.SingleGPUExecution(dm)
            ex.setHostGraph(hg)
            ex.setHostMotif(core.data.HostMotifData(query.pattern()))
            ex.setWorker(w)
            t = time.perf_counter()
            ex.run(86400)
            ms = (time.perf_counter() - t) * 1e3
            cn = w.count()
            del ex, w
            log("pass %d %-3s %s: %d matches, %.1f ms" % (rep, name, pat, cn, ms))
            if rep == 1:
                res.append(dict(query=name, pattern=pat, constraints=lines, delta=86400, count=int(cn), ms=round(ms, 2)))
    line = json.dumps({"constraint_ladder": res, "graph": args.graph,
                       "source": "system/queries/additional-features/WI-M4-*.yaml shapes"})
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")


def canon(rows):
    """Rows in lexicographic order (row sets compare as arrays)."""
    import numpy as np
    if len(rows) == 0:
        return rows
    return rows[np.lexsort(rows.T[::-1])]


def calibrate_copy(nbytes):
    """One device-to-device copy of nbytes: under a counter run its
    FETCH_SIZE / WRITE_SIZE give the scale of those counters."""
    if nbytes <= 0:
        return
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log("calibration copy: %d bytes read, %d written" % (a.numel() * 4, b.numel() * 4))


def reduce_pmc(fetch_dir, write_dir, out):
    """Per-launch HBM bytes of the roofline kernel from two counter runs of
    the plain pass.  The roofline kernel is the single-query counting
    variant (`mine_kernel<..., false, 0, ..., false>`) with the largest
    total duration in the FETCH_SIZE run, as bench.py picks it by kernel
    time.  Each counter is scaled by what it reads for the 1 GiB
    device-to-device calibration copy (the largest copyBuffer dispatch).
    Launches with a one-wavefront grid (queries left to the window sweep:
    the mining launch has no roots) are dropped; the rest are the plain
    pass's launches of that kernel in query order."""
    import csv
    import glob
    import re

    def load(d, counter):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if len(files) != 1:
            raise SystemExit("expected one counter_collection.csv under %s, found %d" % (d, len(files)))
        with open(files[0]) as f:
            r = [x for x in csv.DictReader(f) if x["Counter_Name"] == counter]
        r.sort(key=lambda x: int(x["Dispatch_Id"]))
        return r

    fr, wr = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    counting = re.compile(r"^void tmk::mine_kernel<8, \d, false, 0, tmk::DynSpec, \d, false>\(tm_kernel_args\)$")
    total = {}
    for x in fr:
        if counting.match(x["Kernel_Name"]) and int(x["Grid_Size"]) > 64:
            total[x["Kernel_Name"]] = total.get(x["Kernel_Name"], 0) + int(x["End_Timestamp"]) - int(x["Start_Timestamp"])
    if not total:
        raise SystemExit("no single-query counting launches in %s" % fetch_dir)
    kname = max(total, key=total.get)

    def per_launch(r):
        cal = max(float(x["Counter_Value"]) for x in r if "copyBuffer" in x["Kernel_Name"])
        return cal, [float(x["Counter_Value"]) for x in r if x["Kernel_Name"] == kname and int(x["Grid_Size"]) > 64]

    fcal, fk = per_launch(fr)
    wcal, wk = per_launch(wr)
    if len(fk) != len(wk):
        raise SystemExit("the two runs counted %d and %d launches" % (len(fk), len(wk)))
    gib = float(1 << 30)
    fscale, wscale = gib / (fcal * 1024.0), gib / (wcal * 1024.0)
    per = [round(f * 1024.0 * fscale + w * 1024.0 * wscale, 1) for f, w in zip(fk, wk)]
    res = {"kernel": kname, "workload": "pl64m plain pass, 256 queries seed 4 (launches of this kernel)",
           "fetch_scale": fscale, "write_scale": wscale, "launches": len(per), "hbm_bytes_per_launch": per,
           "mean_hbm_bytes": round(sum(per) / max(1, len(per)), 1),
           "fetch_bytes": round(sum(fk) * 1024.0 * fscale, 1), "write_bytes": round(sum(wk) * 1024.0 * wscale, 1)}
    line = json.dumps(res)
    if out:
        with open(out, "w") as f:
            f.write(line + "\n")
    print(json.dumps({k: v for k, v in res.items() if k != "hbm_bytes_per_launch"}))


def cpu_baseline(hg, arrs, motifs, dg, budget_s):
    """The C oracle (the reference's search restated, multithreaded) on four
    seeded root windows of every query; per query, the time per root over
    its windows times its root count, summed over the set -> queries/s; the
    median of 3 repetitions (the same windows), with the spread of the
    per-window estimates.  Window sizes are set by a sizing pass so that one
    repetition takes about budget_s of wall time.  Every window's count is
    checked against the GPU on the same roots."""
    import numpy as np
    from oracle import oracle

    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    # the host cores this process may use (16 per GPU on the measurement hosts)
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    E = hg.numEdges
    nwin, cal = 4, 4
    target = budget_s / (len(motifs) * nwin)
    rng = np.random.default_rng(17)
    # sizing pass: each window grows from a seeded start in chunks of 4, 4, 8,
    # 16, ... roots until it has taken `target` seconds (or reaches 8192
    # roots); per-root cost is heavy-tailed (hub roots), so a window is sized
    # by its own time, not by a sample elsewhere
    windows = []
    t = time.perf_counter()
    for i, (m, d) in enumerate(motifs):
        wins = []
        for b in rng.integers(0, E - 8192, nwin):
            beg = cur = int(b)
            chunk, spent = cal, 0.0
            while cur - beg < 8192 and (cur == beg or spent < target):
                t1 = time.perf_counter()
                oracle.count(g, m, d, cur, cur + chunk, nthreads=threads)
                spent += time.perf_counter() - t1
                cur += chunk
                chunk = min(8192 - (cur - beg), chunk * 2 if cur - beg >= 2 * cal else chunk)
            wins.append((beg, cur - beg))
        windows.append(wins)
        if (i + 1) % 64 == 0:
            log("cpu baseline: sized the windows of %d/%d queries, %.1fs" % (i + 1, len(motifs), time.perf_counter() - t))
    reps, spread = [], []
    for rep in range(3):
        est_s, used = 0.0, 0.0
        lo_hi = [0.0, 0.0]
        for i, (m, d) in enumerate(motifs):
            roots = E - len(m) + 1
            per_root = []
            for beg, span in windows[i]:
                t = time.perf_counter()
                c = oracle.count(g, m, d, beg, beg + span, nthreads=threads)
                dt = time.perf_counter() - t
                used += dt
                per_root.append(dt / span)
                if rep == 0:
                    assert dg.count(m, d, beg, beg + span) == c, "GPU/oracle mismatch on a CPU-baseline window"
            est_s += float(np.mean(per_root)) * roots
            lo_hi[0] += min(per_root) * roots
            lo_hi[1] += max(per_root) * roots
            if (i + 1) % 64 == 0:
                log("cpu baseline repetition %d: %d/%d queries, %.1fs" % (rep, i + 1, len(motifs), used))
        reps.append(len(motifs) / est_s)
        spread.append((len(motifs) / lo_hi[1], len(motifs) / lo_hi[0]))
        log("cpu baseline repetition %d: %.6f queries/s (windows %.1fs of wall time)" % (rep, reps[-1], used))
        if used > budget_s * 3:
            break
    order = np.argsort(reps)
    med = int(order[len(order) // 2])
    spans = [w[0][1] for w in windows]
    return {"value": round(reps[med], 6), "unit": "queries/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "repetitions": [round(x, 6) for x in reps],
            "window_spread_qps": [round(spread[med][0], 6), round(spread[med][1], 6)],
            "sample": "C oracle (the reference's search restated), all %d queries x %d seeded windows of %d-%d "
                      "roots (median %d; each grown from %d roots until it took its share of the time), time per root "
                      "extrapolated to each query's %d roots; median of %d repetitions on %d threads; GPU "
                      "bit-exact on every window" % (len(motifs), nwin, min(spans), max(spans),
                                                      int(np.median(spans)), cal, E, len(reps), threads)}


if __name__ == "__main__":
    main()
