#!/usr/bin/env python3
"""Benchmark: temporal-motif queries on one MI355X.

Workload (default): 256 random motif queries (3-6 edges on <= 4 nodes, delta
log-uniform in [600, 604800] s) on the resident 64M-edge power-law graph.

  plain   one query per launch, sequentially (tm_count), host wall latency each
  batch   all queries in one persistent launch (tm_batch_count); a step is one
          such launch over the whole query set

Every batched count is checked bit-exact against the plain count of the same
query.  The heaviest queries are left out of every pass: for the default set a
fixed list (the 18 whose count took over 2 s when the workload was fixed, so
that the query set does not change as the kernel gets faster); for any other
set, or with --screen, the queries over --qlimit-ms in a screening pass.  The
excluded indexes are reported in the JSON line.

value = batched queries/s.  roofline: algorithmic bytes of the reference's own
work (12 R + 8 X1 + 16 X2 + 16 P, counted exactly by the instrumented kernel)
per launch / mean launch duration (HIP events on the library stream), for the
single-query kernel instantiation with the largest total time in the plain
pass.  cpu_baseline: the C oracle (a multithreaded restatement of the
reference DFS) on a bounded root-range sample of the same queries.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

# The default set's queries whose count took over 2 s (reference work of about
# 10^13 probes and more each): left out so the workload stays the same set.
EXCLUDED_PL64M_256_SEED4 = (14, 17, 21, 22, 45, 52, 61, 64, 82, 87, 126, 132, 149, 181, 234, 235, 241, 244)
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("#", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--graph", default="pl64m")
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--qlimit-ms", type=int, default=2000, help="per-query budget in the plain pass")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget (0 = skip)")
    ap.add_argument("--json-out", default="")
    ap.add_argument("--plain-only", action="store_true", help="stop after the plain pass (counter runs)")
    ap.add_argument("--over", default="", help="comma-separated excluded query indexes (skips the screening pass)")
    ap.add_argument("--screen", action="store_true", help="screen the set against --qlimit-ms instead of the fixed list")
    ap.add_argument("--no-tuning", action="store_true", help="ignore the autotuned table (library defaults)")
    ap.add_argument("--serve-requests", type=int, default=240, help="requests in the serving run (0 = skip)")
    ap.add_argument("--serve-qps", type=float, default=200.0, help="Poisson arrival rate of the serving run")
    ap.add_argument("--serve-enum-frac", type=float, default=0.2, help="share of enumerate requests")
    ap.add_argument("--serve-enum-cap", type=int, default=1000000, help="row cap of enumerate requests")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also copy this many bytes device-to-device (HBM counter scale for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X "
              "(--gpus %d requested; multi-GPU is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from everest_amd import abi, core
    from everest_amd.gen import graphs, queries

    if abi.lib().tm_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)

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

    # ---- screen: queries over the per-query budget are left out of both
    # passes.  Screening runs each query alone through the batch kernel (a
    # different instantiation), so the single-query kernel's launches below
    # are exactly the kept queries.
    lim = abi.tuning(time_limit_ms=args.qlimit_ms)
    over = [int(x) for x in args.over.split(",") if x.strip()]
    fixed = (args.graph, args.queries, args.seed) == ("pl64m", 256, 4) and not args.over and not args.screen
    if fixed:
        over = list(EXCLUDED_PL64M_256_SEED4)
        log("excluded (fixed list of the default set): %s" % over)
    t = time.perf_counter()
    for i, (m, d) in enumerate(motifs if not (args.over or fixed) else []):
        try:
            dg.batch_count([(m, d)], tune=lim)
        except abi.TMError as e:
            if e.status != abi.TM_ERR_TIMEOUT:
                raise
            over.append(i)
            log("query %d over budget (%d ms): %s delta=%d" % (i, args.qlimit_ms, qs[i][0], d))
        if (i + 1) % 64 == 0:
            log("screen: %d/%d queries, %.1fs" % (i + 1, len(motifs), time.perf_counter() - t))

    # ---- plain: one query per launch
    plain_ms, plain_counts, kernel_ms, variant = [], [], [], {}
    for i, (m, d) in enumerate(motifs):
        if i in over:
            continue
        t = time.perf_counter()
        c = dg.count(m, d)
        plain_ms.append((time.perf_counter() - t) * 1e3)
        kernel_ms.append(dg.last_kernel_ms())
        plain_counts.append((i, c))
        st = dg.last_stats()
        nmap = 4 if max(max(a, b) for a, b, _ in m) + 1 <= 4 else 8
        variant[i] = (nmap, st["wpe"])
        if (i + 1) % 32 == 0:
            log("plain: %d/%d queries, %.1fs so far" % (i + 1, len(motifs), sum(plain_ms) / 1e3))
    keep = [i for i, _ in plain_counts]
    plain_total_s = sum(plain_ms) / 1e3
    plain_qps = len(keep) / plain_total_s
    log("plain: %d queries in %.3fs (%.1f q/s), %d over budget" % (len(keep), plain_total_s, plain_qps, len(over)))

    if args.plain_only:
        log("plain pass only (--plain-only): no reference work, batch or CPU baseline")
        calibrate_copy(args.calibrate_copy)
        return

    # ---- reference work of the kept queries (instrumented kernel, untimed)
    work = np.zeros(4, dtype=np.float64)
    balg = {}
    t = time.perf_counter()
    for n, i in enumerate(keep):
        c, w = dg.count_work(motifs[i][0], motifs[i][1])
        assert c == dict(plain_counts)[i]
        work += np.array(w, dtype=np.float64)
        balg[i] = 12.0 * w[0] + 8.0 * w[1] + 16.0 * w[2] + 16.0 * w[3]
        if (n + 1) % 32 == 0:
            log("reference work: %d/%d queries, %.1fs" % (n + 1, len(keep), time.perf_counter() - t))
    b_alg = 12 * work[0] + 8 * work[1] + 16 * work[2] + 16 * work[3]
    log("reference work: R=%d X1=%d X2=%d P=%d -> %.3f GB" % (*work, b_alg / 1e9))
    # roofline of the dominant single-query kernel instantiation (largest
    # total time): algorithmic bytes per launch / mean launch duration (HIP
    # events on the library's stream), over that instantiation's launches
    kms = dict(zip(keep, kernel_ms))
    groups = {}
    for i in keep:
        groups.setdefault(variant[i], []).append(i)
    dom = max(groups, key=lambda k: sum(kms[i] for i in groups[k]))
    dom_n = len(groups[dom])
    dom_bytes = sum(balg[i] for i in groups[dom]) / dom_n
    dom_ms = sum(kms[i] for i in groups[dom]) / dom_n
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    # the instantiation's name as the profiler prints it: mine_kernel<MAXE, MAXN,
    # ENUM, INSTR, spec, waves-per-SIMD register cap, multi-query>
    kname = "void tmk::mine_kernel<8, %d, false, false, tmk::DynSpec, %d, false>(tm_kernel_args)" % dom
    log("dominant kernel %s: %d launches, %.3f GB and %.3f ms per launch -> %.1f GB/s" % (
        kname, dom_n, dom_bytes / 1e9, dom_ms, achieved))
    # HBM bytes per launch of that kernel on this workload, from the committed
    # counter summary (two rocprofv3 --pmc runs, scaled by a 1 GiB copy)
    # (per query, so the mean covers exactly this run's launches of that kernel)
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", "mine_kernel_pl64m_config4_hbm.json")
    if os.path.exists(tf) and args.graph == "pl64m" and args.queries == 256 and args.seed == 4:
        with open(tf) as f:
            t = json.load(f)
        perq = t.get("hbm_bytes_per_query", {})
        if t.get("kernel") == kname and all(str(i) in perq for i in groups[dom]):
            traffic = round(sum(perq[str(i)] for i in groups[dom]) / dom_n, 1)
            traffic_src = "profiles/mine_kernel_pl64m_config4_hbm.md (FETCH_SIZE x %.3f + WRITE_SIZE x %.3f)" % (
                t["fetch_scale"], t["write_scale"])

    # ---- batched: all kept queries in one launch per step
    batch = [motifs[i] for i in keep]
    want = [c for _, c in plain_counts]
    for _ in range(args.warmup):
        got = dg.batch_count(batch)
        assert got == want, "batched counts differ from single-query counts"
    step_ms, bkern_ms = [], []
    for s in range(args.steps):
        t = time.perf_counter()
        got = dg.batch_count(batch)
        step_ms.append((time.perf_counter() - t) * 1e3)
        bkern_ms.append(dg.last_kernel_ms())
        assert got == want, "batched counts differ from single-query counts"
        log("step %d: %.2f ms (kernel %.2f ms)" % (s, step_ms[-1], bkern_ms[-1]))
    ms_per_step = float(np.mean(step_ms))
    batch_qps = len(keep) / (ms_per_step / 1e3)
    matches = float(sum(want))

    # ---- serving: Poisson arrivals through the server vs one query per launch, FIFO
    serving = None
    if args.serve_requests > 0:
        serving = serve(dg, motifs, keep, dict(plain_counts), dict(zip(keep, plain_ms)), args)

    # ---- CPU baseline: the oracle on a bounded sample
    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(hg, arrs, motifs, keep, dict(plain_counts), dg, args.cpu_seconds)

    res = {
        "metric": "batched queries/s (256 random motif queries, resident 64M-edge graph)",
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
            "workload": "config4-batched: %d queries (%d kept, %d heaviest excluded: %s), M=3..6 on <=4 nodes, "
                        "delta log-uniform [600,604800]s, graph %s E=%d" % (
                            len(motifs), len(keep), len(over),
                            "fixed list, each over 2000 ms" if fixed else "over %d ms when screened" % args.qlimit_ms,
                            args.graph, hg.numEdges),
            "plain_qps": round(plain_qps, 3),
            "batched_over_plain": round(batch_qps / plain_qps, 3),
            "plain_latency_ms_p50": round(float(np.percentile(plain_ms, 50)), 3),
            "plain_latency_ms_p99": round(float(np.percentile(plain_ms, 99)), 3),
            "batch_latency_ms_p50": round(float(np.percentile(step_ms, 50)), 3),
            "batch_latency_ms_p99": round(float(np.percentile(step_ms, 99)), 3),
            "matches_per_s": round(matches / (ms_per_step / 1e3), 1),
            "over_budget_queries": over,
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
            "mean_launch_ms": round(dom_ms, 4),
            "algorithmic_bytes_all_kept_queries": b_alg,
        },
        "cpu_baseline": cpu,
    }
    line = json.dumps(res)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")


def serve(dg, motifs, keep, counts, plain_ms, args):
    """Serving run: requests drawn from the kept queries (seeded), a share of
    them enumerate requests, arriving as a Poisson stream; submitted to a
    server (continuous batching over launch slots) by a client thread at
    their arrival times.  Latency = arrival -> result on the host.  The plain
    comparator serves the same stream one query per launch in arrival order
    (a FIFO queue in front of tm_count / tm_enum), its service times measured
    by running each request alone."""
    import threading
    import numpy as np
    from everest_amd import abi

    rng = np.random.default_rng(args.seed + 100)
    n = args.serve_requests
    pick = [keep[int(j)] for j in rng.integers(0, len(keep), n)]
    enum = rng.random(n) < args.serve_enum_frac
    gaps = rng.exponential(1.0 / args.serve_qps, n)
    arrive = np.cumsum(gaps) - gaps[0]
    cap = args.serve_enum_cap

    # plain comparator: service time of each request alone (count: the plain
    # pass's wall time; enumerate: measured here), then a FIFO queue
    svc = np.zeros(n)
    enum_ms = {}
    for k in range(n):
        i = pick[k]
        if enum[k]:
            if i not in enum_ms:
                m, d = motifs[i]
                t = time.perf_counter()
                c, rows = dg.enum(m, d, cap=cap)
                enum_ms[i] = (time.perf_counter() - t) * 1e3
            svc[k] = enum_ms[i]
        else:
            svc[k] = plain_ms[i]
    done = np.zeros(n)
    free_at = 0.0
    for k in range(n):
        start = max(arrive[k] * 1e3, free_at)
        free_at = start + svc[k]
        done[k] = free_at
    plain_lat = done - arrive * 1e3
    plain_span = done.max() / 1e3

    log("serving: %d requests (%d enumerate, cap %d) at %.0f/s offered" % (n, int(enum.sum()), cap, args.serve_qps))
    srv = abi.Server(dg, max_batch=256)  # the library's slot count
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
    lat, batch, ok = np.zeros(n), np.zeros(n), True
    for k in range(n):
        while tickets[k] is None:
            time.sleep(0.0005)
        r = srv.wait(tickets[k])
        lat[k], batch[k] = r["latency_ms"], r["batch"]
        want = counts[pick[k]]
        if enum[k]:
            ok &= r["status"] == 0 and (r["count"] == want if want <= cap else r["count"] >= cap)
            srv.take_rows(tickets[k], len(motifs[pick[k]][0]), min(want, cap))
        else:
            ok &= r["status"] == 0 and r["count"] == want
            srv.release(tickets[k])
    span = time.perf_counter() - t0
    th.join()
    st = srv.stats()
    srv.close()
    assert ok, "served results differ from single-query results"
    res = {
        "requests": n, "enumerate_requests": int(enum.sum()), "enum_cap": cap, "offered_qps": args.serve_qps,
        "served_qps": round(n / span, 3), "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "plain_fifo_qps": round(n / plain_span, 3),
        "plain_fifo_latency_ms_p50": round(float(np.percentile(plain_lat, 50)), 3),
        "plain_fifo_latency_ms_p99": round(float(np.percentile(plain_lat, 99)), 3),
        "served_over_plain_qps": round((n / span) / (n / plain_span), 3),
        "launches": st["launches"], "largest_batch": st["max_batch"], "mean_batch": round(float(batch.mean()), 2),
        "bit_exact_served_vs_plain": True,
    }
    log("serving: %.1f q/s served (plain FIFO %.1f), p50 %.1f ms p99 %.1f ms (plain FIFO p50 %.1f p99 %.1f)" % (
        res["served_qps"], res["plain_fifo_qps"], res["latency_ms_p50"], res["latency_ms_p99"],
        res["plain_fifo_latency_ms_p50"], res["plain_fifo_latency_ms_p99"]))
    return res


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


def cpu_baseline(hg, arrs, motifs, keep, counts, dg, budget_s):
    """Oracle throughput on root-range samples of the kept queries, in queries/s
    extrapolated per query from (sampled roots / all roots); each sample's count
    is also checked against the GPU on the same root range."""
    import numpy as np
    from oracle import oracle

    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    threads = min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(17)
    E = hg.numEdges
    span = 1 << 18  # about 10 s of CPU work over the kept queries on 16 threads
    est_full_s, n, checked, t_used = 0.0, 0, 0, 0.0
    for i in keep:
        if t_used > budget_s:
            break
        m, d = motifs[i]
        beg = int(rng.integers(0, E - span))
        t = time.perf_counter()
        c = oracle.count(g, m, d, beg, beg + span, nthreads=threads)
        dt = time.perf_counter() - t
        t_used += dt
        gc, _ = dg.count_work(m, d, beg, beg + span)  # the instrumented kernel: not the one timed above
        assert gc == c, "GPU/oracle mismatch on sample"
        checked += 1
        roots = E - len(m) + 1
        est_full_s += dt * roots / span
        n += 1
    if n == 0:
        return None
    return {"value": round(n / est_full_s, 6), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": "C oracle (reference DFS restated), %d queries x %d seeded roots each, extrapolated to all "
                      "%d roots; %.1fs of CPU time; GPU bit-exact on every sample" % (n, span, E, t_used)}


if __name__ == "__main__":
    main()
