"""Autotuner: launch parameters per motif shape class, chosen by measurement.

A shape class (tm_shape_class: LDS stack depth and node-map width of a
launch) fixes a launch's occupancy, so the best register cap, donation
trigger and last-edge scan threshold are chosen per class: every candidate
setting is timed on the class's heaviest queries among those that finish
within a short screening limit (single-query launches, HIP-event kernel
time, each capped by a time limit that counts as its time when hit), and the
candidate with the smallest total wins (ties within 2% go to the default).  The
table is saved as JSON and installed on a graph with tm_set_class_tuning,
after which tm_count, tm_batch_count and the server use it whenever a call
brings no launch parameters of its own.

    python -m everest_amd.autotune --graph pl64m --out everest_amd/tuned/pl64m.json
"""
import argparse
import json
import os
import sys
import time

from . import abi

# register cap (5: 96 VGPRs, 6: 80 VGPRs), donation trigger, scan threshold,
# survivor compaction (the library default uses it where its LDS buffer costs
# no resident wavefronts; the last three force it off or on)
CANDIDATES = [
    dict(waves_per_simd_hint=5),
    dict(waves_per_simd_hint=6),
    dict(waves_per_simd_hint=5, steal_after_us=250),
    dict(waves_per_simd_hint=5, scan_max=32),
    dict(waves_per_simd_hint=5, sb_flush=-1),
    dict(waves_per_simd_hint=5, sb_flush=32, sb_inline=32),
    dict(waves_per_simd_hint=5, sb_flush=48, sb_inline=48),
]


def autotune(dg, queries, candidates=CANDIDATES, per_class=4, limit_ms=1500, screen_ms=300, log=None):
    """queries: list of (edges as (a, b, k) triples, delta).  Returns
    (table {class: tuning dict}, measurements {class: [(candidate, total_ms)]}).

    The sample of a class is its per_class heaviest queries among those that
    finish within screen_ms under the first candidate (the heavy queries
    decide a batch's time; ones that do not finish would only measure the
    limit)."""
    by_class = {}
    for m, d in queries:
        by_class.setdefault(abi.shape_class(m), []).append((m, d))
    table, meas = {}, {}
    for cls in sorted(by_class):
        qs = by_class[cls]
        timed = []
        for m, d in qs:
            try:
                dg.count(m, d, tune=abi.tuning(time_limit_ms=screen_ms, **candidates[0]))
                timed.append((dg.last_kernel_ms(), m, d))
            except abi.TMError as e:
                if e.status != abi.TM_ERR_TIMEOUT:
                    raise
        timed.sort(key=lambda x: -x[0])
        pick = [(m, d) for _, m, d in timed[:per_class]]
        if not pick:
            continue
        rows = []
        for cand in candidates:
            total = 0.0
            for m, d in pick:
                try:
                    dg.count(m, d, tune=abi.tuning(time_limit_ms=limit_ms, **cand))
                    total += dg.last_kernel_ms()
                except abi.TMError as e:
                    if e.status != abi.TM_ERR_TIMEOUT:
                        raise
                    total += limit_ms
            rows.append((cand, round(total, 3)))
            if log:
                log("class %d (%d queries, %d sampled): %s -> %.1f ms" % (cls, len(qs), len(pick), cand, total))
        # ties within 2% go to the earlier candidate (the first is the library default)
        best_ms = min(r[1] for r in rows)
        best = next(c for c, t in rows if t <= best_ms * 1.02)
        table[cls] = dict(best)
        meas[cls] = rows
    return table, meas


def apply(dg, table):
    for cls, t in table.items():
        dg.set_class_tuning(int(cls), abi.tuning(**t))


def save(path, table, meas, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"meta": meta, "table": {str(k): v for k, v in table.items()},
                   "measured_ms": {str(k): v for k, v in meas.items()}}, f, indent=1)


def load(path):
    with open(path) as f:
        d = json.load(f)
    return {int(k): v for k, v in d["table"].items()}, d.get("meta", {})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graph", default="pl64m")
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--per-class", type=int, default=6)
    ap.add_argument("--limit-ms", type=int, default=1500)
    ap.add_argument("--screen-ms", type=int, default=300)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    from . import core
    from .gen import graphs, queries

    def log(*a):
        print("#", *a, file=sys.stderr, flush=True)

    U, V, T = graphs.make(args.graph)
    el = core.data.EdgeListLoader.fromArrays(args.graph, U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    dg = abi.DeviceGraph(hg.array("Eg"), hg.array("inEdgesV"), hg.array("outEdgesV"), hg.array("inEdgesR"),
                         hg.array("outEdgesR"), hg.array("Euid"))
    qs = [(queries.as_motif(e), d) for e, d in queries.batch_queries(args.queries, seed=args.seed)]
    t = time.time()
    table, meas = autotune(dg, qs, per_class=args.per_class, limit_ms=args.limit_ms, screen_ms=args.screen_ms,
                           log=log)
    meta = dict(graph=args.graph, queries=args.queries, seed=args.seed, per_class=args.per_class,
                limit_ms=args.limit_ms, screen_ms=args.screen_ms, seconds=round(time.time() - t, 1),
                command="python -m everest_amd.autotune --graph %s --queries %d --seed %d --per-class %d "
                        "--limit-ms %d --screen-ms %d --out <file>" % (
                            args.graph, args.queries, args.seed, args.per_class, args.limit_ms, args.screen_ms))
    log("table", table)
    if args.out:
        save(args.out, table, meas, meta)
    print(json.dumps({str(k): v for k, v in table.items()}))


if __name__ == "__main__":
    main()
