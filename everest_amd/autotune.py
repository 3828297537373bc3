"""Autotuner: launch parameters per motif shape class, chosen by measurement.

A shape class (tm_shape_class: LDS stack depth and node-map width of a
launch) fixes a launch's occupancy, so launch parameters are chosen per
class.  Every query of the set is first timed once with the library
defaults (single-query launches, HIP-event kernel time); each class's
sample is its heaviest queries, the ones that decide a batch step's time.
Then, per class, a coordinate descent over the axes below: starting from the
defaults, each axis's values are tried in turn on top of the best setting so
far, and a value is kept when it cuts the sample's total by more than 2%.  A
candidate stops as soon as its running total (heaviest query first) can no
longer win, and every run's count must equal the default run's.  The table
is saved as JSON and installed on a graph with tm_set_class_tuning, after
which tm_count, tm_batch_count and the server use it whenever a call brings
no launch parameters of its own.

    python -m everest_amd.autotune --graph pl64m --out everest_amd/tuned/pl64m.json
"""
import argparse
import json
import os
import sys
import time

from . import abi

# (field, values): the first value, 0, is the library default
AXES = [
    ("waves_per_simd_hint", [0, 6, 4, 1]),  # register cap: 0 = 5 (96 VGPRs), 6 = 80, 4 = 128, 1 = none
    ("tail_swap", [0, 2, 4, 8, -1]),        # two-edge tail threshold (0 = 32, -1 = off)
    ("sb_flush", [0, -1, 48, 64]),          # survivor compaction (0 = 32 where it costs no occupancy)
    ("sb_inline", [0, 16, 48]),             # survivors a step expands itself with none buffered (0 = 32)
    ("scan_max", [0, 8, 32]),               # last-edge ranges scanned up to this size (0 = 16)
    ("claim_blocks", [0, 1, 2, 8, 32]),     # 256-root blocks per claim (0 = ~8 claims per wavefront)
    ("waves_per_cu", [0, 16, 24]),          # resident wavefronts per CU (0 = occupancy-limited, <= 32)
    ("steal_after_us", [0, 250, 4000]),     # donation trigger (0 = 1000 us)
    ("steal_min", [0, 16, 256]),            # smallest donated half-frame (0 = 64 candidates)
    ("steal_max", [0, 2, 8]),               # frames donated per look (0 = 4)
]


def _run(dg, m, d, tune):
    """Kernel ms of one count, or None when it hit its time limit; and the count."""
    try:
        c = dg.count(m, d, tune=abi.tuning(**tune))
        return dg.last_kernel_ms(), c
    except abi.TMError as e:
        if e.status != abi.TM_ERR_TIMEOUT:
            raise
        return None, None


def autotune(dg, queries, axes=AXES, per_class=6, limit_ms=20000, log=None, start=None, passes=2):
    """queries: list of (edges as (a, b, k) triples, delta).  Returns
    (table {class: tuning dict}, measurements {class: [(candidate, total_ms
    or None when abandoned)]}, screen [(class, ms)] of the default pass).
    start: an earlier table {class: tuning dict}; a class's entry there is
    the descent's starting point when it beats the defaults.  The descent
    repeats over the axes until a pass keeps no new value (at most passes)."""
    base = {}
    screen = []
    for i, (m, d) in enumerate(queries):
        ms, c = _run(dg, m, d, dict(time_limit_ms=limit_ms))
        if ms is None:
            raise RuntimeError("query %d exceeds %d ms with the defaults" % (i, limit_ms))
        base[i] = (ms, c)
        screen.append((abi.shape_class(m), round(ms, 3)))
    if log:
        log("default pass: %d queries, %.1f s of kernel time" % (len(queries), sum(v[0] for v in base.values()) / 1e3))
    by_class = {}
    for i, (m, d) in enumerate(queries):
        by_class.setdefault(abi.shape_class(m), []).append(i)
    table, meas = {}, {}
    for cls in sorted(by_class):
        pick = sorted(by_class[cls], key=lambda i: -base[i][0])[:per_class]
        best, best_ms = {}, sum(base[i][0] for i in pick)
        rows = [({}, round(best_ms, 3))]
        if log:
            log("class %d: %d queries, sample %s, defaults %.1f ms" % (cls, len(by_class[cls]), pick, best_ms))

        def measure(cand, bound):
            total = 0.0
            for i in pick:  # heaviest first
                m, d = queries[i]
                lim = int(min(limit_ms, max(50.0, 3.0 * base[i][0])))
                ms, c = _run(dg, m, d, dict(cand, time_limit_ms=lim))
                if ms is not None and c != base[i][1]:
                    raise RuntimeError("count of query %d differs under %s" % (i, cand))
                total += lim if ms is None else ms
                if total > bound:
                    break
            return total

        first = dict((start or {}).get(cls, {}))
        if first:
            total = measure(first, best_ms * 1.02)
            rows.append((first, round(total, 3) if total <= best_ms * 1.02 else None))
            if log:
                log("class %d: start %s -> %.1f ms" % (cls, first, total))
            if total < best_ms:
                best, best_ms = first, total
        seen = {tuple(sorted(best.items()))}
        for _ in range(passes):
            moved = False
            for field, values in axes:
                for v in values:
                    if v == best.get(field, 0):
                        continue
                    cand = dict(best)
                    if v == 0:
                        cand.pop(field, None)
                    else:
                        cand[field] = v
                    key = tuple(sorted(cand.items()))
                    if key in seen:
                        continue
                    seen.add(key)
                    total = measure(cand, best_ms * 1.02)
                    won = total < best_ms * 0.98
                    rows.append((cand, round(total, 3) if total <= best_ms * 1.02 or won else None))
                    if log:
                        log("class %d: %s -> %s" % (cls, cand, "%.1f ms" % total if rows[-1][1] is not None
                                                      else "abandoned above %.1f ms" % (best_ms * 1.02)))
                    if won:
                        best, best_ms = cand, total
                        moved = True
            if not moved:
                break
        table[cls] = best
        meas[cls] = rows
        if log:
            log("class %d: chose %s (%.1f ms)" % (cls, best, best_ms))
    return table, meas, screen


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
    ap.add_argument("--limit-ms", type=int, default=20000)
    ap.add_argument("--out", default="")
    ap.add_argument("--start", default="", help="an earlier table (JSON) whose entries start the descent")
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
    start = load(args.start)[0] if args.start else None
    table, meas, screen = autotune(dg, qs, per_class=args.per_class, limit_ms=args.limit_ms, log=log, start=start)
    meta = dict(graph=args.graph, queries=args.queries, seed=args.seed, per_class=args.per_class,
                limit_ms=args.limit_ms, seconds=round(time.time() - t, 1), axes=[f for f, _ in AXES],
                default_pass_ms=screen,
                command="python -m everest_amd.autotune --graph %s --queries %d --seed %d --per-class %d "
                        "--limit-ms %d%s --out <file>" % (args.graph, args.queries, args.seed, args.per_class,
                                                          args.limit_ms, " --start " + args.start if args.start else ""))
    log("table", table)
    if args.out:
        save(args.out, table, meas, meta)
    print(json.dumps({str(k): v for k, v in table.items()}))


if __name__ == "__main__":
    main()
