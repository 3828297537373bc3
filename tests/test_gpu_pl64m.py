"""GPU parity at the benchmark's full size: the 64M-edge power-law graph
(2.6M nodes, 7.5 years), with the deeper-search motifs (4-cycle, 5-edge hub,
6-edge bidirectional triangle at delta 3600 and 86400), the whole 256-query
batch set, and the serving loop on the config-4 query mix.

Counts are checked by size-independent properties (additivity over root
ranges, batched == single, pruned == the instrumented kernel's unpruned
search, two-edge tail on == off) and against the oracle on root windows that
hold matches.  Every test prints its case before the GPU call.
"""
import sys
import threading

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pl64m():
    from everest_amd import abi, core
    from everest_amd.gen import graphs
    say("pl64m: generating and preparing the graph")
    U, V, T = graphs.make("pl64m")
    el = core.data.EdgeListLoader.fromArrays("pl64m", U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    say("pl64m: upload")
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    yield g, dg
    dg.close()


def _window_with_matches(g, dg, mot, delta, rng, width=20_000, tries=12):
    """A seeded root window whose oracle count is positive (or the last tried)."""
    for _ in range(tries):
        beg = int(rng.integers(0, g["E"] - width))
        want = oracle.count(g, mot, delta, beg, beg + width, nthreads=16)
        if want > 0:
            return beg, beg + width, want
    return beg, beg + width, want


@pytest.mark.parametrize("name", ["cycle4", "hub5", "bitri6"])
@pytest.mark.parametrize("delta", [3600, 86400])
def test_config3_deeper_motifs(pl64m, name, delta):
    from everest_amd import abi
    from everest_amd.gen import queries
    g, dg = pl64m
    E = g["E"]
    mot = queries.as_motif(queries.CONFIG3_MOTIFS[name])
    say("config3", name, "delta", delta, ": full count")
    full = dg.count(mot, delta)
    ms = dg.last_kernel_ms()
    say("  count", full, "kernel ms %.1f" % ms)
    cuts = [0, E // 3, 2 * E // 3, -1]
    parts = [dg.count(mot, delta, cuts[k], cuts[k + 1]) for k in range(3)]
    assert sum(parts) == full, (parts, full)
    say("  two-edge tail off")
    assert dg.count(mot, delta, tune=abi.tuning(tail_swap=-1)) == full
    beg = 20_000_000
    say("  pruned vs unpruned on [%d, %d)" % (beg, beg + 1_000_000))
    got = dg.count(mot, delta, beg, beg + 1_000_000)
    want, _ = dg.count_work(mot, delta, beg, beg + 1_000_000)
    assert got == want, (got, want)
    rng = np.random.default_rng(sum(map(ord, name)) * 100003 + delta)
    b0, b1, want = _window_with_matches(g, dg, mot, delta, rng)
    say("  oracle window [%d, %d): %d matches" % (b0, b1, want))
    assert dg.count(mot, delta, b0, b1) == want


def test_config4_all_256_batch_equals_single(pl64m):
    """All 256 queries of the batch workload: one tm_batch_count call equals
    the 256 single-query counts (single-query launches back to back, and one
    multi-query launch per shape class); the two-edge tail off gives the same counts
    for the queries it applies to; oracle windows on a sample."""
    from everest_amd import abi
    from everest_amd.gen import queries
    g, dg = pl64m
    qs = queries.batch_queries(256, seed=4)
    motifs = [(queries.as_motif(e), d) for e, d in qs]
    single = []
    for i, (m, d) in enumerate(motifs):
        if i % 32 == 0:
            say("config4 single", i)
        single.append(dg.count(m, d))
    say("config4 batch of 256")
    assert dg.batch_count(motifs) == single
    say("config4 batch of 256, one multi-query launch per shape class")
    assert dg.batch_count(motifs, tune=abi.tuning(batch_mode=1)) == single
    rng = np.random.default_rng(44)
    for i in [int(x) for x in rng.choice(256, 8, replace=False)]:
        m, d = motifs[i]
        say("config4 query", i, qs[i], "tail off / oracle window")
        assert dg.count(m, d, tune=abi.tuning(tail_swap=-1)) == single[i]
        b0 = int(rng.integers(0, g["E"] - 5_000))
        assert dg.count(m, d, b0, b0 + 5_000) == oracle.count(g, m, d, b0, b0 + 5_000, nthreads=16)


def test_config5_server_mix(pl64m):
    """The serving loop on the 64M-edge graph with the config-4 query mix, 20%
    enumerate requests capped at 1M rows: every result is the same alone
    (requests one at a time) and in a mix (submitted together from three
    threads), and equals tm_count / tm_enum; tm_enum equals the oracle's rows
    on a root window."""
    from everest_amd import abi
    from everest_amd.gen import queries
    g, dg = pl64m
    qs = queries.batch_queries(256, seed=4)
    rng = np.random.default_rng(55)
    reqs = []
    for k in range(40):
        i = int(rng.integers(0, 256))
        reqs.append((queries.as_motif(qs[i][0]), qs[i][1], 1_000_000 if k % 5 == 0 else -1, i))
    want = []
    for mot, d, cap, i in reqs:
        say("server reference: query", i, "cap", cap)
        want.append(dg.enum(mot, d, cap=cap) if cap >= 0 else (dg.count(mot, d), None))

    def canon(rows):
        return rows[np.lexsort(rows.T[::-1])] if len(rows) else rows

    def check(k, r, srv, t):
        mot, d, cap, i = reqs[k]
        c, rows_want = want[k]
        assert r["status"] == 0, (k, r)
        assert r["count"] == c, (k, i, r["count"], c)
        if cap < 0:
            srv.release(t)
            return
        rows = srv.take_rows(t, len(mot), cap)
        assert np.array_equal(canon(rows), canon(rows_want)), (k, i)

    say("server alone")
    with abi.Server(dg) as srv:
        for k, (mot, d, cap, i) in enumerate(reqs):
            t = srv.submit(mot, d, cap)
            check(k, srv.wait(t), srv, t)
    say("server mix from 3 threads")
    with abi.Server(dg) as srv:
        tickets = [None] * len(reqs)

        def client(part):
            for k in range(part, len(reqs), 3):
                mot, d, cap, i = reqs[k]
                tickets[k] = srv.submit(mot, d, cap)

        th = [threading.Thread(target=client, args=(p,)) for p in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        biggest = 0
        for k in range(len(reqs)):
            r = srv.wait(tickets[k])
            biggest = max(biggest, r["batch"])
            check(k, r, srv, tickets[k])
        say("server stats", srv.stats(), "largest batch", biggest)
    # tm_enum against the oracle's rows on a root window with matches
    for mot, d, cap, i in reqs[:10]:
        if cap < 0:
            continue
        b0, b1, n = _window_with_matches(g, dg, mot, d, rng, width=5_000, tries=6)
        say("enum window query", i, b0, b1, n)
        total, orows = oracle.enum(g, mot, d, n + 1, b0, b1)
        c, rows = dg.enum(mot, d, cap=n + 1, beg=b0, end=b1)
        assert c == total == n
        assert np.array_equal(canon(rows), canon(orows))
