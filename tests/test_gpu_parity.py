"""GPU parity: HIP kernels (through the C ABI) against the oracle, bit-exact.

Every test prints its case (seed, sizes, motif, delta) before the GPU call.
"""
import sys

import numpy as np
import pytest

from oracle import oracle, prep
from helpers import random_graph, random_motif, all_3edge_motifs, deterministic_cut

pytestmark = pytest.mark.gpu


def _abi():
    from everest_amd import abi
    return abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_cases():
    rng = np.random.default_rng(11)
    cases = []
    for i in range(60):
        E = int(rng.integers(1, 400))
        N = int(rng.integers(2, 30))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, 200)))
        M = int(rng.integers(2, 7))
        mot = random_motif(rng, M)
        delta = int(rng.integers(0, 80))
        cases.append((i, U, V, T, mot, delta))
    return cases


def test_count_small_random(small_cases):
    abi = _abi()
    for (i, U, V, T, mot, delta) in small_cases:
        g = prep.preprocess(U, V, T)
        say("case", i, "E", g["E"], "motif", mot, "delta", delta)
        want = oracle.count(g, mot, delta, nthreads=4)
        dg = abi.DeviceGraph.from_dict(g)
        for scan in (1, 16):
            got = dg.count(mot, delta, tune=abi.tuning(scan_max=scan))
            assert got == want, (i, mot, delta, scan, got, want)
        dg.close()


def test_enum_small_random(small_cases):
    """Below the cap the row set is the oracle's; at a cap of a third or a
    tenth of the matches it is the deterministic cut (every row of the roots
    before the root that reaches the cap, then that root's smallest rows),
    whatever the schedule: two runs with different launch shapes agree."""
    abi = _abi()
    beyond = 0
    for (i, U, V, T, mot, delta) in small_cases[:30]:
        g = prep.preprocess(U, V, T)
        total, rows = oracle.enum(g, mot, delta, 1 << 20)
        say("enum case", i, "motif", mot, "delta", delta, "total", total)
        dg = abi.DeviceGraph.from_dict(g)
        cnt, grows = dg.enum(mot, delta, cap=max(total, 1) + 1)
        assert cnt == total
        assert sorted(map(tuple, grows.tolist())) == sorted(map(tuple, rows.tolist()))
        for cap in sorted({max(1, total // 3), max(1, total // 10), total}):
            if total < 2:
                continue
            want, want_cnt = deterministic_cut(g, rows, cap)
            for tune in (None, abi.tuning(waves_per_cu=1, claim_blocks=1)):
                cnt, grows = dg.enum(mot, delta, cap=cap, tune=tune)
                got = sorted(map(tuple, grows.tolist()))
                assert got == sorted(want), (i, cap)
                assert cnt == want_cnt, (i, cap, cnt, want_cnt)
            beyond += cap < total
        dg.close()
    assert beyond > 10


def test_batch_equals_single(small_cases):
    abi = _abi()
    (i, U, V, T, _, _) = small_cases[5]
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    rng = np.random.default_rng(3)
    qs = [(random_motif(rng, int(rng.integers(2, 7))), int(rng.integers(0, 100))) for _ in range(40)]
    say("batch of", len(qs), "on case", i)
    got = dg.batch_count(qs)
    per_class = dg.batch_count(qs, tune=abi.tuning(batch_mode=1))
    fused = dg.batch_count(qs, tune=abi.tuning(batch_mode=2))
    single = [dg.count(m, d) for (m, d) in qs]
    want = [oracle.count(g, m, d, nthreads=4) for (m, d) in qs]
    dg.close()
    assert got == single == want
    assert per_class == want and fused == want


def test_zipf100k_all36():
    abi = _abi()
    from everest_amd.gen import graphs
    U, V, T = graphs.make("zipf100k")
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    for mot in all_3edge_motifs():
        say("zipf100k motif", mot)
        want = oracle.count(g, mot, 3600, nthreads=8)
        got = dg.count(mot, 3600)
        assert got == want, (mot, got, want)
    dg.close()


def test_reference_work_counters(small_cases):
    """The instrumented kernel's R/X1/X2/P equal the oracle's counters."""
    abi = _abi()
    for (i, U, V, T, mot, delta) in small_cases[:40]:
        g = prep.preprocess(U, V, T)
        say("work case", i, "motif", mot, "delta", delta)
        want_c, want_w = oracle.count(g, mot, delta, nthreads=2, with_work=True)
        dg = abi.DeviceGraph.from_dict(g)
        got_c, got_w = dg.count_work(mot, delta)
        dg.close()
        assert got_c == want_c and got_w == want_w, (i, got_w, want_w)


def test_pl8m_root_subranges():
    """Bit-exact on the 8M-edge power-law graph over seeded root subranges
    (counts are additive over root ranges), all 36 3-edge motifs."""
    abi = _abi()
    from everest_amd import core
    from everest_amd.gen import graphs, queries
    U, V, T = graphs.make("pl8m")
    el = core.data.EdgeListLoader.fromArrays("pl8m", U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    rng = np.random.default_rng(21)
    for edges in queries.all_3edge_motifs():
        mot = queries.as_motif(edges)
        beg = int(rng.integers(0, hg.numEdges - 200_000))
        end = beg + 100_000
        say("pl8m motif", edges, "roots", beg, end)
        want = oracle.count(g, mot, 86400, beg, end, nthreads=8)
        got = dg.count(mot, 86400, beg, end)
        assert got == want, (edges, got, want)
    dg.close()


def test_pl8m_pruned_equals_unpruned():
    """The pruning rules drop only subtrees without matches: on the 8M-edge
    graph, 4- to 6-edge motifs counted by the pruned kernel equal the
    instrumented kernel's counts (which walks the reference's search tree
    unpruned) over 200,000-root ranges, and the oracle on a 20,000-root
    sample of each range.  Five fixed star and fan motifs with a one-week
    window have matches in every range; eight random motifs add shapes."""
    abi = _abi()
    from everest_amd import core
    from everest_amd.gen import graphs, queries
    U, V, T = graphs.make("pl8m")
    el = core.data.EdgeListLoader.fromArrays("pl8m", U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    rng = np.random.default_rng(29)
    cases = [(e, 604800) for e in ([(0, 1), (0, 1), (0, 2), (0, 1)], [(0, 1), (2, 1), (0, 1), (2, 1)],
                                   [(0, 1), (0, 2), (0, 3), (0, 1)], [(0, 1), (0, 1), (0, 1), (0, 2), (0, 1)],
                                   [(0, 1), (0, 2), (0, 1), (0, 2), (0, 1), (0, 2)])]
    for i in range(8):
        cases.append((queries.random_motif(rng, 4 + i % 3, max_nodes=4 if i % 4 else 5),
                      int(rng.choice([3600, 86400, 604800]))))
    nonzero = 0
    for edges, delta in cases:
        mot = queries.as_motif(edges)
        beg = int(rng.integers(0, hg.numEdges - 300_000))
        end = beg + 200_000
        say("pl8m pruned/unpruned", edges, "delta", delta, "roots", beg, end)
        got = dg.count(mot, delta, beg, end)
        want, _ = dg.count_work(mot, delta, beg, end)
        assert got == want, (edges, delta, got, want)
        assert dg.count(mot, delta, beg, beg + 20_000) == oracle.count(g, mot, delta, beg, beg + 20_000, nthreads=8)
        nonzero += got > 0
    say("pl8m pruned/unpruned: cases with matches", nonzero)
    assert nonzero >= 5
    dg.close()


def test_pl64m_full_size_properties():
    """At the benchmark's full size (64M edges, the benchmark's own queries):
    counts are additive over root ranges, the batch equals the single-query
    counts, the pruned counts equal the unpruned instrumented kernel's on a
    1M-root range, and the oracle agrees on a 20,000-root sample."""
    abi = _abi()
    from everest_amd import core
    from everest_amd.gen import graphs, queries
    U, V, T = graphs.make("pl64m")
    el = core.data.EdgeListLoader.fromArrays("pl64m", U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    E = hg.numEdges
    qs = queries.batch_queries(256, seed=4)
    pick = [0, 3, 230, 169, 114, 220, 5, 106]  # benchmark queries of 3-6 edges, each under 0.3 s on the GPU
    motifs = [(queries.as_motif(qs[i][0]), qs[i][1]) for i in pick]
    full = []
    for i, (mot, d) in zip(pick, motifs):
        say("pl64m query", i, qs[i][0], "delta", d)
        c = dg.count(mot, d)
        cuts = [0, E // 3, 2 * E // 3, -1]
        parts = sum(dg.count(mot, d, cuts[k], cuts[k + 1]) for k in range(3))
        assert c == parts, (i, c, parts)
        beg = 20_000_000
        got = dg.count(mot, d, beg, beg + 1_000_000)
        want, _ = dg.count_work(mot, d, beg, beg + 1_000_000)
        assert got == want, (i, got, want)
        full.append(c)
    assert dg.batch_count(motifs) == full
    for i, (mot, d) in zip(pick[-3:], motifs[-3:]):  # samples with matches (10^2 to 10^6)
        beg = 10_000_000
        want = oracle.count(g, mot, d, beg, beg + 20_000, nthreads=8)
        assert want > 0 and dg.count(mot, d, beg, beg + 20_000) == want, i
    say("pl64m counts", full)
    dg.close()


def test_pl8m_batch_classes():
    """A batch mixing every shape class (3..6 edges, up to 4 and up to 8
    nodes) runs as one launch per class, one after another; each
    count equals the single-query count, and a root-range sample of each
    equals the oracle."""
    abi = _abi()
    from everest_amd import core
    from everest_amd.gen import graphs, queries
    U, V, T = graphs.make("pl8m")
    el = core.data.EdgeListLoader.fromArrays("pl8m", U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    rng = np.random.default_rng(23)
    qs = []
    for i in range(40):
        M = 3 + i % 4
        edges = queries.random_motif(rng, M, max_nodes=4 if i % 5 else 6)
        qs.append((queries.as_motif(edges), int(rng.choice([600, 1800, 3600, 7200]))))
    say("pl8m batch of", len(qs), "queries, classes", sorted({(len(m), max(max(a, b) for a, b, _ in m) + 1)
                                                               for m, _ in qs}))
    got = dg.batch_count(qs)
    single = []
    for i, (m, d) in enumerate(qs):
        say("pl8m single", i, m, d)
        single.append(dg.count(m, d))
    assert got == single
    for i, (m, d) in enumerate(qs[:12]):
        beg = int(rng.integers(0, hg.numEdges - 100_000))
        say("pl8m sample", i, beg)
        assert dg.count(m, d, beg, beg + 50_000) == oracle.count(g, m, d, beg, beg + 50_000, nthreads=8)
    dg.close()


def test_server_alone_and_in_mix():
    """Serving: each request's result is the same alone and in a mix of
    concurrently submitted count and enumerate requests (continuous batching
    over several slots), and equals the single-query calls."""
    abi = _abi()
    import threading
    rng = np.random.default_rng(41)
    U, V, T = random_graph(rng, 3000, 60, 4000)
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    reqs = []
    for k in range(60):
        mot = random_motif(rng, int(rng.integers(2, 7)))
        d = int(rng.integers(0, 400))
        cap = int(rng.integers(0, 5000)) if k % 5 == 0 else -1
        reqs.append((mot, d, cap))
    want = []  # (exact count, (count, rows) of tm_enum at the request's cap)
    for mot, d, cap in reqs:
        c = dg.count(mot, d)
        want.append((c, dg.enum(mot, d, cap=cap) if cap >= 0 else None))

    def check(k, r, srv, t):
        mot, d, cap = reqs[k]
        total, rows_want = want[k]
        assert r["status"] == 0, (k, r)
        if cap < 0 or total < cap:
            assert r["count"] == total, (k, r, total)
        else:  # at the cap: the deterministic cut, as tm_enum returns it
            assert r["count"] == rows_want[0], (k, r, rows_want[0])
        if cap < 0:
            srv.release(t)
            return
        rows = srv.take_rows(t, len(mot), cap)
        want_rows = rows_want[1]
        assert sorted(map(tuple, rows.tolist())) == sorted(map(tuple, want_rows.tolist())), k

    say("server: alone")
    with abi.Server(dg, slots=4, max_batch=8) as srv:
        for k, (mot, d, cap) in enumerate(reqs):
            t = srv.submit(mot, d, cap)
            check(k, srv.wait(t), srv, t)
    say("server: mix from 3 threads")
    with abi.Server(dg, slots=3, max_batch=8) as srv:
        tickets = [None] * len(reqs)

        def client(part):
            for k in range(part, len(reqs), 3):
                mot, d, cap = reqs[k]
                tickets[k] = srv.submit(mot, d, cap)

        th = [threading.Thread(target=client, args=(p,)) for p in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        batched = 0
        for k in range(len(reqs)):
            r = srv.wait(tickets[k])
            batched = max(batched, r["batch"])
            check(k, r, srv, tickets[k])
        st = srv.stats()
        say("server stats", st, "largest batch", batched)
        assert st["count_served"] + st["enum_served"] == len(reqs)
    dg.close()


@pytest.mark.skip(reason="graph-captured batches are switched off: on the MI355X two replays passed but an eager "
                         "tm_batch_count after them did not finish; cause not established (tm_batch.hip)")
def test_prepared_batch_graph_replays():
    """A prepared batch (its launches captured into one HIP graph) replayed
    several times, with eager single-query and batch calls on the same slots
    in between: every replay equals the eager counts."""
    abi = _abi()
    from everest_amd.gen import graphs
    U, V, T = graphs.make("zipf100k")
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    rng = np.random.default_rng(43)
    qs = [(random_motif(rng, int(rng.integers(2, 7))), int(rng.integers(600, 20000))) for _ in range(50)]
    want = dg.batch_count(qs)
    say("prepared batch of", len(qs))
    pb = dg.prepare_batch(qs)
    for rep in range(4):
        got = pb.run()
        say("replay", rep)
        assert got == want, rep
        # eager work on the same slots between replays
        m, d = qs[rep]
        assert dg.count(m, d) == want[rep]
        assert dg.batch_count(qs[:7]) == want[:7]
    pb.close()
    assert [oracle.count(g, m, d, nthreads=8) for (m, d) in qs[:10]] == want[:10]
    dg.close()


def test_two_edge_tail_and_meter(small_cases):
    """The two-edge tail (summing the last two edges over the last edge's
    candidates when that list is shorter) forced on every frame it applies
    to, and switched off: counts equal the oracle either way; the metered
    kernel (bytes of the search) counts the same."""
    abi = _abi()
    for (i, U, V, T, mot, delta) in small_cases:
        if len(mot) < 3:
            continue
        g = prep.preprocess(U, V, T)
        say("two-edge tail case", i, "motif", mot, "delta", delta)
        want = oracle.count(g, mot, delta, nthreads=4)
        dg = abi.DeviceGraph.from_dict(g)
        for ts in (1, -1, 0):
            got = dg.count(mot, delta, tune=abi.tuning(tail_swap=ts))
            assert got == want, (i, ts, got, want)
        c, b = dg.count_bytes(mot, delta, tune=abi.tuning(tail_swap=1))
        assert c == want and (b > 0 or want == 0), (i, c, b)
        assert dg.batch_count([(mot, delta)] * 3, tune=abi.tuning(tail_swap=1)) == [want] * 3
        dg.close()


def test_repeated_pair_rewrites():
    """Motifs that repeat a pair (the root's pair at edge 1; a new node's pair
    at edge M-2), on graphs dense in parallel edges: the counting kernel's
    weighted rewrites, with and without survivor compaction, forced
    donation, the two-edge tail and the batch kernel, equal the oracle."""
    abi = _abi()
    rng = np.random.default_rng(17)
    shapes = [[(0, 1), (0, 1), (2, 1), (0, 1), (2, 3)], [(0, 1), (0, 2), (0, 2), (3, 2)],
              [(0, 1), (0, 1), (0, 2), (0, 2)], [(0, 1), (0, 1), (0, 2), (0, 2), (3, 2)],
              [(0, 1), (2, 1), (2, 1), (1, 3)], [(0, 1), (0, 1), (1, 2), (1, 2), (2, 0)],
              [(0, 1), (0, 1), (0, 1), (0, 2)], [(0, 1), (1, 0), (1, 2), (1, 2), (2, 3), (3, 2)],
              [(0, 1), (0, 2), (3, 2), (0, 1)], [(0, 1), (0, 2), (0, 2), (3, 2), (3, 0)]]
    tunes = [None, abi.tuning(sb_flush=-1), abi.tuning(tail_swap=1),
             abi.tuning(steal_after_us=1, steal_every=1, steal_min=1)]
    for case in range(40):
        # Time span a tenth to a third of the edge count: dense in parallel
        # edges (counts up to ~1e8) while the oracle stays near a second for
        # all 40 cases; denser settings put single cases at hours of CPU.
        E = int(rng.integers(200, 1500))
        N = int(rng.integers(3, 8))
        U, V, T = random_graph(rng, E, N, int(rng.integers(E // 10, E // 3)))
        edges = shapes[case % len(shapes)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        delta = int(rng.integers(0, 60))
        g = prep.preprocess(U, V, T)
        say("repeated pairs case", case, "E", E, edges, "delta", delta)
        want = oracle.count(g, mot, delta, nthreads=4)
        dg = abi.DeviceGraph.from_dict(g)
        for t in tunes:
            assert dg.count(mot, delta, tune=t) == want, (case, edges, delta)
        assert dg.batch_count([(mot, delta), (mot, delta + 1)])[0] == want
        c, _ = dg.count_bytes(mot, delta)
        assert c == want
        dg.close()


PAIR_SHAPES = [[(0, 1), (s1, 2) if d1 else (2, s1), (2, 3) if d2 else (3, 2), (0, 1)]
               for s1 in (0, 1) for d1 in (1, 0) for d2 in (0, 1)]


SWEEP_SHAPES = [[(0, 1), (0, 2), (0, 3)], [(0, 1), (2, 1), (3, 1)], [(0, 1), (1, 2), (1, 3)],
                [(0, 1), (2, 0), (3, 0)], [(0, 1), (0, 2), (0, 2)], [(0, 1), (2, 1), (2, 1)],
                [(0, 1), (1, 2), (1, 2)], [(0, 1), (2, 0), (2, 0)]]


def test_window_sweep():
    """3-edge stars and a new node's pair twice on one list of a root node go
    to the window-sweep kernel: on graphs with self-loops, parallel edges and
    tied times, over full and partial root ranges (empty ones too) and a
    delta of 0, its counts equal the oracle and the mining kernel's
    (sweep=-1); a batch mixing swept and mined queries, and the server, give
    the same counts; an enumeration at a cap (whose cut counts root ranges
    through the sweep) is the deterministic cut."""
    abi = _abi()
    rng = np.random.default_rng(23)
    for case in range(48):
        E = int(rng.integers(3, 2500))
        N = int(rng.integers(2, 40))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, max(2, E // 2))))
        g = prep.preprocess(U, V, T)
        edges = SWEEP_SHAPES[case % len(SWEEP_SHAPES)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        delta = int(rng.integers(0, 100)) if case % 6 else 0
        dg = abi.DeviceGraph.from_dict(g)
        nroots = g["E"] - 2
        ranges = [(0, -1), (int(rng.integers(0, nroots)), int(rng.integers(0, nroots + 1))), (nroots // 2, nroots // 2)]
        say("window sweep case", case, "E", g["E"], edges, "delta", delta, "ranges", ranges)
        for beg, end in ranges:
            want = oracle.count(g, mot, delta, beg, end if end >= 0 else nroots, nthreads=4) \
                if (end < 0 or end > beg) else 0
            got = dg.count(mot, delta, beg, end, tune=abi.tuning(sweep=1))
            assert dg.last_stats()["sweep"] == 1
            assert got == want, (case, edges, delta, beg, end, got, want)
            assert dg.count(mot, delta, beg, end, tune=abi.tuning(sweep=-1)) == want
            assert dg.last_stats()["sweep"] == 0
            assert dg.count(mot, delta, beg, end) == want
        want = oracle.count(g, mot, delta, nthreads=4)
        other = [(0, 1), (1, 2), (2, 0)]
        omot = [(a, b, i) for i, (a, b) in enumerate(other)]
        got = dg.batch_count([(mot, delta), (omot, delta), (mot, delta + 7)], tune=abi.tuning(sweep=1))
        assert got[0] == want and got[1] == oracle.count(g, omot, delta, nthreads=4)
        assert got[2] == oracle.count(g, mot, delta + 7, nthreads=4)
        if case % 8 == 0:
            total, rows = oracle.enum(g, mot, delta, 1 << 20)
            if total >= 3:
                cap = max(1, total // 3)
                cut, cut_cnt = deterministic_cut(g, rows, cap)
                cnt, grows = dg.enum(mot, delta, cap=cap, tune=abi.tuning(sweep=1))
                assert sorted(map(tuple, grows.tolist())) == sorted(cut) and cnt == cut_cnt, case
        dg.close()


def test_window_sweep_server():
    """Swept queries (window sweep and root-pair sweep) through the server,
    mixed with mined ones, equal the oracle."""
    abi = _abi()
    rng = np.random.default_rng(29)
    U, V, T = random_graph(rng, 3000, 60, 1500)
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    qs = [([(a, b, i) for i, (a, b) in enumerate(e)], int(rng.integers(0, 200))) for e in SWEEP_SHAPES]
    qs += [([(0, 1, 0), (1, 2, 1), (2, 0, 2)], 150), ([(0, 1, 0), (0, 2, 1), (2, 3, 2), (3, 0, 3)], 150)]
    qs += [([(a, b, i) for i, (a, b) in enumerate(e)], int(rng.integers(0, 200))) for e in PAIR_SHAPES]
    want = [oracle.count(g, m, d, nthreads=4) for m, d in qs]
    say("window sweep server: %d queries" % len(qs))
    srv = abi.Server(dg, max_batch=16, tune=abi.tuning(sweep=1))
    tickets = [srv.submit(m, d, -1) for m, d in qs]
    got = []
    for t in tickets:
        r = srv.wait(t)
        assert r["status"] == 0
        got.append(r["count"])
        srv.release(t)
    srv.close()
    dg.close()
    assert got == want


def test_window_ends_at_last_root_pair_edge():
    """Motifs whose last edge closes the root's two nodes (the root's window
    ends at that pair's last edge in it): counts equal the oracle with the
    two-edge tail forced and off, through the metered kernel and in a batch."""
    abi = _abi()
    rng = np.random.default_rng(73)
    shapes = [[(0, 1), (0, 2), (3, 2), (0, 1)], [(0, 1), (1, 2), (0, 1)], [(0, 1), (2, 0), (1, 0)],
              [(0, 1), (0, 2), (2, 1), (1, 0)], [(0, 1), (2, 1), (2, 3), (0, 1)],
              [(0, 1), (0, 2), (3, 2), (2, 3), (0, 1)]]
    for case in range(36):
        E = int(rng.integers(50, 2500))
        U, V, T = random_graph(rng, E, int(rng.integers(2, 30)), int(rng.integers(E // 8 + 1, E)))
        g = prep.preprocess(U, V, T)
        edges = shapes[case % len(shapes)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        delta = int(rng.integers(0, 200))
        say("last root-pair edge case", case, "E", g["E"], edges, "delta", delta)
        want = oracle.count(g, mot, delta, nthreads=4)
        dg = abi.DeviceGraph.from_dict(g)
        for ts in (1, -1):
            assert dg.count(mot, delta, tune=abi.tuning(tail_swap=ts)) == want, (case, ts)
        c, _ = dg.count_bytes(mot, delta)
        assert c == want
        assert dg.batch_count([(mot, delta), (mot, delta)]) == [want, want]
        dg.close()


def test_pair_sweep():
    """4-edge motifs whose last edge repeats the root's pair (edge 1 a new
    node on a root node's list, edge 2 a new node on its list) go to the
    root-pair sweep kernel: on graphs with self-loops, parallel edges and tied
    times, over full, partial and empty root ranges and a delta of 0, its
    counts equal the oracle and the mining kernel's (sweep=-1); a batch
    mixing swept and mined queries gives the same counts; an enumeration at a
    cap (its cut counts root ranges through the sweep) is the deterministic
    cut."""
    abi = _abi()
    rng = np.random.default_rng(41)
    for case in range(40):
        E = int(rng.integers(4, 2500))
        N = int(rng.integers(2, 30))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, max(2, E // 2))))
        g = prep.preprocess(U, V, T)
        edges = PAIR_SHAPES[case % len(PAIR_SHAPES)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        delta = int(rng.integers(0, 120)) if case % 6 else 0
        dg = abi.DeviceGraph.from_dict(g)
        nroots = g["E"] - 3
        ranges = [(0, -1), (int(rng.integers(0, nroots)), int(rng.integers(0, nroots + 1))), (nroots // 2, nroots // 2)]
        say("pair sweep case", case, "E", g["E"], edges, "delta", delta, "ranges", ranges)
        for beg, end in ranges:
            want = oracle.count(g, mot, delta, beg, end if end >= 0 else nroots, nthreads=4) \
                if (end < 0 or end > beg) else 0
            got = dg.count(mot, delta, beg, end, tune=abi.tuning(sweep=1))
            assert dg.last_stats()["sweep"] == 1
            assert got == want, (case, edges, delta, beg, end, got, want)
            assert dg.count(mot, delta, beg, end, tune=abi.tuning(sweep=-1)) == want
            assert dg.last_stats()["sweep"] == 0
            assert dg.count(mot, delta, beg, end) == want
        want = oracle.count(g, mot, delta, nthreads=4)
        other = [(0, 1), (1, 2), (2, 0)]
        omot = [(a, b, i) for i, (a, b) in enumerate(other)]
        got = dg.batch_count([(mot, delta), (omot, delta), (mot, delta + 9)], tune=abi.tuning(sweep=1))
        assert got[0] == want and got[1] == oracle.count(g, omot, delta, nthreads=4)
        assert got[2] == oracle.count(g, mot, delta + 9, nthreads=4)
        if case % 8 == 0:
            total, rows = oracle.enum(g, mot, delta, 1 << 20)
            if total >= 3:
                cap = max(1, total // 3)
                cut, cut_cnt = deterministic_cut(g, rows, cap)
                cnt, grows = dg.enum(mot, delta, cap=cap, tune=abi.tuning(sweep=1))
                assert sorted(map(tuple, grows.tolist())) == sorted(cut) and cnt == cut_cnt, case
        dg.close()


def test_pair_sweep_pl8m():
    """On the 8M-edge power-law graph the root-pair sweep equals the oracle
    on seeded root subranges and the mining kernel (itself oracle-exact) on
    the full root range, for each shape at deltas of an hour to a week."""
    abi = _abi()
    from everest_amd import core
    from everest_amd.gen import graphs
    U, V, T = graphs.make("pl8m")
    el = core.data.EdgeListLoader.fromArrays("pl8m", U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    rng = np.random.default_rng(43)
    for i, edges in enumerate(PAIR_SHAPES):
        mot = [(a, b, k) for k, (a, b) in enumerate(edges)]
        delta = (3600, 86400, 604800)[i % 3]
        beg = int(rng.integers(0, hg.numEdges - 200_000))
        end = beg + 20_000
        say("pl8m pair sweep", edges, "delta", delta, "roots", beg, end)
        want = oracle.count(g, mot, delta, beg, end, nthreads=8)
        assert dg.count(mot, delta, beg, end, tune=abi.tuning(sweep=1)) == want, (edges, delta, beg)
        got = dg.count(mot, delta, tune=abi.tuning(sweep=1))
        ms_sweep = dg.last_kernel_ms()
        mined = dg.count(mot, delta, tune=abi.tuning(sweep=-1))
        say("  full range: sweep", got, "mined", mined, "ms sweep %.2f mined %.2f" % (ms_sweep, dg.last_kernel_ms()))
        assert got == mined, (edges, delta)
    dg.close()
