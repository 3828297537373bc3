"""The CPU model of the gfx950 kernel (tests/kernel_model.py) against the
oracle: same counts and enumerated row sets, every index in bounds, frame stack
within capacity, for all start rules, scan thresholds and chunk sizes."""
import numpy as np

import kernel_model as km
from oracle import oracle, prep
from helpers import random_graph, random_motif


def test_model_matches_oracle():
    rng = np.random.default_rng(7)
    for _ in range(150):
        E = int(rng.integers(1, 250))
        U, V, T = random_graph(rng, E, int(rng.integers(2, 12)), int(rng.integers(1, 60)))
        M = int(rng.integers(2, 7))
        mot = random_motif(rng, M)
        delta = int(rng.integers(0, 30))
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        md = km.build_motif([(a, b) for a, b, _ in mot])
        want = oracle.count(g, mot, delta, nthreads=1)
        scan = int(rng.choice([0, 2, 16]))
        got, _, st = km.run_wave(gv, md, delta, 0, g["E"] - M + 1, scan_max=scan,
                                 chunk_roots=int(rng.choice([64, 128, 256])),
                                 sb_flush=int(rng.choice([0, 1, 32, 64])), sb_inline=int(rng.choice([1, 8, 32, 64])))
        assert got == want
        if want < 3000:
            total, rows = oracle.enum(g, mot, delta, 100000)
            c2, rows2, _ = km.run_wave(gv, md, delta, 0, g["E"] - M + 1, enumerate_=True)
            assert c2 == total
            assert sorted(rows2) == sorted(map(tuple, rows.tolist()))


def test_model_root_ranges_are_additive():
    rng = np.random.default_rng(8)
    U, V, T = random_graph(rng, 400, 10, 100)
    g = prep.preprocess(U, V, T)
    gv = km.build_view(g)
    for _ in range(10):
        mot = random_motif(rng, 4)
        md = km.build_motif([(a, b) for a, b, _ in mot])
        cut = int(rng.integers(0, 397))
        a, _, _ = km.run_wave(gv, md, 40, 0, cut)
        b, _, _ = km.run_wave(gv, md, 40, cut, 397)
        assert a + b == oracle.count(g, mot, 40)


def test_model_reference_work_counters_match_oracle():
    """The instrumented variant's R/X1/X2/P equal the C oracle's counters."""
    rng = np.random.default_rng(12)
    for _ in range(80):
        U, V, T = random_graph(rng, int(rng.integers(1, 200)), int(rng.integers(2, 10)), int(rng.integers(1, 40)))
        M = int(rng.integers(2, 6))
        mot = random_motif(rng, M)
        delta = int(rng.integers(0, 25))
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        md = km.build_motif([(a, b) for a, b, _ in mot])
        want_c, want_w = oracle.count(g, mot, delta, nthreads=1, with_work=True)
        got, _, st = km.run_wave(gv, md, delta, 0, g["E"] - M + 1, instr=True)
        assert got == want_c
        assert st["ref"] == want_w


def test_model_work_donation_between_wavefronts():
    """Several simulated wavefronts with frame splitting and a shared pool count
    exactly what the oracle counts (no frame lost or processed twice)."""
    rng = np.random.default_rng(31)
    donated = 0
    for _ in range(40):
        U, V, T = random_graph(rng, int(rng.integers(50, 400)), int(rng.integers(2, 8)), int(rng.integers(1, 30)))
        M = int(rng.integers(3, 6))
        mot = random_motif(rng, M)
        delta = int(rng.integers(5, 40))
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        md = km.build_motif([(a, b) for a, b, _ in mot])
        want = oracle.count(g, mot, delta, nthreads=1)
        got, nd = km.run_waves(gv, md, delta, 0, g["E"] - M + 1, nwaves=int(rng.integers(2, 6)),
                               steal_every=int(rng.integers(1, 4)), steal_max=int(rng.integers(1, 5)),
                               pool_cap=int(rng.choice([1, 2, 4, 1 << 20])), seed=int(rng.integers(0, 1 << 30)),
                               idle_cap=int(rng.choice([1, 2, 1 << 30])))
        assert got == want
        donated += nd
    assert donated > 0


def test_model_batch_dispatch_covers_every_match_once():
    """The multi-query root dispatch (one space of 256-root blocks, query
    after query, in fixed-size claims; the gap filter) expands
    each (query, root) pair
    at most once and never drops a root that has a match: per-query sums of
    single-root oracle counts over the dispatched roots equal the full counts."""
    rng = np.random.default_rng(31)
    for it in range(12):
        E = int(rng.integers(1, 700))
        U, V, T = random_graph(rng, E, int(rng.integers(2, 14)), int(rng.integers(1, 300)))
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        E = g["E"]
        qs, mots = [], []
        for _ in range(int(rng.integers(1, 9))):
            M = int(rng.integers(2, 6))
            mot = random_motif(rng, M)
            d = int(rng.integers(-2, 60))
            b = int(rng.integers(0, E + 1))
            e = int(rng.integers(b, E + 1)) if rng.random() < 0.7 else E
            qs.append((km.build_motif([(a, c) for a, c, _ in mot]), d, b, e))
            mots.append(mot)
        pairs = km.batch_dispatch(gv, qs, nwaves=int(rng.integers(1, 12)), per_claim=int(rng.integers(1, 6)),
                                  seed=it)
        assert len(pairs) == len(set(pairs))
        for qi, (md, d, b, e) in enumerate(qs):
            roots = [r for (q, r) in pairs if q == qi]
            assert all(b <= r < e for r in roots)
            got = sum(oracle.count(g, mots[qi], d, r, r + 1, nthreads=1) for r in roots)
            assert got == oracle.count(g, mots[qi], d, b, e, nthreads=1), (it, qi)
        # instrumented launches expand every in-range root of every query
        ipairs = km.batch_dispatch(gv, qs, instr=True)
        for qi, (md, d, b, e) in enumerate(qs):
            assert sorted(r for (q, r) in ipairs if q == qi) == list(range(b, e))


def test_model_two_edge_tail_swap():
    """Frames of the second-to-last edge summed over the last edge's
    candidates (swapped when that list is shorter): counts equal the oracle
    with and without survivor compaction, at several thresholds, and the
    side table entry of a swapped frame is never shared by a live one."""
    rng = np.random.default_rng(9)
    km.SWAPS[0] = 0
    for _ in range(60):
        E = int(rng.integers(20, 300))
        U, V, T = random_graph(rng, E, int(rng.integers(3, 10)), int(rng.integers(2, 60)))
        M = int(rng.integers(3, 7))
        mot = random_motif(rng, M, max_nodes=int(rng.integers(3, 6)))
        md = km.build_motif([(a, b) for a, b, _ in mot])
        delta = int(rng.integers(0, 50))
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        want = oracle.count(g, mot, delta, nthreads=1)
        for t2 in (1, 2, 8):
            got, _, _ = km.run_wave(gv, md, delta, 0, g["E"] - M + 1, t2_min=t2,
                                    sb_flush=int(rng.choice([0, 1, 32])), sb_inline=int(rng.choice([1, 32])))
            assert got == want, (mot, delta, t2)
    assert km.SWAPS[0] > 100


def test_model_concurrent_persistent_launches():
    """Persistent launches running at the same time, each with a grid the size
    of the whole GPU: with every idle wavefront waiting for donations (no
    bound), the waiting wavefronts of the launches can fill every slot while
    each launch still has wavefronts that never started, which hold its busy
    count above zero: the GPU hangs (the first captured batch's failure).
    With at most idle_cap waiters per launch and launches * idle_cap below
    the GPU's slots, every interleaving completes."""
    rng = np.random.default_rng(3)
    cap = 64
    hangs = 0
    for trial in range(40):
        K = int(rng.integers(2, 5))
        grids = [cap] * K
        work = [[int(x) for x in rng.integers(0, 6, cap)] for _ in range(K)]
        hangs += km.run_concurrent_launches(grids, cap, 1 << 30, work, rng) == "deadlock"
        assert km.run_concurrent_launches(grids, cap, cap // (2 * K), work, rng) == "done"
    assert hangs > 0
    # one launch alone never hangs, bounded or not: its whole grid is resident
    for trial in range(20):
        work = [[int(x) for x in rng.integers(0, 6, cap)]]
        assert km.run_concurrent_launches([cap], cap, 1 << 30, work, rng) == "done"


def test_model_window_ends_at_last_root_pair_edge():
    """A motif whose last edge closes the root's two nodes: the root's window
    ends at that pair's last edge inside it (positions and pruning time);
    the counts stay the oracle's, with and without the two-edge tail."""
    rng = np.random.default_rng(71)
    shapes = [[(0, 1), (0, 2), (3, 2), (0, 1)], [(0, 1), (1, 2), (0, 1)], [(0, 1), (2, 0), (1, 0)],
              [(0, 1), (0, 2), (2, 1), (1, 0)], [(0, 1), (2, 1), (2, 3), (0, 1)]]
    nz = 0
    for case in range(30):
        E = int(rng.integers(20, 220))
        U, V, T = random_graph(rng, E, int(rng.integers(2, 10)), int(rng.integers(5, 120)))
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        edges = shapes[case % len(shapes)]
        md = km.build_motif(edges)
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        M = len(edges)
        delta = int(rng.integers(0, 60))
        want = oracle.count(g, mot, delta, nthreads=1)
        nz += want > 0
        for t2 in (None, 1):
            got, _, _ = km.run_wave(gv, md, delta, 0, g["E"] - M + 1, t2_min=t2)
            assert got == want, (case, edges, delta, t2, got, want)
    assert nz > 10


def test_model_donor_give_up_aborts_every_wavefront():
    """A donor whose bounded wait on its ring slot gives up raises error bit 8
    and the abort flag, and every wavefront then leaves (no wavefront keeps
    waiting for the record that will never be written); runs that do not
    give up count exactly."""
    rng = np.random.default_rng(37)
    gave_up = exact = 0
    for _ in range(30):
        U, V, T = random_graph(rng, int(rng.integers(80, 300)), int(rng.integers(2, 6)), int(rng.integers(1, 20)))
        M = int(rng.integers(3, 6))
        mot = random_motif(rng, M)
        delta = int(rng.integers(5, 40))
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        md = km.build_motif([(a, b) for a, b, _ in mot])
        got, _, err = km.run_waves(gv, md, delta, 0, g["E"] - M + 1, nwaves=int(rng.integers(3, 7)),
                                   steal_every=1, steal_max=4, pool_cap=1, seed=int(rng.integers(0, 1 << 30)),
                                   spin_limit=int(rng.choice([0, 1, 3])), with_error=True)
        if err:
            assert err == 8
            gave_up += 1
        else:
            assert got == oracle.count(g, mot, delta, nthreads=1)
            exact += 1
    assert gave_up > 0 and exact > 0


def test_model_ring_across_consecutive_launches():
    """Consecutive launches share the donation ring and its release words, as
    on the device: records and releases are tagged with the launch
    generation, the host clears the ring before the generation wraps (2-bit
    generations here, so it wraps every 3 launches), and every launch counts
    exactly with small rings that force many laps."""
    rng = np.random.default_rng(41)
    U, V, T = random_graph(rng, 300, 5, 25)
    g = prep.preprocess(U, V, T)
    gv = km.build_view(g)
    for cap in (1, 2, 4):
        ring = km.Ring(cap, gen_bits=2)
        for launch in range(10):
            M = int(rng.integers(3, 6))
            mot = random_motif(rng, M)
            md = km.build_motif([(a, b) for a, b, _ in mot])
            delta = int(rng.integers(5, 40))
            got, _ = km.run_waves(gv, md, delta, 0, g["E"] - M + 1, nwaves=int(rng.integers(2, 6)),
                                  steal_every=1, steal_max=int(rng.integers(1, 5)),
                                  seed=int(rng.integers(0, 1 << 30)), ring=ring)
            assert got == oracle.count(g, mot, delta, nthreads=1), (cap, launch)
