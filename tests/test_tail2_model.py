"""CPU check of the two-edge tail: for every frame of the second-to-last
motif edge, summing over the last edge's candidates (the kernel's path when
that list is the shorter) equals summing over the second-to-last edge's
candidates (its leaf-count path), and both equal the oracle in total."""
import numpy as np

import kernel_model as km
from helpers import random_graph, random_motif
from oracle import oracle, prep


def _frames(g, md, delta, visit, with_p=False):
    """Walk every partial match through edge M-3 (roots in order, full DFS
    without pruning) and call visit(m, A, pend) for edge M-2's list."""
    M = md["M"]
    Eg = g["Eg"]
    for r in range(g["E"] - M + 1):
        u, v, t = (int(x) for x in Eg[r])
        if u == v:
            continue
        m = [0] * 8
        m[0], m[1] = u, v
        pend = km.pend_of(g, t + delta)

        def rec(j, m, p):
            L = km.cand_start(g, md, j, m, p)
            pos, nbr, lo, hi, d = L
            hi = km.gallop_lb(pos, lo, hi, pend)
            if hi <= lo:
                return
            if j == M - 2:
                if with_p:
                    visit(m, (pos, nbr, lo, hi, d), pend, p)
                else:
                    visit(m, (pos, nbr, lo, hi, d), pend)
                return
            for s in range(lo, hi):
                mm = list(m)
                if md["kind"][j] == km.K_NEW_DST:
                    if nbr[s] in m[:md["mapped"][j]]:
                        continue
                    mm[md["b"][j]] = nbr[s]
                elif md["kind"][j] == km.K_NEW_SRC:
                    if nbr[s] in m[:md["mapped"][j]]:
                        continue
                    mm[md["a"][j]] = nbr[s]
                rec(j + 1, mm, pos[s])

        rec(1, m, r)


def test_tail2_swapped_sum_equals_direct():
    rng = np.random.default_rng(5)
    checked = 0
    modes = {km.T2_INDEP: 0, km.T2_JOIN: 0}
    joins = [[(0, 1), (2, 1), (2, 0)], [(0, 1), (0, 2), (2, 1)], [(0, 1), (1, 2), (0, 2), (3, 2), (3, 0)],
             [(0, 1), (0, 2), (3, 2), (1, 3)]]
    for case in range(160):
        E = int(rng.integers(20, 300))
        N = int(rng.integers(3, 12))
        U, V, T = random_graph(rng, E, N, int(rng.integers(5, 80)))
        M = int(rng.integers(3, 6))
        mot = random_motif(rng, M, max_nodes=int(rng.integers(3, 6)))
        if case % 4 == 0:
            mot = [(a, b, i) for i, (a, b) in enumerate(joins[(case // 4) % len(joins)])]
        md = km.build_motif([(a, b) for a, b, _ in mot])
        if not km.tail2_eligible(md):
            continue
        modes[km.tail2_mode(md)] += 1
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        delta = int(rng.integers(0, 60))
        tot = [0, 0]

        def visit(m, A, pend):
            a = km.tail2_count(gv, md, m, A, pend)
            b = km.tail2_direct(gv, md, m, A, pend)
            assert a == b, (case, mot, delta, m, a, b)
            tot[0] += a
            tot[1] += 1

        _frames(gv, md, delta, visit)
        assert tot[0] == oracle.count(g, mot, delta), (case, mot)
        checked += tot[1]
    assert checked > 500 and min(modes.values()) > 10, (checked, modes)


def test_tail_patterns_equal_direct():
    """Last two edges on one list (same pair twice, star, new node then its
    pair again): the closed form per frame plus one same-pair term per slot
    equals the leaf-count sum, and the total equals the oracle."""
    rng = np.random.default_rng(8)
    seen = {km.TP_STAR: 0, km.TP_PAIR: 0, km.TP_SAME: 0}
    shapes = [[(0, 1), (2, 1), (3, 1)], [(0, 1), (0, 2), (0, 3)], [(0, 1), (0, 2), (0, 2)],
              [(0, 1), (2, 1), (2, 1)], [(0, 1), (0, 1), (0, 1)], [(0, 1), (1, 2), (1, 2), (1, 2)],
              [(0, 1), (0, 1), (0, 2), (0, 2)], [(0, 1), (1, 2), (3, 2), (4, 2)], [(0, 1), (1, 0), (0, 2), (0, 3)]]
    for case in range(90):
        E = int(rng.integers(30, 300))
        N = int(rng.integers(3, 9))
        U, V, T = random_graph(rng, E, N, int(rng.integers(5, 60)))
        if case % 3 == 0:
            U[: E // 10] = V[: E // 10]  # some self-loops
        edges = shapes[case % len(shapes)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        md = km.build_motif(edges)
        pat = km.tail_pattern(md)
        assert pat != km.TP_NONE
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        delta = int(rng.integers(0, 40))
        tot = [0]

        def visit(m, A, pend, p):
            a = km.tail_pattern_count(gv, md, m, A, p, pend)
            b = km.tail2_direct(gv, md, m, A, pend)
            assert a == b, (case, edges, delta, a, b)
            tot[0] += a
            seen[pat] += 1

        _frames(gv, md, delta, visit, with_p=True)
        assert tot[0] == oracle.count(g, mot, delta), (case, edges)
    assert min(seen.values()) > 50, seen
