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


def _count_rewritten(gv, md, delta, root_pair=True, pair_tail=True):
    """The count with the kernel's two rewrites of repeated pair edges (a
    restatement of their algebra, no pruning): root pair (edge 1 repeats the
    root's pair: edge 2 expanded once from the first parallel edge, each
    candidate weighted by its rank among them) and pair tail (edge M-3 brings
    a node, edge M-2 repeats its pair: only the first edge of each pair after
    the frame's start expands, edge M-2's candidates weighted by their offset
    in the pair list)."""
    M = md["M"]
    Eg = gv["Eg"]
    rp = root_pair and M >= 4 and md["kind"][1] == km.K_CLOSE and md["smode"][1] == km.S_PAIR_SELF
    pt = -1
    if pair_tail and M >= 4 and md["kind"][M - 3] != km.K_CLOSE and md["kind"][M - 2] == km.K_CLOSE \
            and md["smode"][M - 2] == km.S_PAIR_SELF and not (M - 3 == 2 and rp):
        pt = M - 3
    total = 0

    def expand(j, m, p, pend, weight, start):
        """Matches of edges j.. given the map m and previous position p."""
        nonlocal total
        L = km.cand_start(gv, md, j, m, p)
        pos, nbr, lo, hi, d = L
        hi = km.gallop_lb(pos, lo, hi, pend)
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
            w = weight
            if j == pt:
                # only the first edge of its pair after the frame's start
                q = pos[s]
                u, v = (int(x) for x in Eg[q][:2])
                earlier = [x for x in range(start + 1, q) if int(Eg[x][0]) == u and int(Eg[x][1]) == v]
                if earlier:
                    continue
                L2 = km.cand_start(gv, md, j + 1, mm, q)
                p2, _, lo2, hi2, _ = L2
                hi2 = km.gallop_lb(p2, lo2, hi2, pend)
                for s2 in range(lo2, hi2):
                    if j + 1 == M - 1:
                        total += w * (s2 - lo2 + 1)
                    else:
                        expand(j + 2, mm, p2[s2], pend, w * (s2 - lo2 + 1), p2[s2])
                continue
            if j == M - 1:
                total += w
            else:
                expand(j + 1, mm, pos[s], pend, w, pos[s])

    for r in range(gv["E"] - M + 1):
        u, v, t = (int(x) for x in Eg[r])
        if u == v:
            continue
        m = [0] * 8
        m[0], m[1] = u, v
        pend = km.pend_of(gv, t + delta)
        if not rp:
            expand(1, m, r, pend, 1, r)
            continue
        L1 = km.cand_start(gv, md, 1, m, r)
        pos1, _, lo1, hi1, _ = L1
        hi1 = km.gallop_lb(pos1, lo1, hi1, pend)
        if hi1 <= lo1:
            continue
        A1 = [pos1[s] for s in range(lo1, hi1)]
        L2 = km.cand_start(gv, md, 2, m, A1[0])
        pos2, nbr2, lo2, hi2, _ = L2
        hi2 = km.gallop_lb(pos2, lo2, hi2, pend)
        for s in range(lo2, hi2):
            rank = sum(1 for a in A1 if a < pos2[s])
            mm = list(m)
            if md["kind"][2] == km.K_NEW_DST:
                if nbr2[s] in m[:md["mapped"][2]]:
                    continue
                mm[md["b"][2]] = nbr2[s]
            elif md["kind"][2] == km.K_NEW_SRC:
                if nbr2[s] in m[:md["mapped"][2]]:
                    continue
                mm[md["a"][2]] = nbr2[s]
            if 2 == M - 1:
                total += rank
            elif 2 == pt:
                raise AssertionError("pair tail at edge 2 under the root-pair rewrite")
            else:
                expand(3, mm, pos2[s], pend, rank, pos2[s])
    return total


def test_repeated_pair_rewrites_equal_oracle():
    """The root-pair and pair-tail rewrites (weights instead of repeated
    subtrees) count exactly what the oracle counts, on graphs dense in
    parallel edges."""
    rng = np.random.default_rng(13)
    shapes = [[(0, 1), (0, 1), (2, 1), (0, 1), (2, 3)], [(0, 1), (0, 2), (0, 2), (3, 2)],
              [(0, 1), (0, 1), (0, 2), (0, 2)], [(0, 1), (0, 1), (0, 2), (0, 2), (3, 2)],
              [(0, 1), (2, 1), (2, 1), (1, 3)], [(0, 1), (0, 1), (1, 2), (1, 2), (2, 0)],
              [(0, 1), (0, 1), (0, 1), (0, 2)], [(0, 1), (1, 0), (1, 2), (1, 2), (2, 3), (3, 2)]]
    for case in range(48):
        E = int(rng.integers(20, 160))
        N = int(rng.integers(3, 6))
        U, V, T = random_graph(rng, E, N, int(rng.integers(5, 40)))
        edges = shapes[case % len(shapes)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        md = km.build_motif(edges)
        g = prep.preprocess(U, V, T)
        gv = km.build_view(g)
        delta = int(rng.integers(0, 30))
        want = oracle.count(g, mot, delta)
        assert _count_rewritten(gv, md, delta) == want, (case, edges, delta)
