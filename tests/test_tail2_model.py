"""CPU check of the two-edge tail: for every frame of the second-to-last
motif edge, summing over the last edge's candidates (the kernel's path when
that list is the shorter) equals summing over the second-to-last edge's
candidates (its leaf-count path), and both equal the oracle in total."""
import numpy as np

import kernel_model as km
from helpers import random_graph, random_motif
from oracle import oracle, prep


def _frames(g, md, delta, visit):
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
    for case in range(120):
        E = int(rng.integers(20, 300))
        N = int(rng.integers(3, 12))
        U, V, T = random_graph(rng, E, N, int(rng.integers(5, 80)))
        M = int(rng.integers(3, 6))
        mot = random_motif(rng, M, max_nodes=int(rng.integers(3, 6)))
        md = km.build_motif([(a, b) for a, b, _ in mot])
        if not km.tail2_eligible(md):
            continue
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
    assert checked > 500
