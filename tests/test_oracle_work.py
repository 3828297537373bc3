"""The oracle's work counters (R roots, X1 / X2 resolveMi calls with io >= 0
/ io == -1, P candidate probes) restated from the match definition instead
of its DFS: every partial match through motif edge l (a row of the oracle's
enumeration of the motif's (l + 1)-edge prefix, or a root with u != v for
l = 0) makes one resolveMi call for edge l + 1 (system/templates/common/
resolveMi.snp:1-29: the base node's list; for an edge between two mapped
nodes the shorter of the destination's in-list and the source's out-list,
ties to the out-list) and scans that list from the first position after its
own (reduceEarlyIdx.snp) until the first edge past the window, that edge
included (nextMatchInTask.snp:3-21): P adds the list's entries after the
position with time <= t0 + delta, plus one when an entry after them exists."""
import numpy as np

from oracle import oracle, prep
from helpers import random_graph
import cons_check as cc

MOTIFS = [[(0, 1), (1, 2), (2, 0)], [(0, 1), (0, 2), (0, 3)], [(0, 1), (1, 0), (0, 1)],
          [(0, 1), (2, 1), (0, 2), (2, 0)], [(0, 1), (1, 2)], [(0, 1), (0, 2), (2, 3), (3, 0)],
          [(0, 1), (2, 0), (1, 2), (1, 0)]]


def restated_work(g, edges, delta, beg, end):
    Eg = np.asarray(g["Eg"])
    T = Eg[:, 2].astype(np.int64)
    outV, outR, inV, inR = g["outV"], g["outR"], g["inV"], g["inR"]
    M = len(edges)
    st, _ = cc.motif_structure(edges)
    w = [max(0, end - beg), 0, 0, 0]
    for lvl in range(M - 1):
        if lvl == 0:
            partial = [([r], [int(Eg[r, 0]), int(Eg[r, 1])]) for r in range(beg, end) if Eg[r, 0] != Eg[r, 1]]
        else:
            pre = edges[:lvl + 1]
            total, rows = oracle.enum(g, [(a, b, i) for i, (a, b) in enumerate(pre)], delta, 1 << 22, beg, end)
            assert total == len(rows)
            partial = cc.match_rows(g, rows.tolist(), pre)
        kind, a, b, _ = st[lvl + 1]
        for pos, m in partial:
            p, tl = pos[-1], int(T[pos[0]]) + delta
            if kind == 0:
                lst = outV[outR[m[a]]:outR[m[a] + 1]]
                w[1] += 1
            elif kind == 1:
                lst = inV[inR[m[b]]:inR[m[b] + 1]]
                w[1] += 1
            else:
                li = inV[inR[m[b]]:inR[m[b] + 1]]
                lo = outV[outR[m[a]]:outR[m[a] + 1]]
                lst = li if len(li) < len(lo) else lo
                w[2] += 1
            after = np.asarray(lst)[np.asarray(lst) > p]
            inside = int((T[after] <= tl).sum())
            w[3] += inside + (1 if inside < len(after) else 0)
    return w


def test_work_counters_equal_the_restatement():
    rng = np.random.default_rng(8)
    nz = 0
    for case in range(42):
        E = int(rng.integers(5, 120))
        U, V, T = random_graph(rng, E, int(rng.integers(3, 10)), int(rng.integers(5, 60)))
        g = prep.preprocess(U, V, T)
        edges = MOTIFS[case % len(MOTIFS)]
        M = len(edges)
        d = int(rng.integers(0, 40))
        beg, end = (0, g["E"] - M + 1) if case % 2 else (int(rng.integers(0, g["E"] - M + 1)), g["E"] - M + 1)
        _, work = oracle.count(g, [(a, b, i) for i, (a, b) in enumerate(edges)], d, beg, end, nthreads=1,
                               with_work=True)
        want = restated_work(g, edges, d, beg, end)
        assert [int(x) for x in work] == want, (case, edges, d, beg, end, list(work), want)
        nz += want[2] > 0
    assert nz > 5
