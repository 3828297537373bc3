"""The constraint checker (tests/cons_check.py) against the oracle: no
constraints gives the oracle's count; a uniform node or edge predicate gives
the oracle's count on the edge list restricted to the edges it keeps;
per-node and per-edge predicates, temporal gaps and filters give the
oracle's enumerated rows that satisfy them (their number and the rows
themselves); anti-edges give the oracle's rows of the motif without them
that have no witness edge."""
import numpy as np

from oracle import oracle, prep
from helpers import random_graph
import cons_check as cc

SHAPES = [[(0, 1), (1, 2), (2, 0)], [(0, 1), (0, 2), (0, 3)], [(0, 1), (1, 0), (0, 1)],
          [(0, 1), (2, 1), (0, 2), (2, 0)], [(0, 1), (1, 2)], [(0, 1), (0, 2), (2, 3), (3, 0)]]


def _case(rng, E=None):
    E = E or int(rng.integers(20, 160))
    U, V, T = random_graph(rng, E, int(rng.integers(3, 12)), int(rng.integers(5, 80)))
    g = prep.preprocess(U, V, T)
    EF = rng.integers(0, 10, g["E"])
    VF = rng.integers(0, 10, g["V"])
    return U, V, T, g, EF, VF


def test_no_constraints_is_the_oracle():
    rng = np.random.default_rng(1)
    for case in range(24):
        U, V, T, g, EF, VF = _case(rng)
        edges = SHAPES[case % len(SHAPES)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        d = int(rng.integers(0, 50))
        assert cc.count_constrained(g, edges, d, EF, VF) == oracle.count(g, mot, d, nthreads=1)


def test_uniform_predicates_are_restricted_graphs():
    rng = np.random.default_rng(2)
    for case in range(30):
        U, V, T, g, EF, VF = _case(rng)
        edges = SHAPES[case % len(SHAPES)]
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        d = int(rng.integers(0, 50))
        k = int(rng.integers(2, 5))
        # labels by raw vertex id and by original edge index, carried to the restricted list
        raw = g["rename"]
        vlab = {int(raw[i]): int(VF[i]) for i in range(g["V"])}
        elab = np.empty(g["E"], dtype=np.int64)
        elab[np.asarray(g["Euid"])] = EF  # label of original edge i
        if case % 2 == 0:
            keep = np.array([vlab[int(u)] % k == 0 and vlab[int(v)] % k == 0 for u, v in zip(U, V)], dtype=bool)
            got = cc.count_constrained(g, edges, d, EF, VF, node_pred={i: (lambda f: f % k == 0) for i in range(8)})
        else:
            keep = elab % k == 0
            got = cc.count_constrained(g, edges, d, EF, VF, edge_pred={j: (lambda f: f % k == 0) for j in range(8)})
        if keep.sum() < len(edges):
            want = 0
        else:
            want = oracle.count(prep.preprocess(U[keep], V[keep], T[keep]), mot, d, nthreads=1)
        assert got == want, (case, edges, d, k)


def test_per_item_predicates_temporal_and_filter_are_filtered_rows():
    rng = np.random.default_rng(3)
    nz = 0
    for case in range(40):
        U, V, T, g, EF, VF = _case(rng, int(rng.integers(20, 90)))
        edges = SHAPES[case % len(SHAPES)]
        M = len(edges)
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        d = int(rng.integers(5, 60))
        gaps = [int(rng.integers(0, d + 1)) for _ in range(M - 1)]
        npred = {1: lambda f: f < 6}
        epred = {M - 1: lambda f: f % 3 != 0}
        filt = lambda el, nl: (el[0] + nl[0]) % 2 == 0
        total, rows = oracle.enum(g, mot, d, 1 << 20)
        ok = 0
        want_rows = []
        for pos, m in cc.match_rows(g, rows.tolist(), edges):
            t = [int(g["Eg"][p, 2]) for p in pos]
            if any(t[j + 1] > t[j] + gaps[j] for j in range(M - 1)):
                continue
            if not (VF[m[1]] < 6 and EF[pos[M - 1]] % 3 != 0):
                continue
            if not filt([int(EF[p]) for p in pos], [int(VF[x]) for x in m]):
                continue
            ok += 1
            want_rows.append(tuple(pos))
        got_rows = []
        got = cc.count_constrained(g, edges, d, EF, VF, node_pred=npred, edge_pred=epred, temporal=gaps, filt=filt,
                                   out_rows=got_rows)
        assert got == ok == len(got_rows), (case, edges, d, gaps)
        assert sorted(got_rows) == sorted(want_rows), (case, edges, d, gaps)
        nz += ok > 0
    assert nz > 10


# (motif, anti flags): the reference's example M4-A with its anti edge, an anti
# edge last, two anti edges in a row, anti edges in the middle and last, a
# repeated pair
ANTI = [([(0, 1), (1, 2), (2, 0), (2, 3), (3, 0)], [False, False, True, False, False]),
        ([(0, 1), (1, 2), (2, 0)], [False, False, True]),
        ([(0, 1), (1, 2), (2, 0), (1, 0), (2, 3)], [False, False, True, True, False]),
        ([(0, 1), (0, 2), (1, 2), (2, 3), (0, 3)], [False, False, True, False, True]),
        ([(0, 1), (1, 0), (0, 1)], [False, False, True]),
        ([(0, 1), (1, 0)], [False, True])]


def _rows_without_witness(g, edges, anti, d, EF, VF, gaps=None, epred=None, npred=None):
    """Anti-edge semantics restated over sets: the oracle enumerates the motif
    without its anti edges (roots below E - M + 1 of the whole motif); a row
    counts when, for every anti edge j, no edge of the graph from m[a] to m[b]
    lies after the position of the nearest earlier non-anti edge and within
    edge j's window (gap from that edge, at most delta after the root)."""
    M = len(edges)
    keep = [j for j in range(M) if not anti[j]]
    red = [edges[j] for j in keep]
    _, prev = cc.anti_prev(anti, M)
    Eg = np.asarray(g["Eg"])
    E = g["E"]
    if len(red) == 1:  # the root alone: every edge that is not a self-loop
        rows = np.asarray(g["Euid"])[np.nonzero(Eg[:, 0] != Eg[:, 1])[0]].reshape(-1, 1)
    else:
        total, rows = oracle.enum(g, [(a, b, i) for i, (a, b) in enumerate(red)], d, 1 << 20)
    epred, npred = epred or {}, npred or {}
    out = 0
    for rpos, m in cc.match_rows(g, rows.tolist(), red):
        pos = [None] * M
        for i, j in enumerate(keep):
            pos[j] = rpos[i]
        if pos[0] >= E - M + 1:
            continue
        t0 = int(Eg[pos[0], 2])
        def tl(j):
            tp = int(Eg[pos[prev[j]], 2])
            if gaps is None:
                return t0 + d
            return tp + gaps[0] if j == 1 else min(tp + gaps[j - 1], t0 + d)
        ok = all(int(Eg[pos[j], 2]) <= tl(j) for j in keep[1:])
        ok = ok and all(epred[j](int(EF[pos[j]])) for j in keep if j in epred)
        ok = ok and all(npred[i](int(VF[x])) for i, x in enumerate(m) if i in npred)
        for j in range(1, M):
            if not ok or not anti[j]:
                continue
            a, b = edges[j]
            w = np.nonzero((Eg[:, 0] == m[a]) & (Eg[:, 1] == m[b]) & (np.arange(E) > pos[prev[j]])
                           & (Eg[:, 2] <= tl(j)))[0]
            if j in epred:
                w = [q for q in w if epred[j](int(EF[q]))]
            ok = len(w) == 0
        out += ok
    return out


def test_anti_edges_are_rows_without_a_witness():
    rng = np.random.default_rng(4)
    nz = 0
    for case in range(48):
        U, V, T, g, EF, VF = _case(rng, int(rng.integers(20, 120)))
        edges, anti = ANTI[case % len(ANTI)]
        M = len(edges)
        d = int(rng.integers(5, 80))
        gaps = [int(rng.integers(0, d + 1)) for _ in range(M - 1)] if case % 3 == 1 else None
        epred = {j: (lambda f: f % 3 != 0) for j in range(M)} if case % 4 == 2 else None
        npred = {1: lambda f: f < 7} if case % 5 == 3 else None
        want = _rows_without_witness(g, edges, anti, d, EF, VF, gaps, epred, npred)
        got = cc.count_constrained(g, edges, d, EF, VF, node_pred=npred, edge_pred=epred, temporal=gaps, anti=anti)
        assert got == want, (case, edges, anti, d, gaps)
        nz += want > 0
    assert nz > 20


def test_anti_edges_partition_the_motif_without_them():
    """Closed form for one anti edge last: the rows of the motif without it
    split into those with a witness and those without, and those with a
    witness at its first candidate are matches of the whole motif, so
    count(without) - count(anti) counts distinct (row, first witness) pairs:
    at most count(whole motif), equal to it when no row has two witnesses."""
    rng = np.random.default_rng(5)
    for case in range(16):
        U, V, T, g, EF, VF = _case(rng, int(rng.integers(20, 100)))
        edges = [(0, 1), (1, 2), (2, 0)]
        d = int(rng.integers(5, 80))
        E = g["E"]
        red = oracle.count(g, [(0, 1, 0), (1, 2, 1)], d, end=E - 2, nthreads=1)
        whole = oracle.count(g, [(a, b, i) for i, (a, b) in enumerate(edges)], d, nthreads=1)
        got = cc.count_constrained(g, edges, d, EF, VF, anti=[False, False, True])
        assert 0 <= red - got <= whole, (case, d, red, got, whole)
        if whole == 0:
            assert got == red
