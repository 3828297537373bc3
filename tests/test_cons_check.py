"""The constraint checker (tests/cons_check.py) against the oracle: no
constraints gives the oracle's count; a uniform node or edge predicate gives
the oracle's count on the edge list restricted to the edges it keeps;
per-node and per-edge predicates, temporal gaps and filters give the number
of the oracle's enumerated rows that satisfy them."""
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
        for pos, m in cc.match_rows(g, rows.tolist(), edges):
            t = [int(g["Eg"][p, 2]) for p in pos]
            if any(t[j + 1] > t[j] + gaps[j] for j in range(M - 1)):
                continue
            if not (VF[m[1]] < 6 and EF[pos[M - 1]] % 3 != 0):
                continue
            if not filt([int(EF[p]) for p in pos], [int(VF[x]) for x in m]):
                continue
            ok += 1
        got = cc.count_constrained(g, edges, d, EF, VF, node_pred=npred, edge_pred=epred, temporal=gaps, filt=filt)
        assert got == ok, (case, edges, d, gaps)
        nz += ok > 0
    assert nz > 10
