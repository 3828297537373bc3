"""Brute-force matcher from the match definition (test infrastructure only).

Written from the definition, not from the reference DFS: a match of an M-edge
motif (renamed node ids, time-ordered edges) is a strictly increasing tuple of
positions p_0 < ... < p_{M-1} in the stably time-sorted edge list such that an
injective node map phi exists with src(p_k) = phi(a_k), dst(p_k) = phi(b_k),
and t(p_{M-1}) - t(p_0) <= delta (inclusive; equal times allowed).  Roots are
p_0 in [beg, end).  Enumerated rows are original line numbers (Euid) of the
matched edges in motif order (system/tmine/KernelGen.py:509-525).

Use only on tiny inputs (E up to ~40 for M<=4): it walks all combinations.
"""
import itertools

import numpy as np


def _sorted_edges(U, V, T):
    order = np.argsort(np.asarray(T), kind="stable")
    return [(int(U[i]), int(V[i]), int(T[i]), int(i)) for i in order]


def _matches(edges, motif, delta, beg, end):
    M = len(motif)
    E = len(edges)
    if end is None or end < 0:
        end = E - M + 1
    for combo in itertools.combinations(range(E), M):
        if not (beg <= combo[0] < end):
            continue
        if edges[combo[-1]][2] - edges[combo[0]][2] > delta:
            continue
        phi = {}
        used = set()
        ok = True
        for k, p in enumerate(combo):
            a, b, _ = motif[k]
            u, v = edges[p][0], edges[p][1]
            for mn, gn in ((a, u), (b, v)):
                if mn in phi:
                    if phi[mn] != gn:
                        ok = False
                        break
                else:
                    if gn in used:
                        ok = False
                        break
                    phi[mn] = gn
                    used.add(gn)
            if not ok:
                break
        if ok:
            yield combo


def brute_count(U, V, T, motif, delta, beg=0, end=None):
    edges = _sorted_edges(U, V, T)
    return sum(1 for _ in _matches(edges, motif, delta, beg, end))


def brute_enum(U, V, T, motif, delta, beg=0, end=None):
    """Sorted list of rows of original line numbers."""
    edges = _sorted_edges(U, V, T)
    rows = [tuple(edges[p][3] for p in c) for c in _matches(edges, motif, delta, beg, end)]
    return sorted(rows)
