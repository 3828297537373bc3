"""Shared helpers for tests: random motifs and graphs."""
import numpy as np


def random_motif(rng, M, max_nodes=8):
    """Prefix-connected motif with first-appearance node ids, no self-loops."""
    edges = [(0, 1)]
    n = 2
    for _ in range(1, M):
        while True:
            a = int(rng.integers(0, min(n + 1, max_nodes)))
            b = int(rng.integers(0, min(n + 1, max_nodes)))
            if a == b or (a == n and b == n):
                continue
            if a == n or b == n:
                if n >= max_nodes:
                    continue
                n += 1
            edges.append((a, b))
            break
    return [(a, b, i) for i, (a, b) in enumerate(edges)]


def random_graph(rng, E, N, tspan):
    U = rng.integers(0, N, E).astype(np.int32)
    V = rng.integers(0, N, E).astype(np.int32)
    T = rng.integers(0, max(1, tspan), E).astype(np.int32)
    return U, V, T


ALL3 = None


def all_3edge_motifs():
    """The 36 prefix-connected, self-loop-free 3-edge motifs on <= 3 nodes."""
    out = []
    pairs2 = [(0, 1), (1, 0)]
    for e1 in [(0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1)]:
        n = 3 if 2 in e1 else 2
        cands = pairs2 + ([(0, 2), (2, 0), (1, 2), (2, 1)])
        for e2 in cands:
            if n == 2 and 2 in e2:
                pass
            out.append([(0, 1, 0), (e1[0], e1[1], 1), (e2[0], e2[1], 2)])
    return out


def deterministic_cut(g, all_rows, cap):
    """The rows tm_enum returns at a cap, from every row of a query in root
    order (oracle.enum with a cap above the total): all rows of the roots
    before r* (the first root at which the cumulative count reaches the cap)
    and, of root r*'s rows, the smallest by their Euid tuples.  Returns
    (rows as a list of tuples, count of the roots up to r*)."""
    rows = [tuple(r) for r in all_rows.tolist()]
    if len(rows) <= cap:
        return rows, len(rows)
    pos_of = np.empty(len(g["Euid"]), dtype=np.int64)
    pos_of[np.asarray(g["Euid"])] = np.arange(len(g["Euid"]))
    roots = [int(pos_of[r[0]]) for r in rows]
    assert roots == sorted(roots), "oracle rows are not in root order"
    rstar = roots[cap - 1]
    head = [r for r, p in zip(rows, roots) if p < rstar]
    star = sorted(r for r, p in zip(rows, roots) if p == rstar)
    return head + star[:cap - len(head)], len(head) + len(star)
