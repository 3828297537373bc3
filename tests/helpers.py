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
