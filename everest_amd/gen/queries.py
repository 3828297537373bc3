"""Query workloads: motif lists and query streams for the benchmark configs.

Motifs are lists of (src, dst) pairs over renamed node ids: edge 0 is 0->1,
every later edge touches a node already used (prefix-connected, the only form
the reference motif pre-decoder accepts, src/core/cpu/data.cc:422-447), a new
node takes the next id, and there are no self-loops.
"""
import math

import numpy as np

DAY = 86400


def all_3edge_motifs():
    """The 36 prefix-connected, self-loop-free 3-edge motifs on <= 3 nodes
    (the 6x6 grid of Paranjape et al.): 4 on two nodes, 32 on three."""
    out = []
    for e1 in [(0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1)]:
        for e2 in [(0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1)]:
            out.append([(0, 1), e1, e2])
    return out


CONFIG3_MOTIFS = {
    "cycle4": [(0, 1), (1, 2), (2, 3), (3, 0)],
    "hub5": [(0, 1), (0, 2), (1, 2), (2, 3), (0, 3)],
    "bitri6": [(0, 1), (1, 0), (1, 2), (2, 1), (2, 0), (0, 2)],
}
CONFIG3_DELTAS = (3600, 86400)


def allowed_next(n, max_nodes):
    """Ordered pairs that may extend a motif currently using nodes 0..n-1."""
    out = []
    top = n + 1 if n < max_nodes else n
    for a in range(top):
        for b in range(top):
            if a == b or (a == n and b == n):
                continue
            out.append((a, b))
    return out


def random_motif(rng, M, max_nodes=4):
    edges = [(0, 1)]
    n = 2
    for _ in range(1, M):
        cands = allowed_next(n, max_nodes)
        a, b = cands[int(rng.integers(0, len(cands)))]
        if a == n or b == n:
            n += 1
        edges.append((a, b))
    return edges


def random_delta(rng, lo=600, hi=604800):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def batch_queries(n=256, seed=4, m_lo=3, m_hi=6, max_nodes=4):
    """Config-4 query set: M ~ U{m_lo..m_hi}, random prefix-connected motif on
    <= max_nodes nodes, delta log-uniform in [600, 604800] s."""
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(n):
        M = int(rng.integers(m_lo, m_hi + 1))
        qs.append((random_motif(rng, M, max_nodes), random_delta(rng)))
    return qs


def as_motif(edges):
    """(a, b) pairs -> (a, b, k) triples in time order."""
    return [(a, b, k) for k, (a, b) in enumerate(edges)]
