"""The root-pair sweep count (4-edge motifs whose last edge repeats the
root's pair, edge 1 a new node on a root node's list, edge 2 a new node on
edge 1's new node's list; each edge-1 slot expanded once per pair and
weighted by the roots whose windows reach each closing edge) equals the
oracle: random graphs with self-loops, parallel edges and tied times, all
eight shapes, root subranges, and a delta of 0."""
import numpy as np
import pytest

from oracle import oracle, prep
from helpers import random_graph
import kernel_model as km

SHAPES = [[(0, 1), (s1, 2) if d1 else (2, s1), (2, 3) if d2 else (3, 2), (0, 1)]
          for s1 in (0, 1) for d1 in (1, 0) for d2 in (0, 1)]


def test_plan_shapes():
    for e in SHAPES:
        assert km.pair_sweep_plan(km.build_motif(e)) is not None, e
    for e in ([(0, 1), (0, 2), (3, 2), (1, 0)], [(0, 1), (0, 2), (3, 0), (0, 1)], [(0, 1), (0, 2), (2, 1), (0, 1)],
              [(0, 1), (0, 2), (3, 2)], [(0, 1), (0, 1), (2, 1), (0, 1)], [(0, 1), (0, 2), (3, 2), (0, 1), (0, 1)]):
        assert km.pair_sweep_plan(km.build_motif(e)) is None, e


@pytest.mark.parametrize("seed", range(6))
def test_pair_sweep_equals_oracle(seed):
    rng = np.random.default_rng(300 + seed)
    nz = 0
    for case in range(10):
        E = int(rng.integers(4, 260))
        N = int(rng.integers(2, 9))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, 100)))
        g = prep.preprocess(U, V, T)
        view = km.build_view(g)
        edges = SHAPES[(seed * 10 + case) % len(SHAPES)]
        md = km.build_motif(edges)
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        delta = int(rng.integers(0, 70)) if case % 5 else 0
        M = 4
        for beg, end in ((0, g["E"] - M + 1), (int(rng.integers(0, g["E"])), int(rng.integers(0, g["E"] + 1)))):
            end = min(end, g["E"] - M + 1)
            want = oracle.count(g, mot, delta, beg, max(beg, end), nthreads=1) if end > beg else 0
            for scan_max in (32, 2, 0):  # direct scans, the switch to ranks, ranks only
                got = km.pair_sweep_count(view, md, delta, beg, max(beg, end), scan_max=scan_max)
                assert got == want, (seed, case, edges, delta, beg, end, scan_max)
            nz += want > 0
    assert nz >= 4
