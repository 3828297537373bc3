"""The window-sweep count (3-edge stars and a new node's pair twice, summed
over each anchor node's roots by window moves) equals the oracle: random
graphs with self-loops, parallel edges and tied times, every eligible shape,
root subranges, and a delta of 0."""
import numpy as np
import pytest

from oracle import oracle, prep
from helpers import random_graph
import kernel_model as km

SHAPES = [[(0, 1), (0, 2), (0, 3)], [(0, 1), (2, 1), (3, 1)], [(0, 1), (1, 2), (1, 3)], [(0, 1), (2, 0), (3, 0)],
          [(0, 1), (0, 2), (0, 2)], [(0, 1), (2, 1), (2, 1)], [(0, 1), (1, 2), (1, 2)], [(0, 1), (2, 0), (2, 0)]]


def test_plan_shapes():
    for e in SHAPES:
        assert km.window_sweep_plan(km.build_motif(e)) is not None, e
    for e in ([(0, 1), (0, 2), (2, 0)], [(0, 1), (0, 2), (1, 2)], [(0, 1), (1, 0), (1, 0)], [(0, 1), (0, 2), (3, 2)],
              [(0, 1), (0, 2), (0, 3), (0, 4)], [(0, 1), (0, 1), (0, 1)]):
        assert km.window_sweep_plan(km.build_motif(e)) is None, e


@pytest.mark.parametrize("seed", range(6))
def test_window_sweep_equals_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    for case in range(12):
        E = int(rng.integers(1, 300))
        N = int(rng.integers(2, 12))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, 120)))
        g = prep.preprocess(U, V, T)
        view = km.build_view(g)
        edges = SHAPES[(seed * 12 + case) % len(SHAPES)]
        md = km.build_motif(edges)
        mot = [(a, b, i) for i, (a, b) in enumerate(edges)]
        delta = int(rng.integers(0, 60)) if case % 5 else 0
        M = 3
        for beg, end in ((0, g["E"] - M + 1), (int(rng.integers(0, g["E"])), int(rng.integers(0, g["E"] + 1)))):
            end = min(end, g["E"] - M + 1)
            want = oracle.count(g, mot, delta, beg, max(beg, end), nthreads=1) if end > beg else 0
            got = km.window_sweep_count(view, md, delta, beg, max(beg, end))
            assert got == want, (seed, case, edges, delta, beg, end, got, want)
