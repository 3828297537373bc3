"""Pins the oracle (oracle/tm_oracle.c, oracle/prep.py, oracle/motif.py).

The reference ships no tests and no golden vectors, so the oracle is pinned by
(1) brute force from the match definition on random tiny graphs, (2) closed
forms, (3) invariants, and (4) the compiled reference motif reader.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import brute, motif as omotif, oracle, prep
from helpers import random_graph, random_motif

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_READER = os.path.join(ROOT, "oracle", "_ref", "motif_ref")
REF_MOTIFS = os.path.join(ROOT, "inputs", "motifs")


def test_oracle_matches_brute_force_counts_and_rows():
    rng = np.random.default_rng(5)
    for _ in range(250):
        E = int(rng.integers(1, 20))
        N = int(rng.integers(2, 6))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, 8)))
        mot = random_motif(rng, int(rng.integers(2, 5)))
        delta = int(rng.integers(0, 8))
        g = prep.preprocess(U, V, T)
        want = brute.brute_count(U, V, T, mot, delta)
        assert oracle.count(g, mot, delta, nthreads=2) == want
        total, rows = oracle.enum(g, mot, delta, 10000)
        assert total == want
        assert sorted(map(tuple, rows.tolist())) == brute.brute_enum(U, V, T, mot, delta)


def test_closed_form_parallel_edges():
    # n parallel edges u->v at distinct times, motif = M copies of 0->1: C(n, M)
    for n in (1, 5, 17):
        for M in (2, 3, 4):
            U = np.zeros(n, np.int32)
            V = np.ones(n, np.int32)
            T = np.arange(n, dtype=np.int32) * 10
            g = prep.preprocess(U, V, T)
            mot = [(0, 1, k) for k in range(M)]
            assert oracle.count(g, mot, 10 * n) == math.comb(n, M)


def test_closed_form_out_star():
    # hub 0 -> leaves 1..n at increasing times; motif 0->1, 0->2, 0->3 = C(n, 3)
    n = 12
    U = np.zeros(n, np.int32)
    V = np.arange(1, n + 1, dtype=np.int32)
    T = np.arange(n, dtype=np.int32)
    g = prep.preprocess(U, V, T)
    assert oracle.count(g, [(0, 1, 0), (0, 2, 1), (0, 3, 2)], n) == math.comb(n, 3)
    # leaves must be distinct: a repeated leaf does not form a 3-star
    U2 = np.zeros(4, np.int32)
    V2 = np.array([1, 1, 2, 3], np.int32)
    g2 = prep.preprocess(U2, V2, np.arange(4, dtype=np.int32))
    assert oracle.count(g2, [(0, 1, 0), (0, 2, 1), (0, 3, 2)], 10) == 2  # (1,2,3) with either 0->1


def test_closed_form_cycles_and_window():
    # r repetitions of the directed 4-cycle, far apart in time
    r, gap = 7, 1000
    U, V, T = [], [], []
    for i in range(r):
        for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 3), (3, 0)]):
            U.append(a); V.append(b); T.append(i * gap + k)
    g = prep.preprocess(np.array(U, np.int32), np.array(V, np.int32), np.array(T, np.int32))
    cyc = [(0, 1, 0), (1, 2, 1), (2, 3, 2), (3, 0, 3)]
    assert oracle.count(g, cyc, 3) == r
    assert oracle.count(g, cyc, 2) == 0          # window is inclusive: span 3 needs delta >= 3


def test_equal_times_use_position_order():
    # two edges with the same time: the stable sort keeps file order, so only
    # the file order 0->1 then 1->2 matches the path motif
    U = np.array([0, 1], np.int32); V = np.array([1, 2], np.int32); T = np.array([5, 5], np.int32)
    g = prep.preprocess(U, V, T)
    assert oracle.count(g, [(0, 1, 0), (1, 2, 1)], 0) == 1
    U2 = np.array([1, 0], np.int32); V2 = np.array([2, 1], np.int32)
    g2 = prep.preprocess(U2, V2, T)
    assert oracle.count(g2, [(0, 1, 0), (1, 2, 1)], 0) == 0


def test_invariants_additivity_monotone_relabel():
    rng = np.random.default_rng(9)
    for _ in range(40):
        U, V, T = random_graph(rng, 300, 12, 200)
        mot = random_motif(rng, int(rng.integers(2, 6)))
        g = prep.preprocess(U, V, T)
        d = int(rng.integers(0, 60))
        total = oracle.count(g, mot, d)
        cut = int(rng.integers(0, g["E"]))
        assert oracle.count(g, mot, d, 0, cut) + oracle.count(g, mot, d, cut, -1) == total
        assert oracle.count(g, mot, d + 7) >= total
        perm = rng.permutation(12).astype(np.int32)
        g2 = prep.preprocess(perm[U], perm[V], T)
        assert oracle.count(g2, mot, d) == total


def test_prep_layout():
    rng = np.random.default_rng(2)
    U, V, T = random_graph(rng, 500, 40, 50)
    g = prep.preprocess(U, V, T)
    Eg = g["Eg"]
    assert np.all(np.diff(Eg[:, 2]) >= 0)
    # stable: equal times keep line order
    ids = g["Euid"]
    same = Eg[1:, 2] == Eg[:-1, 2]
    assert np.all(ids[1:][same] > ids[:-1][same])
    # first-appearance renaming, source before destination
    seen = []
    for u, v in Eg[:, :2]:
        for x in (u, v):
            if x not in seen:
                assert x == len(seen)
                seen.append(x)
    for R, P, col in ((g["outR"], g["outV"], 0), (g["inR"], g["inV"], 1)):
        for x in range(g["V"]):
            lst = P[R[x]:R[x + 1]]
            assert np.all(np.diff(lst) > 0)
            assert np.array_equal(lst, np.flatnonzero(Eg[:, col] == x))


@pytest.mark.skipif(not os.path.exists(REF_READER), reason="reference motif reader not built (no /root/reference)")
def test_motif_reader_matches_reference_build(tmp_path):
    files = sorted(os.listdir(REF_MOTIFS))
    rng = np.random.default_rng(4)
    for i in range(20):  # random motif files: shuffled lines, arbitrary ids, tied times
        M = int(rng.integers(2, 7))
        mot = random_motif(rng, M)
        ids = rng.permutation(1000)[:8] + 1
        lines = ["%d %d %d" % (ids[a], ids[b], int(rng.integers(1, 4))) for a, b, _ in mot]
        p = tmp_path / ("rnd%d.txt" % i)
        p.write_text("\n".join(lines) + "\n")
        files.append(str(p))
    for f in files:
        path = f if os.path.isabs(f) else os.path.join(REF_MOTIFS, f)
        out = subprocess.check_output([REF_READER, path], text=True).split("\n")
        nn, ne = map(int, out[0].split())
        ref_edges = [tuple(map(int, l.split())) for l in out[1:1 + ne]]
        ours = omotif.read_motif(path)
        assert ours == ref_edges, path
        assert omotif.num_nodes(ours) == nn
