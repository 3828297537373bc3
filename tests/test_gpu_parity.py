"""GPU parity: HIP kernels (through the C ABI) against the oracle, bit-exact.

Every test prints its case (seed, sizes, motif, delta) before the GPU call.
"""
import sys

import numpy as np
import pytest

from oracle import oracle, prep
from helpers import random_graph, random_motif, all_3edge_motifs

pytestmark = pytest.mark.gpu


def _abi():
    from everest_amd import abi
    return abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_cases():
    rng = np.random.default_rng(11)
    cases = []
    for i in range(60):
        E = int(rng.integers(1, 400))
        N = int(rng.integers(2, 30))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, 200)))
        M = int(rng.integers(2, 7))
        mot = random_motif(rng, M)
        delta = int(rng.integers(0, 80))
        cases.append((i, U, V, T, mot, delta))
    return cases


def test_count_small_random(small_cases):
    abi = _abi()
    for (i, U, V, T, mot, delta) in small_cases:
        g = prep.preprocess(U, V, T)
        say("case", i, "E", g["E"], "motif", mot, "delta", delta)
        want = oracle.count(g, mot, delta, nthreads=4)
        dg = abi.DeviceGraph.from_dict(g)
        for scan in (1, 16):
            got = dg.count(mot, delta, tune=abi.tuning(scan_max=scan))
            assert got == want, (i, mot, delta, scan, got, want)
        dg.close()


def test_enum_small_random(small_cases):
    abi = _abi()
    for (i, U, V, T, mot, delta) in small_cases[:30]:
        g = prep.preprocess(U, V, T)
        total, rows = oracle.enum(g, mot, delta, 1 << 20)
        say("enum case", i, "motif", mot, "delta", delta, "total", total)
        dg = abi.DeviceGraph.from_dict(g)
        cnt, grows = dg.enum(mot, delta, cap=max(total, 1))
        dg.close()
        assert cnt == total
        assert sorted(map(tuple, grows.tolist())) == sorted(map(tuple, rows.tolist()))


def test_batch_equals_single(small_cases):
    abi = _abi()
    (i, U, V, T, _, _) = small_cases[5]
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    rng = np.random.default_rng(3)
    qs = [(random_motif(rng, int(rng.integers(2, 7))), int(rng.integers(0, 100))) for _ in range(40)]
    say("batch of", len(qs), "on case", i)
    got = dg.batch_count(qs)
    single = [dg.count(m, d) for (m, d) in qs]
    want = [oracle.count(g, m, d, nthreads=4) for (m, d) in qs]
    dg.close()
    assert got == single == want


def test_zipf100k_all36():
    abi = _abi()
    from everest_amd.gen import graphs
    U, V, T = graphs.make("zipf100k")
    g = prep.preprocess(U, V, T)
    dg = abi.DeviceGraph.from_dict(g)
    for mot in all_3edge_motifs():
        say("zipf100k motif", mot)
        want = oracle.count(g, mot, 3600, nthreads=8)
        got = dg.count(mot, 3600)
        assert got == want, (mot, got, want)
    dg.close()
