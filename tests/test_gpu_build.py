"""The graph arrays built on the GPU (tm_time_order, tm_csr_build) equal the
oracle's restatement of the reference loader (oracle/prep.py) bit for bit,
and the corelib loaders give the same arrays with the build on the GPU and
on the host.  Every test prints its case before the GPU call."""
import sys
import time

import numpy as np
import pytest

from oracle import prep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _abi():
    from everest_amd import abi
    return abi


def test_time_order_is_stable():
    abi = _abi()
    rng = np.random.default_rng(5)
    for case in range(20):
        n = int(rng.integers(0, 20000))
        lo, hi = (-50, 50) if case % 3 == 0 else (0, int(rng.integers(1, 1 << 30)))
        T = rng.integers(lo, hi, n).astype(np.int32)
        say("time order case", case, "n", n, "range", lo, hi)
        got = abi.time_order(T)
        assert np.array_equal(got, np.argsort(T, kind="stable").astype(np.int32)), case


def test_csr_build_equals_oracle_prep():
    """Random edge lists (dense and sparse raw ids, negative ids, self-loops,
    tied times), time-sorted as the loader sorts them: Eg, both CSR indexes
    and offsets, and the renaming equal prep.preprocess's."""
    abi = _abi()
    rng = np.random.default_rng(9)
    for case in range(40):
        n = int(rng.integers(1, 30000))
        nodes = int(rng.integers(1, 2000))
        if case % 4 == 1:
            ids = rng.integers(-(1 << 30), 1 << 30, nodes)   # sparse, negative
        else:
            ids = np.arange(nodes) * int(rng.integers(1, 7))
        U = ids[rng.integers(0, nodes, n)].astype(np.int32)
        V = ids[rng.integers(0, nodes, n)].astype(np.int32)
        T = rng.integers(0, int(rng.integers(1, 5000)), n).astype(np.int32)
        g = prep.preprocess(U, V, T)
        o = np.argsort(T, kind="stable")
        say("csr build case", case, "n", n, "nodes", nodes)
        b = abi.csr_build(U[o], V[o], T[o])
        assert b["V"] == g["V"]
        assert np.array_equal(b["Eg"], g["Eg"])
        for k in ("outV", "inV", "outR", "inR"):
            assert np.array_equal(b[k], g[k]), (case, k)
        assert np.array_equal(b["rename"], g["rename"].astype(np.int32))


def test_loaders_build_on_gpu_as_on_host():
    """corelib's EdgeListLoader + GraphDataLoader with the build on the GPU
    and on the host give identical arrays on the 8M-edge graph; the 64M-edge
    graph's build times are printed."""
    from everest_amd import core
    from everest_amd.gen import graphs
    d = core.data
    names = ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid", "verticesFeatures", "edgeFeatures")
    for name in ("pl8m", "pl64m"):
        U, V, T = graphs.make(name)
        got = {}
        for dev in (-1, 0):
            d.setBuildDevice(dev)
            say(name, "build on", "host" if dev < 0 else "GPU %d" % dev)
            t = time.time()
            el = d.EdgeListLoader.fromArrays(name, U, V, T)
            hg = d.GraphDataLoader(el, d.FeatureLoader(el, "")).createGraphData()
            say("  %.2f s" % (time.time() - t))
            got[dev] = {k: hg.array(k).copy() for k in names}
            got[dev]["IDS"] = el.IDS().copy()
            del hg, el
        d.setBuildDevice(-2)
        for k in got[-1]:
            assert np.array_equal(got[-1][k], got[0][k]), (name, k)
        del got
