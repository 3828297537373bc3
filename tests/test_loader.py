"""The product loader / preprocessor (libeverest_core through corelib) against
the oracle restatement, bit-exact, plus the reference file and cache formats."""
import os

import numpy as np

from everest_amd import core
from everest_amd.gen import graphs
from oracle import motif as omotif, prep
from helpers import random_graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _product(U, V, T, name="t"):
    el = core.data.EdgeListLoader.fromArrays(name, U, V, T)
    fl = core.data.FeatureLoader(el, "")
    return core.data.GraphDataLoader(el, fl).createGraphData()


def _same(hg, g):
    return (np.array_equal(hg.array("Eg"), g["Eg"].ravel()) and np.array_equal(hg.array("outEdgesV"), g["outV"])
            and np.array_equal(hg.array("inEdgesV"), g["inV"]) and np.array_equal(hg.array("outEdgesR"), g["outR"])
            and np.array_equal(hg.array("inEdgesR"), g["inR"]) and np.array_equal(hg.array("Euid"), g["Euid"]))


def test_loader_matches_oracle_random():
    rng = np.random.default_rng(1)
    for i in range(30):
        U, V, T = random_graph(rng, int(rng.integers(1, 3000)), int(rng.integers(1, 400)), int(rng.integers(1, 100)))
        if i % 3 == 0:  # arbitrary, sparse, negative ids
            U = (U.astype(np.int64) * 7919 - 1000).astype(np.int32)
            V = (V.astype(np.int64) * 7919 - 1000).astype(np.int32)
        hg = _product(U, V, T)
        assert _same(hg, prep.preprocess(U, V, T))
        assert hg.numEdges == U.size


def test_loader_matches_oracle_zipf100k():
    U, V, T = graphs.make("zipf100k")
    assert _same(_product(U, V, T), prep.preprocess(U, V, T))


def test_edge_file_and_caches(tmp_path):
    rng = np.random.default_rng(3)
    U, V, T = random_graph(rng, 5000, 300, 1000)
    path = tmp_path / "g.txt"
    graphs.write_edge_file(str(path), U, V, T)
    el = core.data.EdgeListLoader(str(path))
    assert el.edgeListLength() == 5000
    for name in ("U.cache", "V.cache", "T.cache", "IDS.cache"):
        assert (tmp_path / "g" / name).exists()
    fl = core.data.FeatureLoader(el, "")
    gl = core.data.GraphDataLoader(el, fl)
    hg = gl.createGraphData()
    assert (tmp_path / "g" / "g-0-5000" / "Eg").exists()
    assert _same(hg, prep.preprocess(U, V, T))
    hg2 = gl.createGraphData()  # from the cache directory
    assert np.array_equal(hg2.array("Eg"), hg.array("Eg"))
    el2 = core.data.EdgeListLoader(str(path))  # from U/V/T/IDS caches
    assert np.array_equal(el2.IDS(), el.IDS())


def test_motif_reader_and_predecode():
    for f in sorted(os.listdir(os.path.join(ROOT, "inputs", "motifs"))):
        p = os.path.join(ROOT, "inputs", "motifs", f)
        m = core.data.HostMotifData(p)
        edges = omotif.read_motif(p)
        assert m.edges() == edges
        want = [(x.baseNode, x.constraintNode, x.mappedNodes, x.io, x.newNode) for x in omotif.predecode(edges)]
        assert m.minfo() == want


def test_generators_are_deterministic():
    a = graphs.make("zipf100k")
    b = graphs.make("zipf100k")
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.any(a[0] == a[1])
