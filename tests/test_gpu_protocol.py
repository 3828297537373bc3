"""GPU tests of the scheduling protocol: work donation forced on every
step, the ring's wrap-around, and the deadline path.

Counts are compared with the oracle (small graphs, 8M-edge root samples)
and with the default launch (same kernel, donation after 1 ms).  Every test
prints its case before the GPU call.
"""
import sys

import numpy as np
import pytest

from oracle import oracle, prep
from helpers import random_graph, random_motif

pytestmark = pytest.mark.gpu

# donation as eagerly as the protocol allows: after 1 us of a task, checked
# every step, frames of 2 candidates split
FORCE = dict(steal_after_us=1, steal_every=1, steal_min=1, steal_max=4)


def _abi():
    from everest_amd import abi
    return abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pl8m():
    from everest_amd import abi, core
    from everest_amd.gen import graphs
    U, V, T = graphs.make("pl8m")
    el = core.data.EdgeListLoader.fromArrays("pl8m", U, V, T)
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    arrs = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid")}
    g = dict(E=hg.numEdges, Eg=arrs["Eg"].reshape(-1, 3), inV=arrs["inEdgesV"], outV=arrs["outEdgesV"],
             inR=arrs["inEdgesR"], outR=arrs["outEdgesR"], Euid=arrs["Euid"])
    dg = abi.DeviceGraph(arrs["Eg"], arrs["inEdgesV"], arrs["outEdgesV"], arrs["inEdgesR"], arrs["outEdgesR"],
                         arrs["Euid"])
    return g, dg


@pytest.fixture(scope="module")
def pl8m():
    g, dg = _pl8m()
    yield g, dg
    dg.close()


def test_forced_donation_small_random():
    """60 random small graphs: with donation forced, counts equal the oracle,
    wavefronts do donate, and every donated frame is taken."""
    abi = _abi()
    rng = np.random.default_rng(11)
    donated = 0
    for i in range(60):
        E = int(rng.integers(50, 2000))
        N = int(rng.integers(2, 30))
        U, V, T = random_graph(rng, E, N, int(rng.integers(1, 200)))
        mot = random_motif(rng, int(rng.integers(3, 7)))
        delta = int(rng.integers(0, 120))
        g = prep.preprocess(U, V, T)
        say("forced donation case", i, "E", E, "motif", mot, "delta", delta)
        want = oracle.count(g, mot, delta, nthreads=4)
        dg = abi.DeviceGraph.from_dict(g)
        got = dg.count(mot, delta, tune=abi.tuning(**FORCE))
        st = dg.last_stats()
        dg.close()
        assert got == want, (i, got, want)
        assert st["taken"] == st["donated"], (i, st)
        donated += st["donated"]
    say("frames donated over the 60 cases:", donated)
    assert donated > 0


def test_forced_donation_pl8m(pl8m):
    """8M-edge graph, 4- to 6-edge motifs over 100,000-root ranges: counts
    with donation forced equal the default launch's and the oracle on a
    20,000-root sample; donations happen and are all taken."""
    abi = _abi()
    from everest_amd.gen import queries
    g, dg = pl8m
    rng = np.random.default_rng(31)
    cases = [([(0, 1), (0, 2), (0, 3), (0, 1)], 604800), ([(0, 1), (2, 1), (0, 1), (2, 1)], 86400)]
    for i in range(6):
        cases.append((queries.random_motif(rng, 4 + i % 3), int(rng.choice([3600, 86400, 604800]))))
    donated = 0
    for edges, delta in cases:
        mot = queries.as_motif(edges)
        beg = int(rng.integers(0, g["E"] - 200_000))
        end = beg + 100_000
        say("pl8m forced donation", edges, "delta", delta, "roots", beg, end)
        want = dg.count(mot, delta, beg, end)
        got = dg.count(mot, delta, beg, end, tune=abi.tuning(**FORCE))
        st = dg.last_stats()
        assert got == want, (edges, delta, got, want)
        assert st["taken"] == st["donated"], st
        donated += st["donated"]
        assert dg.count(mot, delta, beg, beg + 20_000, tune=abi.tuning(**FORCE)) == \
            oracle.count(g, mot, delta, beg, beg + 20_000, nthreads=8)
    say("pl8m frames donated:", donated)
    assert donated > 0


def test_donation_ring_wraps(pl8m):
    """A launch that donates more frames than the ring has slots (2^18):
    donors reuse slots after their takers released them; the count equals
    the default launch's."""
    abi = _abi()
    from everest_amd.gen import queries
    g, dg = pl8m
    mot = queries.as_motif([(0, 1), (0, 2), (0, 3), (0, 1), (0, 2)])
    beg, end = 0, 3_000_000
    say("pl8m ring wrap", mot, "roots", beg, end)
    want = dg.count(mot, 604800, beg, end)
    got = dg.count(mot, 604800, beg, end, tune=abi.tuning(**FORCE))
    st = dg.last_stats()
    say("donated", st["donated"], "taken", st["taken"])
    assert got == want
    assert st["taken"] == st["donated"]
    if st["donated"] <= (1 << 18):
        pytest.skip("only %d frames donated: the ring did not wrap on this run" % st["donated"])


def test_deadline_then_exact(pl8m):
    """A 1 ms time limit on a query that runs far longer ends the launch with
    TM_ERR_TIMEOUT; the next call on the same graph is bit-exact (against
    the unlimited count and the oracle on a sample)."""
    abi = _abi()
    from everest_amd.gen import queries
    g, dg = pl8m
    mot = queries.as_motif([(0, 1), (0, 2), (0, 3), (0, 1), (0, 2), (0, 3)])
    say("pl8m deadline: unlimited run", mot)
    want = dg.count(mot, 604800)
    ms = dg.last_kernel_ms()
    say("unlimited kernel ms", ms)
    assert ms > 20.0, "query too light to test a 1 ms deadline"
    say("pl8m deadline: 1 ms limit")
    with pytest.raises(abi.TMError) as ei:
        dg.count(mot, 604800, tune=abi.tuning(time_limit_ms=1))
    assert ei.value.status == abi.TM_ERR_TIMEOUT
    say("pl8m deadline: 1 ms limit, donation forced")
    with pytest.raises(abi.TMError) as ei:
        dg.count(mot, 604800, tune=abi.tuning(time_limit_ms=1, **FORCE))
    assert ei.value.status == abi.TM_ERR_TIMEOUT
    say("pl8m deadline: batch with 1 ms limit")
    with pytest.raises(abi.TMError) as ei:
        dg.batch_count([(mot, 604800), (queries.as_motif([(0, 1), (1, 2), (2, 0)]), 3600)],
                       tune=abi.tuning(time_limit_ms=1))
    assert ei.value.status == abi.TM_ERR_TIMEOUT
    say("pl8m deadline: exact after the timeouts")
    assert dg.count(mot, 604800) == want
    beg = 3_000_000
    assert dg.count(mot, 604800, beg, beg + 5_000) == oracle.count(g, mot, 604800, beg, beg + 5_000, nthreads=8)
