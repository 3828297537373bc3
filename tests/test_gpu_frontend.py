"""End to end on the GPU through the reference-shaped API: YAML query ->
generated gfx950 worker plugin -> GPUWorkerLibManager -> SingleGPUExecution,
and the stock plugin libTM.so as in the reference's legacy runner; counts and
enumerated rows bit-exact against the oracle."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle, prep
from helpers import deterministic_cut

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def inputs(tmp_path_factory):
    from everest_amd.gen import graphs
    root = tmp_path_factory.mktemp("inputs")
    (root / "graphs").mkdir()
    os.symlink(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "inputs", "motifs"),
               root / "motifs")
    U, V, T = graphs.make("zipf100k")
    graphs.write_edge_file(str(root / "graphs" / "zipf.txt"), U, V, T)
    os.environ["EVEREST_INPUTS"] = str(root)
    return root, prep.preprocess(U, V, T)


def _query(root, name, pattern, delta, extra=""):
    p = root / name
    p.write_text("pattern: %s\ngraph: zipf.txt\nconstraints:\n  delta: %d\noptions:\n  gpu: 1\n%s" % (
        pattern, delta, extra))
    return str(p)


def test_generated_worker_counts(inputs):
    from everest_amd.tmine import frontend
    from oracle import motif as omotif
    root, g = inputs
    m = None
    for pat, d in (("cycle3.txt", 3600), ("M1.txt", 1800), ("M4.txt", 7200), ("hub5.txt", 3600)):
        q = _query(root, "q_%s.yaml" % pat, pat, d)
        say("frontend query", pat, d)
        if m is None:
            m = frontend.MiningSingle(q)
            cn, _, _ = m.run()
        else:  # later queries reuse the resident graph (Mining.run takes a query path)
            cn, _, _ = frontend.Mining.run(m, q)
        mot = omotif.read_motif(os.path.join(str(root), "motifs", pat))
        assert cn == oracle.count(g, mot, d)


def test_generated_worker_enumerates(inputs):
    """Below the cap the rows are exactly the oracle's; at the cap (where the
    reference returns whichever rows its schedule produced) they are the
    deterministic cut: the rows of the roots before the root that reaches
    the cap, then that root's smallest rows."""
    from everest_amd.tmine import frontend
    from oracle import motif as omotif
    root, g = inputs
    mot = omotif.read_motif(os.path.join(str(root), "motifs", "cycle3.txt"))
    total, want = oracle.enum(g, mot, 1200, 1 << 20)
    assert total == len(want)
    want_set = set(map(tuple, want.tolist()))
    q = _query(root, "qe.yaml", "cycle3.txt", 1200, "  enumerate: %d\n" % (total + 10))
    say("frontend enumerate cycle3 1200 cap", total + 10)
    cn, _, _, rows = frontend.MiningSingle(q).run()
    assert cn == total
    assert sorted(map(tuple, rows.tolist())) == sorted(want_set)
    cap = total // 3
    q = _query(root, "qe2.yaml", "cycle3.txt", 1200, "  enumerate: %d\n" % cap)
    say("frontend enumerate cycle3 1200 cap", cap)
    cn, _, _, rows = frontend.MiningSingle(q).run()
    want_rows, want_cnt = deterministic_cut(g, want, cap)
    assert cn == want_cnt
    assert sorted(map(tuple, rows.tolist())) == sorted(want_rows)


def test_stock_plugin_and_executions(inputs):
    """The reference's legacy flow (system/legacy.py) with libTM.so."""
    from everest_amd import core
    from everest_amd.abi import LIB_DIR
    from oracle import motif as omotif
    root, g = inputs
    el = core.data.EdgeListLoader(str(root / "graphs" / "zipf.txt"))
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    motif_path = os.path.join(str(root), "motifs", "M3.txt")
    hm = core.data.HostMotifData(motif_path)
    dm = core.data.MakeGraphDataManager(1)
    lm = core.GPUWorkerLibManager()
    lm.openLib(os.path.join(LIB_DIR, "libTM.so"))
    w = lm.construct("TM", 0)
    ex = core.SingleGPUExecution(dm)
    ex.setHostGraph(hg)
    ex.setHostMotif(hm)
    ex.setWorker(w)
    say("legacy flow M3 delta 3600")
    want = oracle.count(g, omotif.read_motif(motif_path), 3600)
    assert ex.run(3600) == want
    assert ex.run(3600) == 2 * want  # run() is cumulative over the worker's lifetime, as the reference's
    div = core.SingleGPUExecutionDiv(dm, 0, 7)
    div.setHostGraph(hg)
    div.setHostMotif(hm)
    w2 = lm.construct("TM", 0)
    div.setWorker(w2)
    assert div.run(3600) == want
    del ex, div, w, w2


def test_handwritten_worker_reads_device_job_data(inputs):
    """A worker written against the reference's DeviceJobData fields (Eg_d,
    minfo()[i].arrR / arrV on the device: libCSR.so, one thread per root,
    plain CSR scans) loads through GPUWorkerLibManager and counts
    oracle-exact through SingleGPUExecution, cumulatively like the
    reference's workers."""
    from everest_amd import core
    from everest_amd.abi import LIB_DIR
    from oracle import motif as omotif
    root, g = inputs
    el = core.data.EdgeListLoader(str(root / "graphs" / "zipf.txt"))
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData()
    dm = core.data.MakeGraphDataManager(1)
    lm = core.GPUWorkerLibManager()
    lm.openLib(os.path.join(LIB_DIR, "libCSR.so"))
    for pat, d in (("cycle3.txt", 3600), ("M1.txt", 1800), ("M3.txt", 3600), ("M4.txt", 7200), ("hub5.txt", 3600)):
        motif_path = os.path.join(str(root), "motifs", pat)
        hm = core.data.HostMotifData(motif_path)
        w = lm.construct("CSR", 0)
        ex = core.SingleGPUExecution(dm)
        ex.setHostGraph(hg)
        ex.setHostMotif(hm)
        ex.setWorker(w)
        say("hand-written worker", pat, d)
        want = oracle.count(g, omotif.read_motif(motif_path), d)
        assert ex.run(d) == want, pat
        assert ex.run(d) == 2 * want, pat  # cumulative
        del ex, w
