"""Constraint features end to end on the GPU: a YAML query with label
predicates on motif nodes / edges, temporal gaps, a filter and anti-edges goes through
the frontend (MiningSingle: loader, generated constrained worker compiled
for gfx950, SingleGPUExecution) and its count equals the constraint checker
(tests/cons_check.py, pinned against the oracle) on the same graph and
labels.  Each case is printed before its GPU call."""
import os
import sys

import numpy as np
import pytest

from oracle import motif as omotif
import cons_check as cc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (motif file, delta, YAML constraint lines, node preds, edge preds, temporal, filter) with
# the C expressions of the YAML restated as Python for the checker
CASES = [
    ("cycle3.txt", 3000, ['node: "f % 2 == 0"'], {i: (lambda f: f % 2 == 0) for i in range(8)}, {}, None, None),
    ("cycle3.txt", 3000, ['edge: "f % 3 != 0"'], {}, {j: (lambda f: f % 3 != 0) for j in range(8)}, None, None),
    ("M4.txt", 4000, ['node1: "f < 600000000"', 'edge2: "f > 200000000"'],
     {1: lambda f: f < 600000000}, {2: lambda f: f > 200000000}, None, None),
    ("M4.txt", 4000, ["temporal: [1500, 800, 2000]"], {}, {}, [1500, 800, 2000], None),
    ("hub5.txt", 5000, ['filter: "(e0 + e4) % 2 == 0 && n3 % 5 != 1"'], {}, {}, None,
     lambda el, nl: (el[0] + el[4]) % 2 == 0 and nl[3] % 5 != 1),
    ("M1.txt", 2500, ['node: "f % 3 == 1"', 'edge0: "f % 2 == 1"', "temporal: [900, 900]",
                      'filter: "e1 < e2"'],
     {i: (lambda f: f % 3 == 1) for i in range(8)}, {0: lambda f: f % 2 == 1}, [900, 900],
     lambda el, nl: el[1] < el[2]),
    # anti edges: the reference's M4-A example (predicates on the four nodes,
    # temporal gaps), an anti edge last under an edge predicate, anti edges
    # in the middle and last, two in a row with gaps, one with a filter
    ("M4-A.txt", 8000, ["node0: f > 500000000", "node1: f < 500000000", "node2: f > 500000000",
                        "node3: f < 500000000", "antiEdge: [False, False, True, False, False]",
                        "temporal: [3000, 4000, 4000, 3000]"],
     {0: lambda f: f > 500000000, 1: lambda f: f < 500000000, 2: lambda f: f > 500000000,
      3: lambda f: f < 500000000}, {}, [3000, 4000, 4000, 3000], None, [False, False, True, False, False]),
    ("cycle3.txt", 3000, ['edge: "f % 3 != 0"', "antiEdge: [false, false, true]"], {},
     {j: (lambda f: f % 3 != 0) for j in range(8)}, None, None, [False, False, True]),
    ("hub5.txt", 3000, ["antiEdge: [false, false, true, false, true]"], {}, {}, None, None,
     [False, False, True, False, True]),
    ("bitri6.txt", 6000, ["antiEdge: [false, false, false, true, true, false]", "temporal: [3000, 3000, 2500, 2500, 3000]"],
     {}, {}, [3000, 3000, 2500, 2500, 3000], None, [False, False, False, True, True, False]),
    ("M4-A.txt", 4000, ["antiEdge: [false, false, true, false, false]", 'filter: "(e0 + e4) % 2 == 0"'], {}, {},
     None, lambda el, nl: (el[0] + el[4]) % 2 == 0, [False, False, True, False, False]),
    # 2-edge motifs: a reply under a node predicate, and roots with no reply in the window
    ("recip2.txt", 3000, ['node: "f % 2 == 0"'], {i: (lambda f: f % 2 == 0) for i in range(8)}, {}, None, None),
    ("recip2.txt", 1500, ["antiEdge: [false, true]"], {}, {}, None, None, [False, True]),
]


@pytest.fixture(scope="module")
def world(tmp_path_factory):
    from everest_amd import core
    from everest_amd.gen import graphs
    root = tmp_path_factory.mktemp("cinputs")
    (root / "graphs").mkdir()
    os.symlink(os.path.join(ROOT, "inputs", "motifs"), root / "motifs")
    rng = np.random.default_rng(61)
    E = 6000
    U = rng.integers(0, 40, E).astype(np.int32)
    V = rng.integers(0, 40, E).astype(np.int32)
    T = rng.integers(0, 40000, E).astype(np.int32)
    graphs.write_edge_file(str(root / "graphs" / "small.txt"), U, V, T)
    os.environ["EVEREST_INPUTS"] = str(root)
    # the graph and labels as the frontend loads them
    el = core.data.EdgeListLoader(str(root / "graphs" / "small.txt"))
    hg = core.data.GraphDataLoader(el, core.data.FeatureLoader(el, "")).createGraphData(0, 0, False)
    a = {k: hg.array(k) for k in ("Eg", "inEdgesV", "outEdgesV", "inEdgesR", "outEdgesR", "Euid",
                                   "edgeFeatures", "verticesFeatures")}
    g = dict(E=hg.numEdges, V=hg.numVertices, Eg=a["Eg"].reshape(-1, 3), inV=a["inEdgesV"], outV=a["outEdgesV"],
             inR=a["inEdgesR"], outR=a["outEdgesR"], Euid=a["Euid"])
    return root, g, a["edgeFeatures"], a["verticesFeatures"]


def test_constraint_queries_through_the_frontend(world):
    from everest_amd.tmine import frontend
    root, g, EF, VF = world
    m = None
    nonzero = 0
    for i, (pat, d, lines, npred, epred, tem, filt, *anti) in enumerate(CASES):
        anti = anti[0] if anti else None
        q = root / ("cq%d.yaml" % i)
        q.write_text("pattern: %s\ngraph: small.txt\nconstraints:\n  delta: %d\n%s\noptions:\n  gpu: 1\n" % (
            pat, d, "".join("  %s\n" % x for x in lines)))
        say("constraint query", i, pat, d, lines)
        if m is None:
            m = frontend.MiningSingle(str(q))
            cn, _, _ = m.run()
        else:
            cn, _, _ = frontend.Mining.run(m, str(q))
        edges = [(a, b) for a, b, _ in omotif.read_motif(os.path.join(ROOT, "inputs", "motifs", pat))]
        want = cc.count_constrained(g, edges, d, EF, VF, node_pred=npred, edge_pred=epred, temporal=tem, filt=filt,
                                    anti=anti)
        say("  count", cn, "want", want)
        assert cn == want, (i, pat, lines, cn, want)
        nonzero += want > 0
    assert nonzero >= 10


# enumerating constraint queries: (motif, delta, YAML lines, node preds, edge preds, temporal, filter)
ENUM_CASES = [
    ("cycle3.txt", 3000, ['node: "f % 2 == 0"', "temporal: [2000, 1500]", 'filter: "e0 < e2"'],
     {i: (lambda f: f % 2 == 0) for i in range(8)}, {}, [2000, 1500], lambda el, nl: el[0] < el[2]),
    ("M4.txt", 4000, ['edge: "f % 3 != 0"', 'node2: "f < 700000000"'], {2: lambda f: f < 700000000},
     {j: (lambda f: f % 3 != 0) for j in range(8)}, None, None),
]


def test_constraint_enumeration_through_the_frontend(world):
    """Rows of enumerating constraint queries: below the cap exactly the
    checker's rows (original edge ids); at a cap, the deterministic cut (the
    rows of the roots before the root that reaches the cap, then that root's
    smallest rows) computed from the checker's rows in root order."""
    from everest_amd.tmine import frontend
    from helpers import deterministic_cut
    root, g, EF, VF = world
    Euid = np.asarray(g["Euid"])
    m = None
    for i, (pat, d, lines, npred, epred, tem, filt) in enumerate(ENUM_CASES):
        edges = [(a, b) for a, b, _ in omotif.read_motif(os.path.join(ROOT, "inputs", "motifs", pat))]
        pos_rows = []
        total = cc.count_constrained(g, edges, d, EF, VF, node_pred=npred, edge_pred=epred, temporal=tem, filt=filt,
                                     out_rows=pos_rows)
        assert total == len(pos_rows) and total > 30
        all_rows = Euid[np.asarray(pos_rows, dtype=np.int64)]
        for cap in (total + 10, total // 3):
            q = root / ("ce%d_%d.yaml" % (i, cap))
            q.write_text("pattern: %s\ngraph: small.txt\nconstraints:\n  delta: %d\n%s\noptions:\n  gpu: 1\n"
                         "  enumerate: %d\n" % (pat, d, "".join("  %s\n" % x for x in lines), cap))
            say("constraint enumeration", i, pat, d, lines, "cap", cap, "total", total)
            if m is None:
                m = frontend.MiningSingle(str(q))
                cn, _, _, rows = m.run()
            else:
                cn, _, _, rows = frontend.Mining.run(m, str(q))
            want_rows, want_cnt = deterministic_cut(g, all_rows, cap)
            say("  count", cn, "want", want_cnt, "rows", len(rows))
            assert cn == want_cnt
            assert sorted(map(tuple, rows.tolist())) == sorted(want_rows)
