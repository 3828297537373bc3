"""KernelGen for constraint features (no GPU): the generated constrained
worker holds the predicates, gaps and filter as device code and compiles
with hipcc for gfx950, anti-edges included (the reference's M4-A example);
enumeration included; anti-edges with a new endpoint or directly after the
root, filters naming an anti edge, and enumeration with anti-edges are
rejected; malformed expressions are refused."""
import os
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _query(tmp, cons_lines, opts="", pattern="hub5.txt"):
    p = os.path.join(tmp, "q.yaml")
    with open(p, "w") as f:
        f.write("pattern: %s\ngraph: g.txt\nconstraints:\n  delta: 100\n%s\noptions:\n  gpu: 1\n%s" % (
            pattern, "".join("  %s\n" % x for x in cons_lines), opts))
    return p


def test_constrained_worker_generates_and_compiles(monkeypatch):
    from everest_amd.tmine import KernelGen, Query, utils
    monkeypatch.setenv("EVEREST_INPUTS", os.path.join(ROOT, "inputs"))
    with tempfile.TemporaryDirectory() as tmp:
        q = Query.Query(_query(tmp, ['node: "f % 2 == 0"', 'edge3: "f > 7"', "temporal: [5, 6, 7, 8]",
                                     'filter: "e0 < e4 && n2 != n3"'], "  enumerate: 1000\n"))
        code = KernelGen.KernelGen(q.pattern(), q).generate()
        assert "constexpr long long kEnumCap = 1000;" in code
        assert "tmc::cons_kernel<CSpec>" in code
        assert "case 3: return (f % 2 == 0);" in code and "case 3: return (f > 7);" in code
        assert "{5, 6, 7, 8, 0, 0, 0, 0}" in code
        assert "return (EF[pos[0]] < EF[pos[4]] && VF[m[2]] != VF[m[3]]);" in code
        src = os.path.join(tmp, "w.hip")
        with open(src, "w") as f:
            f.write(code)
        utils.compileModule(src, os.path.join(tmp, "w.so"))
        assert os.path.getsize(os.path.join(tmp, "w.so")) > 0


def test_unsupported_constraint_forms(monkeypatch):
    from everest_amd.tmine import KernelGen, Query
    monkeypatch.setenv("EVEREST_INPUTS", os.path.join(ROOT, "inputs"))
    with tempfile.TemporaryDirectory() as tmp:
        # hub5: (1,2) (1,3) (2,3) (3,4) (1,4); edge 1 and edge 3 bring new nodes
        for flags in ["[false, true, false, false, false]", "[false, false, false, true, false]"]:
            q = Query.Query(_query(tmp, ["antiEdge: %s" % flags]))
            with pytest.raises(NotImplementedError):
                KernelGen.KernelGen(q.pattern(), q)
        for bad in ["antiEdge: [true, false, false, false, false]", "antiEdge: [false, false, true]",
                    "antiEdge: [false, false, maybe, false, false]"]:
            q = Query.Query(_query(tmp, [bad]))
            with pytest.raises(ValueError):
                KernelGen.KernelGen(q.pattern(), q)
        q = Query.Query(_query(tmp, ["antiEdge: [false, false, true, false, false]", 'filter: "e2 > 1"']))
        with pytest.raises(ValueError):
            KernelGen.KernelGen(q.pattern(), q).generate()
        q = Query.Query(_query(tmp, ["antiEdge: [false, false, true, false, false]"], "  enumerate: 10\n",
                               pattern="M4-A.txt"))
        with pytest.raises(NotImplementedError):
            KernelGen.KernelGen(q.pattern(), q)
        for bad in ['node: "f > 1; return 0"', 'filter: "e9 > 1"', "temporal: [1, 2]"]:
            q = Query.Query(_query(tmp, [bad]))
            with pytest.raises(ValueError):
                KernelGen.KernelGen(q.pattern(), q).generate()


def test_anti_edge_worker_generates_and_compiles(monkeypatch):
    """The reference's anti-edge example (M4-A with edge 2 an anti edge,
    label predicates on the four nodes, temporal gaps; additional-features
    queries): tables of anti flags and nearest earlier non-anti edges."""
    from everest_amd.tmine import KernelGen, Query, utils
    monkeypatch.setenv("EVEREST_INPUTS", os.path.join(ROOT, "inputs"))
    with tempfile.TemporaryDirectory() as tmp:
        q = Query.Query(_query(tmp, ["node0: f > 500000000", "node1: f < 500000000", "node2: f > 500000000",
                                     "node3: f < 500000000", "antiEdge: [False, False, True, False, False]",
                                     "temporal: [21600, 43200, 43200, 21600]"], pattern="M4-A.txt"))
        kg = KernelGen.KernelGen(q.pattern(), q)
        assert kg.anti == [False, False, True, False, False] and kg.prev == [-1, 0, 1, 1, 3]
        code = kg.generate()
        assert "{0, 0, 1, 0, 0, 0, 0, 0}" in code and "{-1, 0, 1, 1, 3, 0, 0, 0}" in code
        src = os.path.join(tmp, "w.hip")
        with open(src, "w") as f:
            f.write(code)
        utils.compileModule(src, os.path.join(tmp, "w.so"))
        assert os.path.getsize(os.path.join(tmp, "w.so")) > 0
