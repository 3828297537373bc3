"""YAML query model (reference: system/tmine/Query.py:1-338).

A query file holds:
    pattern: <motif file under inputs/motifs>
    graph:   <edge list under inputs/graphs>
    constraints: {delta: <seconds>, ...}
    options: {gpu, warpCentric, warpBalancing, tailExpansion, enumerate}

Options default to gpu=1 and warpCentric/warpBalancing/tailExpansion=True
(Query.py:6-12).  Constraint features: label predicates (`node`, `nodeK`,
`edge`, `edgeK`: C expressions over the label `f`), temporal gaps
(`temporal`), a filter over the matched labels (`filter`: eK / nK) and
anti-edges (`antiEdge`: one flag per motif edge; an anti edge's two
endpoints mapped by earlier edges) are compiled into a constrained worker by
KernelGen (counting); enumeration under constraints is rejected.
"""
import os

import yaml

from . import locations


class Options:
    def __init__(self, opts=None):
        if opts is None:
            opts = {"gpu": 1, "warpCentric": True, "warpBalancing": True, "tailExpansion": True}
        self.opts = dict(opts)
        if self.multigpu():
            self.opts["multigpu"] = True

    def __str__(self):
        return "Options:\n %s" % self.opts

    __repr__ = __str__

    def gpu(self):
        return self.opts.get("gpu", 1)

    def multigpu(self):
        return self.opts.get("gpu", 1) > 1

    def enumerate(self):
        return "enumerate" in self.opts

    def enumNum(self):
        return int(self.opts["enumerate"])

    def update(self, rhs):
        return self.opts.update(rhs.opts)


FEATURE_KEYS = ("temporal", "antiEdge", "filter")


class Constraints:
    def __init__(self, constraints):
        self.cons = dict(constraints or {})

    def __str__(self):
        return "Constraints:\n %s" % self.cons

    __repr__ = __str__

    def isDelta(self):
        return "delta" in self.cons

    def isTemporal(self):
        return "temporal" in self.cons

    def isAnti(self):
        return "antiEdge" in self.cons

    def isFilter(self):
        return "filter" in self.cons

    def delta(self):
        return int(self.cons["delta"])

    def hasVertexCons(self):
        return any(k.startswith("node") for k in self.cons)

    def hasEdgeCons(self):
        return any(k.startswith("edge") for k in self.cons)

    def features(self):
        """Constraint keys beyond the delta window."""
        return [k for k in self.cons if k != "delta"]

    def genConsTemplates(self):
        """The reference writes constraint snippets here (Query.py:136-166);
        this build emits them into the worker itself (KernelGen
        constraint_bodies)."""
        return None


class Query:
    def __init__(self, pathToQuery):
        with open(pathToQuery, "r") as f:
            self.queryFile = yaml.safe_load(f)
        self.name = pathToQuery
        self.cons = Constraints(self.queryFile.get("constraints", {}))
        opts = Options()
        opts.update(Options(self.queryFile.get("options") or {}))
        self.opts = opts
        self.queryFile["pattern"] = os.path.join(locations.inputMotifsRoot(), self.queryFile["pattern"])
        self.queryFile["graph"] = os.path.join(locations.inputGraphRoot(), self.queryFile["graph"])

    def __str__(self):
        return "Query %s:\n %s" % (self.name, self.queryFile)

    __repr__ = __str__

    def pattern(self):
        return self.queryFile["pattern"]

    def graph(self):
        return self.queryFile["graph"]

    def constraints(self):
        return self.cons

    def options(self):
        return self.opts
