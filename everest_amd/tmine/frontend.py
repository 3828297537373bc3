"""Query frontend (reference: system/tmine/frontend.py:1-223).

HeavyRes loads and preprocesses a graph once; Mining keeps it resident on the
GPU, generates and compiles the per-query worker plugin, loads it through the
library manager and runs the query; results are printed between the
reference's "@@@ [output starts]" / "@@@ [output ends]" markers as
"count, microseconds, seconds", or as enumerated rows of original line
numbers.
"""
import logging
from timeit import default_timer as timer

import numpy as np

from . import KernelGen, Query, core, locations, utils
from .KernelGen import InspectMotif

log = logging.getLogger("everest.tmine")


class Result:
    def __init__(self, line):
        self.count, self.microsec = [int(s) for s in line.split(",")[:2]]

    def __str__(self):
        return "count: %d, microsec: %d" % (self.count, self.microsec)

    __repr__ = __str__


class HeavyRes:
    """Graph loaded and preprocessed once per Mining instance."""

    def __init__(self, graph, gpu):
        if gpu != 1:
            raise NotImplementedError("this build drives one GPU (gpu = %d)" % gpu)
        self.graph = graph
        self.eloader = core.data.EdgeListLoader(graph)
        self.floader = core.data.FeatureLoader(self.eloader, locations.randFilePath())
        self.gloader = core.data.GraphDataLoader(self.eloader, self.floader)
        self.gpu = gpu
        self.hostGraph = self.gloader.createGraphData()


class Mining:
    def __init__(self, pathToGraph, gpu, tuning=None):
        self.heavy = HeavyRes(pathToGraph, gpu)
        self.libManager = core.GPUWorkerLibManager()
        self.dataManager = core.data.MakeGraphDataManager(gpu)
        self.tuning = tuning

    def makePlugin(self, query):
        query.cons.genConsTemplates()
        code = KernelGen.KernelGen(query.pattern(), query, self.tuning).generate()
        utils.createDirIfNotExist(locations.dynLibSourceRoot())
        with open(locations.dynLibSource(), "w") as f:
            f.write(code)
        utils.compilePlugin()

    def processSingle(self, query):
        hostMotif = core.data.HostMotifData(query.pattern())
        w = self.libManager.construct(utils.libFileToLibName(locations.dynLibPath()), 0)
        ex = core.SingleGPUExecution(self.dataManager)
        ex.setHostGraph(self.heavy.hostGraph)
        ex.setHostMotif(hostMotif)
        ex.setWorker(w)

        start = timer()
        ex.run(query.constraints().delta())
        end = timer()
        cn = w.count()
        tm = w.time()

        print("@@@ [output starts]", flush=True)
        rows = None
        if not query.options().enumerate():
            print("%d, %d, %s" % (w.count(), w.time(), end - start))
        else:
            e = w.printEnum(query.options().enumNum())
            _, _, allE = InspectMotif(query.pattern())
            rows = np.array(e, dtype=np.int64).reshape(-1, allE)
            for k in rows:
                print(" ".join(map(str, k)))
        print("@@@ [output ends]", flush=True)
        del ex
        del w  # workers go before their library is unloaded
        return (cn, tm, (end - start)) if rows is None else (cn, tm, (end - start), rows)

    def processQuery(self, query):
        if query.graph() != self.heavy.graph or query.options().gpu() != self.heavy.gpu:
            raise ValueError("cannot process query %s on this Mining instance" % query.name)
        self.makePlugin(query)
        self.libManager.openLib(locations.dynLibPath())
        return self.processSingle(query)

    def run(self, pathToQuery):
        return self.processQuery(Query.Query(pathToQuery))


class MiningSingle(Mining):
    def __init__(self, pathToQuery, tuning=None):
        query = Query.Query(pathToQuery)
        super().__init__(query.graph(), query.options().gpu(), tuning)
        self.query = query

    def run(self):
        return self.processQuery(self.query)
