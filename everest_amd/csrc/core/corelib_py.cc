// corelib_py.cc -- Python module "corelib" (pybind11), the same module,
// submodule and class names as the reference binding (src/mods/corelib.cc:17-191),
// plus array accessors and an in-memory edge-list constructor.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl/filesystem.h>
#include <pybind11/stl_bind.h>

#include <sstream>

#include "everest/data.h"
#include "everest/gen.h"

namespace py = pybind11;
using namespace corelib;

PYBIND11_MAKE_OPAQUE(corelib::data::HostGraphDataList);
PYBIND11_MAKE_OPAQUE(corelib::data::SubPartitionList);

template <class T>
static std::string repr(const std::string &head, const T &x) {
  std::stringstream s;
  s << head << x;
  return s.str();
}

static py::array_t<int> ints_of(const data::SinkMemory *m) {
  size_t n = m ? m->size() / sizeof(int) : 0;
  py::array_t<int> a(n);
  if (n) memcpy(a.mutable_data(), m->data(), n * sizeof(int));
  return a;
}

static py::array_t<int> ints_of_src(const data::SourceMemory *m) {
  size_t n = m ? m->size() / sizeof(int) : 0;
  py::array_t<int> a(n);
  if (n) memcpy(a.mutable_data(), m->data(), n * sizeof(int));
  return a;
}

PYBIND11_MODULE(corelib, m) {
  m.doc() = "Temporal-motif query core: loaders, resident graphs, worker plugins and executions (MI355X)";

  auto d = m.def_submodule("data");
  py::bind_vector<data::HostGraphDataList>(d, "HostGraphDataList");
  py::bind_vector<data::SubPartitionList>(d, "SubPartitionList");

  py::class_<data::EdgeListLoader>(d, "EdgeListLoader", "Edge-list file (u v t per line) parsed and stably sorted by time")
      .def(py::init<std::filesystem::path>())
      .def_static(
          "fromArrays",
          [](const std::string &name, py::array_t<int, py::array::c_style | py::array::forcecast> U,
             py::array_t<int, py::array::c_style | py::array::forcecast> V,
             py::array_t<int, py::array::c_style | py::array::forcecast> T) {
            if (U.size() != V.size() || U.size() != T.size()) throw std::runtime_error("U, V, T sizes differ");
            py::gil_scoped_release rel;
            return std::make_unique<data::EdgeListLoader>(name, U.data(), V.data(), T.data(), (size_t)U.size());
          },
          py::arg("name"), py::arg("U"), py::arg("V"), py::arg("T"))
      .def("name", &data::EdgeListLoader::name)
      .def("edgeListLength", &data::EdgeListLoader::edgeListLength)
      .def("U", [](const data::EdgeListLoader &e) { return ints_of_src(e.getUPtr()); })
      .def("V", [](const data::EdgeListLoader &e) { return ints_of_src(e.getVPtr()); })
      .def("T", [](const data::EdgeListLoader &e) { return ints_of_src(e.getTPtr()); })
      .def("IDS", [](const data::EdgeListLoader &e) { return ints_of_src(e.getIDSPtr()); })
      .def("__repr__", [](const data::EdgeListLoader &e) { return "<EdgeListLoader " + e.name() + ">"; });

  py::class_<data::FeatureLoader>(d, "FeatureLoader", "Edge and vertex labels for a loaded edge list")
      .def(py::init<const data::EdgeListLoader *, std::filesystem::path>(), py::keep_alive<1, 2>())
      .def("name", &data::FeatureLoader::name)
      .def("edgeFeaturesLength", &data::FeatureLoader::edgeFeaturesLength)
      .def("verticesFeaturesLength", &data::FeatureLoader::verticesFeaturesLength)
      .def("__repr__", [](const data::FeatureLoader &e) { return "<FeatureLoader " + e.name() + ">"; });

  py::class_<data::HostGraphData>(d, "HostGraphData", "Renamed, CSR-indexed host arrays of an edge range")
      .def_readonly("name", &data::HostGraphData::name)
      .def_readonly("numEdges", &data::HostGraphData::numEdges)
      .def_readonly("numVertices", &data::HostGraphData::numVertices)
      .def_readonly("totalBytes", &data::HostGraphData::totalBytes)
      .def_readonly("beg", &data::HostGraphData::beg)
      .def_readonly("end", &data::HostGraphData::end)
      .def("array",
           [](const data::HostGraphData &h, const std::string &which) {
             if (which == "Eg") return ints_of(h.Eg_h.get());
             if (which == "inEdgesV") return ints_of(h.inEdgesV_h.get());
             if (which == "outEdgesV") return ints_of(h.outEdgesV_h.get());
             if (which == "inEdgesR") return ints_of(h.inEdgesR_h.get());
             if (which == "outEdgesR") return ints_of(h.outEdgesR_h.get());
             if (which == "edgeFeatures") return ints_of(h.edgeFeatures.get());
             if (which == "verticesFeatures") return ints_of(h.verticesFeatures.get());
             if (which == "Euid") return ints_of(h.Euid_h.get());
             throw std::runtime_error("unknown array " + which);
           })
      .def("__repr__", [](const data::HostGraphData &e) { return repr("<HostGraphData>\n", e); });

  py::class_<data::SubPartition>(d, "SubPartition", "Edge range [beg, end) of one time partition")
      .def_readonly("beg", &data::SubPartition::beg)
      .def_readonly("end", &data::SubPartition::end)
      .def_readonly("n", &data::SubPartition::n)
      .def("__repr__", [](const data::SubPartition &e) {
        return "<SubPartition: " + std::to_string(e.beg) + " " + std::to_string(e.end) + " " + std::to_string(e.n) +
               ">";
      });

  d.def("divideList", &data::divideList);
  d.def("setBuildDevice", &data::setBuildDevice, "GPU for the time sort and CSR build (-1 host, -2 auto)");
  d.def("buildDevice", &data::buildDevice);
  d.def("removeMinor", &data::removeMinor);
  d.def("includeMinor", &data::includeMinor);

  py::class_<data::GraphDataLoader>(d, "GraphDataLoader", "Builds HostGraphData (renaming, adjacency, caches) from a loaded edge list")
      .def(py::init<const data::EdgeListLoader *, const data::FeatureLoader *>(), py::keep_alive<1, 2>(),
           py::keep_alive<1, 3>())
      .def("createGraphData", &data::GraphDataLoader::createGraphData, py::arg("beg") = 0, py::arg("end") = 0,
           py::arg("cache") = true, py::call_guard<py::gil_scoped_release>())
      .def("createPartitionsData", &data::GraphDataLoader::createPartitionsData)
      .def("createPartitionsDataMajor", &data::GraphDataLoader::createPartitionsDataMajor)
      .def("createPartitionsDataMinor", &data::GraphDataLoader::createPartitionsDataMinor)
      .def("__repr__", [](const data::GraphDataLoader &e) { return "<GraphDataLoader: " + e.name() + ">"; });

  py::class_<data::DeviceGraphData>(d, "DeviceGraphData", "Host graph resident on one GPU")
      .def("__repr__", [](const data::DeviceGraphData &e) { return repr("<DeviceGraphData>\n", e); });

  py::class_<data::GraphDataManager>(d, "GraphDataManager", "Keeps host graphs resident on the GPUs it manages")
      .def("allocBatch", &data::GraphDataManager::allocBatch)
      .def("alloc", &data::GraphDataManager::alloc)
      .def("moveAsync", &data::GraphDataManager::moveAsync)
      .def("waitMove", &data::GraphDataManager::waitMove)
      .def("__repr__", [](const data::GraphDataManager &) { return std::string("<GraphDataManager>"); });

  d.def("MakeGraphDataManager", &data::MakeGraphDataManager);

  py::class_<data::HostMotifData>(d, "HostMotifData", "Motif file read, nodes renamed, per-edge expansion info")
      .def(py::init<const std::filesystem::path &>())
      .def_readonly("name", &data::HostMotifData::name)
      .def_readonly("numEdges", &data::HostMotifData::numEdges)
      .def_readonly("numVertices", &data::HostMotifData::numVertices)
      .def("edges",
           [](const data::HostMotifData &h) {
             std::vector<std::tuple<int, int, int>> out;
             for (auto &e : h.edges) out.emplace_back(e.u, e.v, e.t);
             return out;
           })
      .def("minfo",
           [](const data::HostMotifData &h) {
             std::vector<std::tuple<int, int, int, int, bool>> out;
             for (auto &mi : h.minfo) out.emplace_back(mi.baseNode, mi.constraintNode, mi.mappedNodes, mi.io, mi.newNode);
             return out;
           })
      .def("__repr__", [](const data::HostMotifData &e) { return repr("<HostMotifData>\n", e); });

  py::class_<data::DeviceJobData>(d, "DeviceJobData", "A motif bound to a resident graph")
      .def(py::init<const data::DeviceGraphData &, const data::HostMotifData &>())
      .def("__repr__", [](const data::DeviceJobData &e) { return repr("<DeviceJobData>\n", e); });

  py::class_<data::MineJob>(d, "MineJob", "Query job: graph + motif, delta and root range [beg, end)")
      .def(py::init<const data::DeviceJobData *, int, int, int>(), py::arg("_data"), py::arg("_delta"),
           py::arg("_beg") = 0, py::arg("_end") = -1, py::keep_alive<1, 2>())
      .def("getWork", &data::MineJob::getWork)
      .def("setWork", &data::MineJob::setWork)
      .def("__repr__", [](const data::MineJob &e) {
        return "<MineJob: " + std::to_string(e.device) + ", " + std::to_string(e.delta) + ", " +
               std::to_string(e.beg) + ", " + std::to_string(e.end) + ">";
      });

  py::class_<GPUWorker>(m, "GPUWorker", "Motif-mining worker of one GPU (plugin object)")
      .def("run", &GPUWorker::run, py::call_guard<py::gil_scoped_release>())
      .def("take", &GPUWorker::take, py::keep_alive<1, 2>())
      .def("update_job", &GPUWorker::update_job)
      .def("timed_run", &GPUWorker::timed_run, py::call_guard<py::gil_scoped_release>())
      .def("time", &GPUWorker::time)
      .def("count", &GPUWorker::count)
      .def("clear_time", &GPUWorker::clear_time)
      .def("pause", &GPUWorker::pause)
      .def("printEnum", &GPUWorker::printEnum)
      .def_readonly("type_", &GPUWorker::type_)
      .def("__repr__", [](const GPUWorker &w) { return repr("<GPUWorker>\n", w); });

  py::class_<GPUWorkerLibManager>(m, "GPUWorkerLibManager", "Opens worker plugin libraries and constructs their workers")
      .def(py::init<>())
      .def("openLib", &GPUWorkerLibManager::openLib)
      .def("unloadLib", &GPUWorkerLibManager::unloadLib)
      .def("construct", &GPUWorkerLibManager::construct)
      .def("__repr__", [](const GPUWorkerLibManager &e) { return repr("<GPUWorkerLibManager>\n", e); });

  py::class_<SingleGPUExecution>(m, "SingleGPUExecution", "Runs one worker over the whole root range on one GPU")
      .def(py::init<const data::GraphDataManager *, int>(), py::arg("dataManager"), py::arg("device") = 0,
           py::keep_alive<1, 2>())
      .def("setHostGraph", &SingleGPUExecution::setHostGraph, py::arg("hostGraph"), py::arg("n") = -1,
           py::keep_alive<1, 2>(), py::call_guard<py::gil_scoped_release>())
      .def("setHostMotif", &SingleGPUExecution::setHostMotif, py::keep_alive<1, 2>())
      .def("setWorker", &SingleGPUExecution::setWorker, py::keep_alive<1, 2>())
      .def("run", &SingleGPUExecution::run, py::call_guard<py::gil_scoped_release>())
      .def("__repr__", [](const SingleGPUExecution &) { return std::string("<SingleGPUExecution: one device>"); });

  py::class_<SingleGPUExecutionDiv>(m, "SingleGPUExecutionDiv", "Runs one worker over the root range split into div parts")
      .def(py::init<const data::GraphDataManager *, int, int>(), py::arg("dataManager"), py::arg("device") = 0,
           py::arg("div") = 64, py::keep_alive<1, 2>())
      .def("setHostGraph", &SingleGPUExecutionDiv::setHostGraph, py::arg("hostGraph"), py::arg("n") = -1,
           py::keep_alive<1, 2>())
      .def("setHostMotif", &SingleGPUExecutionDiv::setHostMotif, py::keep_alive<1, 2>())
      .def("setWorker", &SingleGPUExecutionDiv::setWorker, py::keep_alive<1, 2>())
      .def("run", &SingleGPUExecutionDiv::run, py::call_guard<py::gil_scoped_release>())
      .def("__repr__", [](const SingleGPUExecutionDiv &) { return std::string("<SingleGPUExecution: one device>"); });

  py::class_<SingleGPUExecutionPause>(m, "SingleGPUExecutionPause", "Runs one worker on one GPU (pausing is not needed here)")
      .def(py::init<const data::GraphDataManager *, int>(), py::arg("dataManager"), py::arg("device") = 0,
           py::keep_alive<1, 2>())
      .def("setHostGraph", &SingleGPUExecutionPause::setHostGraph, py::arg("hostGraph"), py::arg("n") = -1,
           py::keep_alive<1, 2>())
      .def("setHostMotif", &SingleGPUExecutionPause::setHostMotif, py::keep_alive<1, 2>())
      .def("setWorker", &SingleGPUExecutionPause::setWorker, py::keep_alive<1, 2>())
      .def("run", &SingleGPUExecutionPause::run, py::call_guard<py::gil_scoped_release>())
      .def("__repr__", [](const SingleGPUExecutionPause &) { return std::string("<SingleGPUExecution: one device>"); });

  m.def("SingleGPURunBatch", &SingleGPURunBatch);

#define BIND_MULTI(CLS, DOC)                                                                \
  py::class_<CLS>(m, #CLS, DOC)                                                             \
      .def(py::init<const data::GraphDataManager *, int>(), py::keep_alive<1, 2>())         \
      .def("setHostGraphs", &CLS::setHostGraphs, py::keep_alive<1, 2>())                    \
      .def("setHostMotif", &CLS::setHostMotif, py::keep_alive<1, 2>())                      \
      .def("setWorkers", &CLS::setWorkers)                                                  \
      .def("run", &CLS::run, py::call_guard<py::gil_scoped_release>())                      \
      .def("numDevice", &CLS::numDevice)                                                    \
      .def("__repr__", [](const CLS &e) { return "<" #CLS ": " + std::to_string(e.numDevice()) + ">"; });
  BIND_MULTI(MultiGPUExecutionNaive, "Static split of roots over GPUs (one GPU here)")
  BIND_MULTI(MultiGPUExecutionDyn, "Roots handed out to GPUs on demand (one GPU here)")
  BIND_MULTI(MultiGPUExecutionDynPause, "Roots handed out to GPUs on demand (one GPU here)")
#undef BIND_MULTI

  m.def("allCount", &allCount);

  auto gen = m.def_submodule("gen", "deterministic synthetic temporal graphs");
  auto arrays3 = [](size_t E) {
    return std::make_tuple(py::array_t<int>(E), py::array_t<int>(E), py::array_t<int>(E));
  };
  gen.def(
      "zipf_uniform",
      [arrays3](size_t N, size_t E, double alpha, int64_t span, uint64_t seed) {
        auto out = arrays3(E);
        int *U = std::get<0>(out).mutable_data(), *V = std::get<1>(out).mutable_data(),
            *T = std::get<2>(out).mutable_data();
        {
          py::gil_scoped_release rel;
          everest::gen::zipf_uniform(N, E, alpha, span, seed, U, V, T);
        }
        return out;
      },
      py::arg("N"), py::arg("E"), py::arg("alpha"), py::arg("span"), py::arg("seed"));
  gen.def(
      "powerlaw_bursty",
      [arrays3](size_t N, size_t E, double alpha, double sigma, int64_t span, uint64_t seed) {
        auto out = arrays3(E);
        int *U = std::get<0>(out).mutable_data(), *V = std::get<1>(out).mutable_data(),
            *T = std::get<2>(out).mutable_data();
        {
          py::gil_scoped_release rel;
          everest::gen::powerlaw_bursty(N, E, alpha, sigma, span, seed, U, V, T);
        }
        return out;
      },
      py::arg("N"), py::arg("E"), py::arg("alpha"), py::arg("sigma"), py::arg("span"), py::arg("seed"));
}
