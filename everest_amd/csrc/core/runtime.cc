// runtime.cc -- device residency, jobs, workers, plugin libraries and the
// single-GPU executions of the core library.
//
// Reference counterparts: src/core/gpu/data.cu:364-544 (device memory,
// GraphDataManager, DeviceJobData, construct) and src/core/cpu/data.cc:452-1007
// (MineJob, GPUWorker, LibHandle, GPUWorkerLibManager, executions).  Device
// work goes through the C ABI of libtm_hip.so (include/tm_abi.h).
#include <dlfcn.h>

#include <chrono>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "everest/data.h"

namespace corelib::data {

static void tm_throw(tm_status s, const char *what) {
  if (s != TM_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string((int)s) + "): " + tm_last_error());
}

std::ostream &operator<<(std::ostream &o, const DeviceGraphData &obj) {
  o << "device: " << obj.device << ", name: " << obj.name << "\n"
    << "beg: " << obj.beg << ", end: " << obj.end << "\n"
    << "numEdges: " << obj.numEdges << ", numVertices: " << obj.numVertices << ", totalBytes: " << obj.totalBytes
    << "\n"
    << "graph: " << (const void *)obj.graph.get();
  return o;
}

// Keeps every uploaded host graph resident per device for the manager's
// lifetime, so executions of many queries on one graph upload it once.
class GraphDataManagerImpl : public GraphDataManager {
  int totalDevice_;
  mutable std::mutex mtx_;
  mutable std::map<std::pair<int, uint64_t>, std::shared_ptr<tm_graph>> resident_;

  static DeviceGraphData describe(int device, const HostGraphData *h) {
    DeviceGraphData d;
    d.device = device;
    d.name = h->name;
    d.beg = h->beg;
    d.end = h->end;
    d.numEdges = h->numEdges;
    d.numVertices = h->numVertices;
    d.totalBytes = h->totalBytes;
    return d;
  }

 public:
  explicit GraphDataManagerImpl(int totalDevice) : totalDevice_(totalDevice) {
    if (totalDevice < 1) throw std::runtime_error("MakeGraphDataManager needs at least one device");
  }

  std::vector<DeviceGraphData> allocBatch(int allDevices,
                                          const std::vector<const HostGraphData *> hostGraphs) const override {
    std::vector<DeviceGraphData> res;
    if (hostGraphs.empty()) return res;
    for (int d = 0; d < allDevices; d++) res.push_back(describe(d, hostGraphs.front()));
    return res;
  }

  DeviceGraphData alloc(int device, const HostGraphData *hostGraph) const override {
    if (!hostGraph) throw std::runtime_error("alloc: null host graph");
    if (device < 0 || device >= totalDevice_) throw std::runtime_error("alloc: device out of range");
    DeviceGraphData d = describe(device, hostGraph);
    std::lock_guard<std::mutex> lk(mtx_);
    auto it = resident_.find({device, hostGraph->uid});
    if (it != resident_.end()) d.graph = it->second;
    return d;
  }

  void moveAsync(DeviceGraphData &s, const HostGraphData &h, int device) const override {
    std::lock_guard<std::mutex> lk(mtx_);
    auto key = std::make_pair(device, h.uid);
    auto it = resident_.find(key);
    if (it == resident_.end()) {
      tm_host_graph hg;
      hg.num_edges = (int32_t)h.numEdges;
      hg.num_vertices = (int32_t)h.numVertices;
      hg.Eg = (const int32_t *)h.Eg_h->data();
      hg.inEdgesV = (const int32_t *)h.inEdgesV_h->data();
      hg.outEdgesV = (const int32_t *)h.outEdgesV_h->data();
      hg.inEdgesR = (const int32_t *)h.inEdgesR_h->data();
      hg.outEdgesR = (const int32_t *)h.outEdgesR_h->data();
      hg.Euid = (const int32_t *)h.Euid_h->data();
      tm_graph *g = nullptr;
      tm_throw(tm_graph_upload(device, &hg, &g), "tm_graph_upload");
      std::shared_ptr<tm_graph> sg(g, [](tm_graph *p) { tm_graph_free(p); });
      // labels travel with the graph, as the reference's moveAsync copies them
      const size_t ne = h.edgeFeatures ? h.edgeFeatures->size() / sizeof(int32_t) : 0;
      const size_t nv = h.verticesFeatures ? h.verticesFeatures->size() / sizeof(int32_t) : 0;
      if ((ne == h.numEdges && ne > 0) || nv > 0)
        tm_throw(tm_graph_set_features(g, ne == h.numEdges ? (const int32_t *)h.edgeFeatures->data() : nullptr,
                                       ne == h.numEdges ? (int32_t)ne : 0,
                                       nv ? (const int32_t *)h.verticesFeatures->data() : nullptr, (int32_t)nv),
                 "tm_graph_set_features");
      it = resident_.emplace(key, std::move(sg)).first;
    }
    s = describe(device, &h);
    s.graph = it->second;
  }

  void waitMove(int) const override {}
};

std::unique_ptr<GraphDataManager> MakeGraphDataManager(int totalDevice) {
  return std::unique_ptr<GraphDataManager>(new GraphDataManagerImpl(totalDevice));
}

// A device copy owned by a shared pointer that frees it on the graph's device
// (the graph handle is held by the deleter, so it outlives the copy).
static std::shared_ptr<void> device_copy(const std::shared_ptr<tm_graph> &g, const void *host, size_t bytes) {
  void *p = nullptr;
  tm_throw(tm_device_copy(g.get(), host, (int64_t)bytes, &p), "tm_device_copy");
  return std::shared_ptr<void>(p, [g](void *q) { tm_device_free(g.get(), q); });
}

DeviceJobData::DeviceJobData(const DeviceGraphData &deviceGraph, const HostMotifData &hostMotif) {
  device = deviceGraph.device;
  graphNumEdges = deviceGraph.numEdges;
  graphNumVertices = deviceGraph.numVertices;
  graph = deviceGraph.graph;
  motifNumEdges = hostMotif.numEdges;
  motifNumVertices = hostMotif.numVertices;
  motif = hostMotif.edges;
  minfol_ = hostMotif.minfo;
  if (!graph) return;  // not resident yet (alloc without moveAsync): no device view
  tm_device_arrays a;
  tm_throw(tm_graph_arrays(graph.get(), &a), "tm_graph_arrays");
  Eg_d = (const TemporalEdge *)a.Eg;
  inEdgesV_d = a.inEdgesV;
  outEdgesV_d = a.outEdgesV;
  inEdgesR_d = a.inEdgesR;
  outEdgesR_d = a.outEdgesR;
  edgefeatures_d = a.edgeFeatures;
  nodefeatures_d = a.verticesFeatures;
  Euid_d = a.Euid;
  for (auto &mi : minfol_) {
    mi.arrR = mi.io == 1 ? outEdgesR_d : inEdgesR_d;
    mi.arrV = mi.io == 1 ? outEdgesV_d : inEdgesV_d;
  }
  Em_d_ = device_copy(graph, motif.data(), sizeof(TemporalEdge) * motif.size());
  minfo_ = device_copy(graph, minfol_.data(), sizeof(MotifEdgeInfoV1) * minfol_.size());
}

std::vector<int32_t> DeviceJobData::motifSrc() const {
  std::vector<int32_t> s;
  for (auto &e : motif) s.push_back(e.u);
  return s;
}

std::vector<int32_t> DeviceJobData::motifDst() const {
  std::vector<int32_t> s;
  for (auto &e : motif) s.push_back(e.v);
  return s;
}

std::ostream &operator<<(std::ostream &o, const DeviceJobData &obj) {
  o << "device: " << obj.device << "\n[Graph]\n"
    << "numEdges: " << obj.graphNumEdges << ", numVertices: " << obj.graphNumVertices << "\n"
    << "graph: " << (const void *)obj.graph.get() << "\n[Motif]\n"
    << "numEdges: " << obj.motifNumEdges << ", numVertices: " << obj.motifNumVertices;
  for (auto &mi : obj.minfol_) o << "\n" << mi;
  return o;
}

MineJob::MineJob(const DeviceJobData *_data, int _delta, int _beg, int _end)
    : device(_data->device), beg(_beg), end(_end), delta(_delta), data(_data) {
  if (end == -1) end = (int)(_data->graphNumEdges - _data->motifNumEdges + 1);
}

std::tuple<int, int> MineJob::getWork() const { return {beg, end}; }

void MineJob::setWork(int _beg, int _end) {
  beg = _beg;
  end = _end;
}

}  // namespace corelib::data

namespace corelib {

GPUWorker::GPUWorker(std::string type, int gpu, int sizeBlock)
    : type_(std::move(type)), gpu_(gpu), time_(0), sizeBlock_(sizeBlock), count_(0) {}

GPUWorker::~GPUWorker() {}

unsigned long long GPUWorker::timed_run() {
  using namespace std::chrono;
  auto st = high_resolution_clock::now();
  auto res = run();
  auto ed = high_resolution_clock::now();
  time_ += (unsigned long long)duration_cast<microseconds>(ed - st).count();
  return res;
}

void GPUWorker::take(data::MineJob &job) { job_ = &job; }
unsigned long long GPUWorker::time() { return time_; }
unsigned long long GPUWorker::count() { return count_; }
void GPUWorker::clear_time() { time_ = 0; }

int GPUWorker::numBlocks() {
  if (!job_) return -1;
  return (job_->end + sizeBlock_ - 1) / sizeBlock_;
}

std::vector<int> GPUWorker::printEnum(int) { return {}; }
std::pair<int, int> GPUWorker::pause() { return {-1, -1}; }

std::ostream &operator<<(std::ostream &o, const GPUWorker &w) {
  o << "type: " << w.type_ << ", gpu: " << w.gpu_ << ", job:" << (const void *)w.job_ << "\n";
  return o;
}

GPUWorkerTM::GPUWorkerTM(int gpu, long long enumCap) : GPUWorker("TM", gpu), enumCap_(enumCap) {
  memset(&tune_, 0, sizeof(tune_));
}

unsigned long long GPUWorkerTM::run() {
  if (!job_ || !job_->data) throw std::runtime_error("GPUWorker::run without a job");
  const data::DeviceJobData *d = job_->data;
  if (!d->graph) throw std::runtime_error("GPUWorker::run: job has no device graph");
  auto src = d->motifSrc();
  auto dst = d->motifDst();
  const int M = (int)src.size();
  uint64_t c = 0;
  if (enumCap_ < 0) {
    data::tm_throw(tm_count(d->graph.get(), M, src.data(), dst.data(), job_->delta, job_->beg, job_->end, &tune_, &c),
                   "tm_count");
  } else {
    long long have = (long long)rows_.size() / M;
    long long cap = std::max(0ll, enumCap_ - have);
    std::vector<int32_t> rows((size_t)std::max(1ll, cap) * M);
    int64_t nw = 0;
    data::tm_throw(tm_enum(d->graph.get(), M, src.data(), dst.data(), job_->delta, job_->beg, job_->end, cap, &tune_,
                           rows.data(), &c, &nw),
                   "tm_enum");
    rows_.insert(rows_.end(), rows.begin(), rows.begin() + (size_t)nw * M);
  }
  count_ += c;  // cumulative over the worker's lifetime, as the reference's count_d
  return count_;
}

void GPUWorkerTM::update_job() { job_->beg = job_->end; }

std::vector<int> GPUWorkerTM::printEnum(int n) {
  if (!job_ || !job_->data) return {};
  size_t M = job_->data->motifNumEdges;
  size_t rows = std::min<size_t>((size_t)std::max(n, 0), std::min<size_t>(count_, rows_.size() / std::max<size_t>(M, 1)));
  return std::vector<int>(rows_.begin(), rows_.begin() + rows * M);
}

LibHandle::LibHandle(void *h, FactoryType f) : handle(h), getWorker(f) {}

LibHandle::LibHandle(const std::string &libpath) {
  handle = dlopen(libpath.c_str(), RTLD_LAZY);
  if (!handle) {
    std::string msg = std::string("Load Lib Error: ") + dlerror();
    std::cerr << msg << std::endl;
    throw std::runtime_error(msg);
  }
  getWorker = (LibHandle::FactoryType)dlsym(handle, "getWorker");
  if (!getWorker) {
    std::string msg = std::string("query func Error: ") + dlerror();
    dlclose(handle);
    handle = nullptr;
    std::cerr << msg << std::endl;
    throw std::runtime_error(msg);
  }
}

LibHandle::LibHandle(LibHandle &&rhs) : handle(nullptr), getWorker(nullptr) { *this = std::move(rhs); }

LibHandle &LibHandle::operator=(LibHandle &&rhs) {
  if (this != &rhs) {
    if (handle) dlclose(handle);
    handle = rhs.handle;
    getWorker = rhs.getWorker;
    rhs.handle = nullptr;
    rhs.getWorker = nullptr;
  }
  return *this;
}

std::string LibHandle::toString() const {
  std::stringstream s;
  s << "handle: " << handle << ", factory: " << (void *)getWorker;
  return s.str();
}

LibHandle::~LibHandle() {
  if (handle) dlclose(handle);
}

std::ostream &operator<<(std::ostream &o, const LibHandle &obj) { return o << obj.toString(); }

void GPUWorkerLibManager::openLib(const std::string &libpath) {
  std::string name;
  std::size_t pos = libpath.find_last_of('/');
  std::string file = pos == std::string::npos ? libpath : libpath.substr(pos + 1);
  if (file.size() < 6 || file.compare(0, 3, "lib") != 0 || file.compare(file.size() - 3, 3, ".so") != 0)
    throw std::runtime_error("plugin file must be named lib<NAME>.so: " + libpath);
  name = file.substr(3, file.size() - 6);
  if (table_.find(name) != table_.end()) unloadLib(name);
  table_.emplace(name, LibHandle(libpath));
  std::cout << "#" << name << " is opened" << std::endl;
}

void GPUWorkerLibManager::unloadLib(const std::string &lib) {
  table_.erase(lib);
  std::cout << "#" << lib << " is unloaded" << std::endl;
}

std::unique_ptr<GPUWorker> GPUWorkerLibManager::construct(const std::string &lib, int gpu) {
  auto it = table_.find(lib);
  if (it == table_.end()) throw std::out_of_range("no plugin library named " + lib);
  return std::unique_ptr<GPUWorker>(it->second.getWorker(gpu));
}

std::string GPUWorkerLibManager::toString() const {
  std::stringstream s;
  bool first = true;
  for (auto &e : table_) {
    if (!first) s << "\n";
    s << e.first << ": " << e.second.toString();
    first = false;
  }
  return s.str();
}

std::ostream &operator<<(std::ostream &o, const GPUWorkerLibManager &obj) { return o << obj.toString(); }

SingleGPUExecution::SingleGPUExecution(const data::GraphDataManager *dataManager, int device)
    : device_(device), dataManager_(dataManager) {}

void SingleGPUExecution::formJobData() {
  if (hostGraph_ && hostMotif_) jobData_ = data::DeviceJobData(deviceGraph_, *hostMotif_);
}

void SingleGPUExecution::setHostGraph(const data::HostGraphData *hostGraph, int n) {
  n_ = n;
  hostGraph_ = hostGraph;
  deviceGraph_ = dataManager_->alloc(device_, hostGraph_);
  dataManager_->moveAsync(deviceGraph_, *hostGraph_, device_);
  dataManager_->waitMove(device_);
  formJobData();
}

void SingleGPUExecution::setHostMotif(const data::HostMotifData *hostMotif) {
  hostMotif_ = hostMotif;
  formJobData();
}

void SingleGPUExecution::setWorker(GPUWorker *w) { w_ = w; }

unsigned long long SingleGPUExecution::run(int delta) {
  if (!w_) throw std::runtime_error("SingleGPUExecution::run without a worker");
  if (!hostGraph_ || !hostMotif_) throw std::runtime_error("SingleGPUExecution::run without graph and motif");
  job_ = data::MineJob(&jobData_, delta, 0, n_);
  w_->take(job_);
  return w_->timed_run();
}

SingleGPUExecutionDiv::SingleGPUExecutionDiv(const data::GraphDataManager *dataManager, int device, int div)
    : SingleGPUExecution(dataManager, device), div_(div) {}

void SingleGPUExecutionDiv::formJobData() {
  if (hostGraph_ && hostMotif_) {
    jobData_ = data::DeviceJobData(deviceGraph_, *hostMotif_);
    queue_.clear();
    int n = n_ >= 0 ? n_ : (int)(jobData_.graphNumEdges - jobData_.motifNumEdges + 1);
    int chunk = n / std::max(div_, 1);
    for (int i = 0; i < div_; i++) queue_.push_back(Task{i * chunk, (i + 1) * chunk});
    if (!queue_.empty()) queue_.back().end = n;
  }
}

// Returns the number of matches over all pieces (the reference sums the
// worker's cumulative counts, data.cc:697-701, which overstates it).
unsigned long long SingleGPUExecutionDiv::run(int delta) {
  if (!w_) throw std::runtime_error("SingleGPUExecutionDiv::run without a worker");
  std::cout << "[SingleGPUExecutionDiv] div: " << queue_.size() << std::endl;
  unsigned long long before = w_->count();
  for (auto &t : queue_) {
    job_ = data::MineJob(&jobData_, delta, t.beg, t.end);
    w_->take(job_);
    w_->timed_run();
  }
  return w_->count() - before;
}

void SingleGPURunBatch(const data::GraphDataManager *dataManager, data::HostGraphDataList &hostGraphs,
                       const data::HostMotifData *hostMotif, GPUWorker *w, int delta) {
  for (auto &hostGraph : hostGraphs) {
    SingleGPUExecution exec(dataManager);
    exec.setHostGraph(&hostGraph);
    exec.setHostMotif(hostMotif);
    exec.setWorker(w);
    exec.run(delta);
  }
}

MultiGPUExecution::MultiGPUExecution(const data::GraphDataManager *dataManager, int deviceNum)
    : deviceNum_(deviceNum), dataManager_(dataManager) {}

void MultiGPUExecution::setHostGraphs(const data::HostGraphDataList &hostGraphs, const data::SubPartitionList &ps) {
  hostGraphs_.clear();
  for (auto &h : hostGraphs) hostGraphs_.push_back(&h);
  ps_ = ps;
}

void MultiGPUExecution::setHostMotif(const data::HostMotifData *hostMotif) { hostMotif_ = hostMotif; }
void MultiGPUExecution::setWorkers(std::vector<GPUWorker *> &ws) { ws_ = ws; }
int MultiGPUExecution::numDevice() const { return deviceNum_; }

// One device: each partition's seeds [0, n) are mined on device 0 by worker 0.
void MultiGPUExecution::run(int delta) {
  if (deviceNum_ != 1)
    throw std::runtime_error("this build drives one GPU; multi-GPU executions need deviceNum == 1");
  if (ws_.empty() || !hostMotif_) throw std::runtime_error("MultiGPUExecution::run without workers or motif");
  for (size_t i = 0; i < hostGraphs_.size(); i++) {
    SingleGPUExecution exec(dataManager_, 0);
    int n = i < ps_.size() ? ps_[i].n : -1;
    exec.setHostGraph(hostGraphs_[i], n);
    exec.setHostMotif(hostMotif_);
    exec.setWorker(ws_[0]);
    exec.run(delta);
  }
}

unsigned long long allCount(std::vector<GPUWorker *> &ws) {
  unsigned long long res = 0;
  for (auto &w : ws) res += w->count();
  return res;
}

}  // namespace corelib
