// everest/data.h -- C++ API of the temporal-motif core library (libeverest_core.so).
//
// Same namespaces, class names, members and signatures as the reference's
// include/data.h (corelib, corelib::data), so code written against the
// reference's corelib compiles against this header.  Differences:
//   * no Boost / fmt / range-v3: mapped files use POSIX mmap, printing uses
//     std::ostream;
//   * device memory is allocated through the C ABI in tm_abi.h: DeviceGraphData
//     holds a resident tm_graph handle, and DeviceJobData exposes that graph's
//     device arrays under the reference's field names;
//   * GraphDataManager keeps uploaded graphs resident per device and hands the
//     same device copy to every execution of the same host graph;
//   * loader errors throw std::runtime_error instead of exiting.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <functional>
#include <memory>
#include <mutex>
#include <ostream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

#include "tm_abi.h"

namespace corelib::util {
// Fill a file with N ints in [0, 1e9] (reference: src/core/cpu/data.cc:22-55).
// Seeded (deterministic), where the reference seeds from std::random_device.
void generateRandToFile(size_t N, std::filesystem::path pathToFile, uint64_t seed = 1);
void memcpyPar(void *dst, const void *src, size_t count);
}  // namespace corelib::util

namespace corelib {

// include/data.h:27-29
struct TemporalEdge {
  int u, v, t;
};

bool operator==(const TemporalEdge &lhs, const TemporalEdge &rhs);

// include/data.h:33-40.  HostMotifData leaves arrR/arrV null; DeviceJobData
// binds them to the device out-CSR (io == 1) or in-CSR (otherwise), as
// src/core/gpu/data.cu:515-523 does.
struct MotifEdgeInfoV1 {
  int baseNode, constraintNode;
  int mappedNodes;
  const int *arrR;
  const int *arrV;
  int io;
  bool newNode;
};

std::ostream &operator<<(std::ostream &o, const MotifEdgeInfoV1 &obj);

}  // namespace corelib

namespace corelib::data {

// ---------------------------------------------------------------- memory
struct SourceMemory {
  virtual const void *data() const = 0;
  virtual size_t size() const = 0;
  virtual ~SourceMemory() {}
};

class MappedFileSourceMemory : public SourceMemory {
  void *ptr_ = nullptr;
  size_t size_ = 0;

 public:
  explicit MappedFileSourceMemory(std::filesystem::path pathToFile);
  const void *data() const override { return ptr_; }
  size_t size() const override { return size_; }
  ~MappedFileSourceMemory() override;
};

struct SinkMemory {
  virtual void *data() const = 0;
  virtual size_t size() const = 0;
  virtual ~SinkMemory() {}
};

class MappedFileSinkMemory : public SinkMemory {
  void *ptr_ = nullptr;
  size_t size_ = 0;

 public:
  MappedFileSinkMemory(std::filesystem::path pathToFile, size_t size);
  explicit MappedFileSinkMemory(std::filesystem::path pathToFile);
  void *data() const override { return ptr_; }
  size_t size() const override { return size_; }
  ~MappedFileSinkMemory() override;
};

class VectorSinkMemory : public SinkMemory {
  std::vector<uint8_t> mem_;

 public:
  explicit VectorSinkMemory(size_t size) : mem_(size) {}
  void *data() const override { return (void *)mem_.data(); }
  size_t size() const override { return mem_.size(); }
  ~VectorSinkMemory() override {}
};

// Source memory owning a vector (graphs built in memory, no cache files).
class VectorSourceMemory : public SourceMemory {
  std::vector<int> mem_;

 public:
  explicit VectorSourceMemory(std::vector<int> v) : mem_(std::move(v)) {}
  const void *data() const override { return mem_.data(); }
  size_t size() const override { return mem_.size() * sizeof(int); }
};

// ---------------------------------------------------------------- caches
// include/data.h:103-122: <graph-dir>/<stem>/{U,V,T,IDS}.cache and
// <stem>-<beg>-<end>/ per-range graph directories.
struct CacheDirectoryAccess {
  const std::string name_;
  const std::filesystem::path graphPath_, cacheDir_;
  explicit CacheDirectoryAccess(std::filesystem::path pathToGraph, bool createDir = true);

  std::filesystem::path cacheUPath() const { return cacheDir_ / "U.cache"; }
  std::filesystem::path cacheVPath() const { return cacheDir_ / "V.cache"; }
  std::filesystem::path cacheTPath() const { return cacheDir_ / "T.cache"; }
  std::filesystem::path cacheIDSPath() const { return cacheDir_ / "IDS.cache"; }
  std::filesystem::path cacheFEPath() const { return cacheDir_ / "EdgesFeatures"; }
  std::filesystem::path cacheFVVPath() const { return cacheDir_ / "VerticesFeaturesValue"; }
  std::filesystem::path cacheFVMPath() const { return cacheDir_ / "VerticesFeaturesMap"; }
  std::filesystem::path cacheGraphDirPath(size_t beg, size_t end) const;
  std::string name() const { return name_; }
};

// Edge list, stably sorted by time, with IDS = original line numbers
// (reference: src/core/cpu/data.cc:175-310).
class EdgeListLoader : public CacheDirectoryAccess {
  std::unique_ptr<SourceMemory> U_, V_, T_, IDS_;
  const int thread_ = 64;
  bool inMemory_ = false;

  bool shouldInit() const;
  void createEdgeListCache();
  void loadEdgeListCache();

 public:
  explicit EdgeListLoader(std::filesystem::path pathToGraph);
  // In-memory edge list (no file, no caches): U, V, T in line order.
  EdgeListLoader(const std::string &name, const int *U, const int *V, const int *T, size_t n);

  size_t edgeListLength() const;
  bool inMemory() const { return inMemory_; }
  const SourceMemory *getUPtr() const { return U_.get(); }
  const SourceMemory *getVPtr() const { return V_.get(); }
  const SourceMemory *getTPtr() const { return T_.get(); }
  const SourceMemory *getIDSPtr() const { return IDS_.get(); }
};

// Random labels for edges and vertices (reference: data.cu:57-79,
// data.cc:316-343).  When the label file does not exist, labels are drawn
// from a seeded generator instead of failing.
class FeatureLoader : public CacheDirectoryAccess {
  std::unique_ptr<SourceMemory> edgeFeatures_, verticesFeatures_, verticesMap_;
  std::filesystem::path pathToRands_;
  const EdgeListLoader *eloader_;

  bool shouldInit() const;
  void createFeatureCache();
  void loadFeatureCache();

 public:
  FeatureLoader(const EdgeListLoader *eloader, std::filesystem::path pathToRands);

  size_t edgeFeaturesLength() const;
  size_t verticesFeaturesLength() const;
  const SourceMemory *getFEPtr() const { return edgeFeatures_.get(); }
  const SourceMemory *getFVVPtr() const { return verticesFeatures_.get(); }
  const SourceMemory *getFVMPtr() const { return verticesMap_.get(); }
};

// include/data.h:174-186
struct HostGraphData {
  std::string name;
  size_t beg{0}, end{0};
  size_t numEdges{0}, numVertices{0}, totalBytes{0};
  std::unique_ptr<SinkMemory> Eg_h, inEdgesV_h, outEdgesV_h;
  std::unique_ptr<SinkMemory> inEdgesR_h, outEdgesR_h;
  std::unique_ptr<SinkMemory> edgeFeatures, verticesFeatures;
  std::unique_ptr<SinkMemory> Euid_h;
  uint64_t uid{0};  // identity used by the device-residency cache

  HostGraphData() = default;
  HostGraphData(HostGraphData &&) = default;
  HostGraphData &operator=(HostGraphData &&) = default;
};

std::ostream &operator<<(std::ostream &o, const HostGraphData &obj);

size_t calcTotalBytes(size_t E, size_t V);

struct SubPartition {
  int beg, end, n;
};

typedef std::vector<SubPartition> SubPartitionList;
SubPartitionList divideList(const EdgeListLoader &eloader, int delta, int num);

// Where the stable time sort (EdgeListLoader) and the renaming + CSR build
// (GraphDataLoader::createGraphData) run: a GPU index (tm_time_order /
// tm_csr_build), -1 for the host, -2 (the default) for GPU 0 when one is
// visible and the host otherwise.  The arrays are the same either way.
void setBuildDevice(int device);
int buildDevice();

typedef std::vector<HostGraphData> HostGraphDataList;
void removeMinor(HostGraphDataList &complete);
void includeMinor(HostGraphDataList &major, HostGraphDataList &minor);

// Renaming + in/out CSR of an edge-list range (reference: data.cu:124-348).
class GraphDataLoader : public CacheDirectoryAccess {
  std::string stem_;
  const EdgeListLoader *eloader_;
  const SourceMemory *UPtr_, *VPtr_, *TPtr_;
  const SourceMemory *FEPtr_, *FVVPtr_, *FVMPtr_;
  const SourceMemory *EuidPtr_;
  size_t edgeListLength_, verticesFeatureLength_;

  HostGraphData loadGraphData(size_t beg, size_t end) const;
  HostGraphData createGraphDataOnDevice(int device, size_t beg, size_t end, bool cache) const;
  bool isCached(size_t beg, size_t end) const;

 public:
  GraphDataLoader(const EdgeListLoader *eloader, const FeatureLoader *floader);
  std::string name() const { return stem_; }

  HostGraphData createGraphData(size_t beg = 0, size_t end = 0, bool cache = true) const;
  HostGraphDataList createPartitionsData(const SubPartitionList &partitions) const;
  HostGraphDataList createPartitionsDataMajor(const SubPartitionList &partitions) const;
  HostGraphDataList createPartitionsDataMinor(const SubPartitionList &partitions) const;
};

// ---------------------------------------------------------------- device
// A resident device graph (shared: several executions may use one copy).
struct DeviceGraphData {
  int device{0};
  std::string name;
  size_t beg{0}, end{0};
  size_t numEdges{0}, numVertices{0}, totalBytes{0};
  std::shared_ptr<tm_graph> graph;
};

std::ostream &operator<<(std::ostream &o, const DeviceGraphData &obj);

struct GraphDataManager {
  virtual std::vector<DeviceGraphData> allocBatch(int allDevices,
                                                  const std::vector<const HostGraphData *> hostGraphs) const = 0;
  virtual DeviceGraphData alloc(int device, const HostGraphData *hostGraph) const = 0;
  virtual void moveAsync(DeviceGraphData &deviceGraph, const HostGraphData &hostGraph, int device) const = 0;
  virtual void waitMove(int device) const = 0;
  virtual ~GraphDataManager() {}
};

std::unique_ptr<GraphDataManager> MakeGraphDataManager(int totalDevice);

// include/data.h:284-292 (reference: src/core/cpu/data.cc:411-449)
struct HostMotifData {
  std::string name;
  size_t numEdges{0}, numVertices{0};
  std::vector<TemporalEdge> edges;
  std::vector<MotifEdgeInfoV1> minfo;

  explicit HostMotifData(const std::filesystem::path &pathToMotif);
  HostMotifData() = default;
  static HostMotifData fromEdges(const std::string &name, const std::vector<TemporalEdge> &rawEdges);
};

std::ostream &operator<<(std::ostream &o, const HostMotifData &obj);

// include/data.h:304-320: the graph on a device plus one motif.  The device
// pointers are the resident graph's arrays (valid while `graph` lives); Em_d()
// and minfo() are device copies of the motif edges and of minfol_, whose
// arrR/arrV point at the device CSR (src/core/gpu/data.cu:496-529), so a
// worker kernel written against the reference's data.h reads the same fields.
struct DeviceJobData {
  int device{0};
  size_t graphNumEdges{0}, graphNumVertices{0};
  const TemporalEdge *Eg_d{nullptr};
  const int *inEdgesV_d{nullptr}, *outEdgesV_d{nullptr}, *inEdgesR_d{nullptr}, *outEdgesR_d{nullptr};
  const int *edgefeatures_d{nullptr}, *nodefeatures_d{nullptr};
  const int *Euid_d{nullptr};
  std::shared_ptr<tm_graph> graph;

  size_t motifNumEdges{0}, motifNumVertices{0};
  std::shared_ptr<void> Em_d_, minfo_;  // device copies (freed with the last holder)
  std::vector<TemporalEdge> motif;      // host copy of the motif edges
  std::vector<MotifEdgeInfoV1> minfol_; // host copy of minfo(), device CSR pointers bound

  const TemporalEdge *Em_d() const { return (const TemporalEdge *)Em_d_.get(); }
  const MotifEdgeInfoV1 *minfo() const { return (const MotifEdgeInfoV1 *)minfo_.get(); }
  std::vector<int32_t> motifSrc() const;
  std::vector<int32_t> motifDst() const;

  DeviceJobData(const DeviceGraphData &deviceGraph, const HostMotifData &hostMotif);
  DeviceJobData() = default;
};

std::ostream &operator<<(std::ostream &o, const DeviceJobData &obj);

// include/data.h:327-336 (reference: src/core/cpu/data.cc:452-466)
struct MineJob {
  int device{-1};
  int beg{-1}, end{-1}, delta{-1};
  const DeviceJobData *data{nullptr};

  MineJob(const DeviceJobData *_data, int _delta, int _beg = 0, int _end = -1);
  MineJob() = default;
  std::tuple<int, int> getWork() const;
  void setWork(int _beg, int _end);
};

}  // namespace corelib::data

namespace corelib {

// include/data.h:342-366 (reference: src/core/cpu/data.cc:529-569)
struct GPUWorker {
  std::string type_;
  int gpu_;
  unsigned long long time_;
  int sizeBlock_;
  unsigned long long count_;

  data::MineJob *job_ = nullptr;

  GPUWorker(std::string type, int gpu = 0, int sizeBlock = 64);
  virtual ~GPUWorker();
  virtual unsigned long long run() = 0;
  virtual void take(data::MineJob &job);
  virtual void update_job() = 0;
  int numBlocks();

  unsigned long long timed_run();
  unsigned long long time();
  unsigned long long count();
  void clear_time();

  virtual std::vector<int> printEnum(int n);
  virtual std::pair<int, int> pause();
};

std::ostream &operator<<(std::ostream &o, const GPUWorker &w);

// The library's stock worker: counts (or enumerates up to enumCap rows) with
// the table-driven kernels behind tm_abi.h.  Generated plugins derive from
// GPUWorker directly.
struct GPUWorkerTM : public GPUWorker {
  long long enumCap_;
  tm_tuning tune_;
  std::vector<int> rows_;

  explicit GPUWorkerTM(int gpu, long long enumCap = -1);
  unsigned long long run() override;
  void update_job() override;
  std::vector<int> printEnum(int n) override;
};

// include/data.h:371-384 (reference: src/core/cpu/data.cc:571-609)
struct LibHandle {
  typedef GPUWorker *(*FactoryType)(int);

  void *handle;
  FactoryType getWorker;
  LibHandle(void *h, FactoryType f);
  explicit LibHandle(const std::string &libpath);

  LibHandle(LibHandle &&rhs);
  LibHandle &operator=(LibHandle &&rhs);
  ~LibHandle();

  std::string toString() const;
};

std::ostream &operator<<(std::ostream &o, const LibHandle &obj);

// include/data.h:388-396 (reference: src/core/cpu/data.cc:611-640, data.cu:540-544)
class GPUWorkerLibManager {
  std::unordered_map<std::string, LibHandle> table_;

 public:
  void openLib(const std::string &libpath);
  void unloadLib(const std::string &lib);
  std::unique_ptr<GPUWorker> construct(const std::string &lib, int gpu);
  std::string toString() const;
};

std::ostream &operator<<(std::ostream &o, const GPUWorkerLibManager &obj);

// include/data.h:407-427 (reference: src/core/cpu/data.cc:642-677)
class SingleGPUExecution {
 protected:
  int device_{0};
  GPUWorker *w_{nullptr};
  const data::HostGraphData *hostGraph_{nullptr};
  int n_{0};
  const data::HostMotifData *hostMotif_{nullptr};
  const data::GraphDataManager *dataManager_{nullptr};

  data::DeviceJobData jobData_;
  data::MineJob job_;
  data::DeviceGraphData deviceGraph_;

  virtual void formJobData();

 public:
  SingleGPUExecution(const data::GraphDataManager *dataManager, int device = 0);
  virtual ~SingleGPUExecution() {}
  void setHostGraph(const data::HostGraphData *hostGraph, int n = -1);
  void setHostMotif(const data::HostMotifData *hostMotif);
  void setWorker(GPUWorker *w);
  virtual unsigned long long run(int delta);
};

// include/data.h:432-442 (reference: src/core/cpu/data.cc:679-703)
class SingleGPUExecutionDiv : public SingleGPUExecution {
 protected:
  void formJobData() override;
  int div_;
  struct Task {
    int beg, end;
  };
  std::vector<Task> queue_;

 public:
  SingleGPUExecutionDiv(const data::GraphDataManager *dataManager, int device = 0, int div = 64);
  unsigned long long run(int delta) override;
};

void SingleGPURunBatch(const data::GraphDataManager *dataManager, data::HostGraphDataList &hostGraphs,
                       const data::HostMotifData *hostMotif, GPUWorker *w, int delta);

// Multi-GPU executions (include/data.h:448-541).  This build drives one GPU:
// with deviceNum == 1 they run the partitions on device 0; with more devices
// they throw std::runtime_error.
class MultiGPUExecution {
 protected:
  int deviceNum_{0};
  std::vector<GPUWorker *> ws_;
  std::vector<const data::HostGraphData *> hostGraphs_;
  data::SubPartitionList ps_;
  const data::HostMotifData *hostMotif_{nullptr};
  const data::GraphDataManager *dataManager_{nullptr};

 public:
  MultiGPUExecution(const data::GraphDataManager *dataManager, int deviceNum);
  virtual ~MultiGPUExecution() {}
  void setHostGraphs(const data::HostGraphDataList &hostGraphs, const data::SubPartitionList &ps);
  void setHostMotif(const data::HostMotifData *hostMotif);
  void setWorkers(std::vector<GPUWorker *> &ws);
  virtual void run(int delta);
  int numDevice() const;
};

class MultiGPUExecutionNaive : public MultiGPUExecution {
 public:
  using MultiGPUExecution::MultiGPUExecution;
};

class MultiGPUExecutionDyn : public MultiGPUExecution {
 public:
  using MultiGPUExecution::MultiGPUExecution;
};

class MultiGPUExecutionDynPause : public MultiGPUExecutionDyn {
 public:
  using MultiGPUExecutionDyn::MultiGPUExecutionDyn;
};

// Runs like SingleGPUExecution (the reference marks it reserved, data.cc:895-918).
class SingleGPUExecutionPause : public SingleGPUExecution {
 public:
  using SingleGPUExecution::SingleGPUExecution;
};

unsigned long long allCount(std::vector<GPUWorker *> &ws);

}  // namespace corelib
