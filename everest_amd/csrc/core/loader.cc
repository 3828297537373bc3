// loader.cc -- edge-list loader and graph preprocessor (host, C++17).
//
// Follows the reference loader's contract (src/core/cpu/data.cc:114-310,
// src/core/gpu/data.cu:57-348): parse "u v t" lines, stable sort by t keeping
// IDS = original line numbers, rename vertices by first occurrence in the
// sorted list (source before destination on the same edge), build the in/out
// CSR with ascending positions, and keep the same cache files.  The
// algorithms are O(E) passes (LSD radix sort on the time key, counting sorts
// for the CSR) rather than the reference's comparison sorts.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "tm_abi.h"
#include "everest/data.h"

namespace corelib::util {

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

void generateRandToFile(size_t N, std::filesystem::path pathToFile, uint64_t seed) {
  data::MappedFileSinkMemory sink(pathToFile, sizeof(int) * N);
  int *r = (int *)sink.data();
  for (size_t i = 0; i < N; i++) r[i] = (int)(splitmix64(seed * 0x100000001b3ull + i) % 1000000001ull);
}

void memcpyPar(void *dst, const void *src, size_t count) {
  const int nt = 8;
  std::vector<std::thread> ts;
  size_t chunk = count / nt;
  for (int i = 0; i < nt; i++) {
    size_t b = chunk * i, e = (i == nt - 1) ? count : chunk * (i + 1);
    ts.emplace_back([=] { memcpy((char *)dst + b, (const char *)src + b, e - b); });
  }
  for (auto &t : ts) t.join();
}

}  // namespace corelib::util

namespace corelib {

bool operator==(const TemporalEdge &lhs, const TemporalEdge &rhs) {
  return lhs.u == rhs.u && lhs.v == rhs.v && lhs.t == rhs.t;
}

std::ostream &operator<<(std::ostream &o, const MotifEdgeInfoV1 &obj) {
  o << "bN: " << obj.baseNode << ", cN: " << obj.constraintNode << ", mN: " << obj.mappedNodes
    << ", arrR: " << (const void *)obj.arrR << ", arrV: " << (const void *)obj.arrV << ", io: " << obj.io
    << ", newNode: " << (obj.newNode ? "true" : "false");
  return o;
}

}  // namespace corelib

namespace corelib::data {

static std::atomic<uint64_t> g_graph_uid{1};

static int hw_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  return (int)std::min(n, 16u);
}

// ------------------------------------------------------------- memory

MappedFileSourceMemory::MappedFileSourceMemory(std::filesystem::path pathToFile) {
  int fd = ::open(pathToFile.c_str(), O_RDONLY);
  if (fd < 0) throw std::runtime_error("cannot open " + pathToFile.string());
  struct stat st;
  if (fstat(fd, &st) != 0) {
    ::close(fd);
    throw std::runtime_error("cannot stat " + pathToFile.string());
  }
  size_ = (size_t)st.st_size;
  if (size_) {
    ptr_ = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd, 0);
    if (ptr_ == MAP_FAILED) {
      ptr_ = nullptr;
      ::close(fd);
      throw std::runtime_error("cannot mmap " + pathToFile.string());
    }
  }
  ::close(fd);
}

MappedFileSourceMemory::~MappedFileSourceMemory() {
  if (ptr_) munmap(ptr_, size_);
}

MappedFileSinkMemory::MappedFileSinkMemory(std::filesystem::path pathToFile, size_t size) : size_(size) {
  int fd = ::open(pathToFile.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) throw std::runtime_error("cannot create " + pathToFile.string());
  if (size_ && ftruncate(fd, (off_t)size_) != 0) {
    ::close(fd);
    throw std::runtime_error("cannot size " + pathToFile.string());
  }
  if (size_) {
    ptr_ = mmap(nullptr, size_, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    if (ptr_ == MAP_FAILED) {
      ptr_ = nullptr;
      ::close(fd);
      throw std::runtime_error("cannot mmap " + pathToFile.string());
    }
  }
  ::close(fd);
}

MappedFileSinkMemory::MappedFileSinkMemory(std::filesystem::path pathToFile) {
  int fd = ::open(pathToFile.c_str(), O_RDWR);
  if (fd < 0) throw std::runtime_error("cannot open " + pathToFile.string());
  struct stat st;
  fstat(fd, &st);
  size_ = (size_t)st.st_size;
  if (size_) {
    ptr_ = mmap(nullptr, size_, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    if (ptr_ == MAP_FAILED) {
      ptr_ = nullptr;
      ::close(fd);
      throw std::runtime_error("cannot mmap " + pathToFile.string());
    }
  }
  ::close(fd);
}

MappedFileSinkMemory::~MappedFileSinkMemory() {
  if (ptr_) munmap(ptr_, size_);
}

// ------------------------------------------------------------- caches

CacheDirectoryAccess::CacheDirectoryAccess(std::filesystem::path pathToGraph, bool createDir)
    : name_(pathToGraph.stem().string()),
      graphPath_(pathToGraph),
      cacheDir_(pathToGraph.parent_path() / pathToGraph.stem()) {
  if (createDir && !std::filesystem::exists(cacheDir_)) std::filesystem::create_directories(cacheDir_);
}

std::filesystem::path CacheDirectoryAccess::cacheGraphDirPath(size_t beg, size_t end) const {
  return cacheDir_ / (name_ + "-" + std::to_string(beg) + "-" + std::to_string(end));
}

// ------------------------------------------------------------- parsing

namespace {

inline const char *skip_blank(const char *p, const char *e) {
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
  return p;
}

inline const char *parse_int(const char *p, const char *e, long long *out, bool *ok) {
  p = skip_blank(p, e);
  bool neg = false;
  if (p < e && (*p == '-' || *p == '+')) {
    neg = (*p == '-');
    p++;
  }
  long long v = 0;
  const char *s = p;
  while (p < e && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
  *ok = (p != s);
  *out = neg ? -v : v;
  return p;
}

// Parse every non-blank line "u v t" of [b, e) (b at a line start).
void parse_range(const char *b, const char *e, std::vector<int> &U, std::vector<int> &V, std::vector<int> &T) {
  const char *p = b;
  while (p < e) {
    const char *nl = (const char *)memchr(p, '\n', (size_t)(e - p));
    const char *le = nl ? nl : e;
    const char *q = skip_blank(p, le);
    if (q < le) {
      long long f[3];
      bool ok = true;
      for (int i = 0; i < 3 && ok; i++) q = parse_int(q, le, &f[i], &ok);
      if (!ok) throw std::runtime_error("malformed edge line: " + std::string(p, le));
      U.push_back((int)f[0]);
      V.push_back((int)f[1]);
      T.push_back((int)f[2]);
    }
    p = nl ? nl + 1 : e;
  }
}

}  // namespace

static std::atomic<int> g_build_device{-2};
void setBuildDevice(int device) { g_build_device = device; }
int buildDevice() {
  const int d = g_build_device;
  if (d != -2) return d;
  return tm_device_count() > 0 ? 0 : -1;
}

namespace {

[[noreturn]] void build_failed(const char *what) {
  throw std::runtime_error(std::string(what) + " on the GPU failed: " + tm_last_error());
}

// Stable permutation sorting by T: on the build device (radix sort of
// (time, position) pairs), else on the host (LSD radix sort, 2 x 16-bit digits).
std::vector<int> stable_time_order(const int *T, size_t n) {
  const int dev = buildDevice();
  if (dev >= 0) {
    std::vector<int> ids(n);
    if (tm_time_order(dev, (int64_t)n, T, ids.data()) != TM_OK) build_failed("stable time sort");
    return ids;
  }
  std::vector<int> a(n), b(n);
  std::iota(a.begin(), a.end(), 0);
  std::vector<size_t> cnt(65537);
  for (int pass = 0; pass < 2; pass++) {
    const int sh = 16 * pass;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; i++) cnt[((((uint32_t)T[i]) ^ 0x80000000u) >> sh & 0xffffu) + 1]++;
    for (size_t k = 1; k < cnt.size(); k++) cnt[k] += cnt[k - 1];
    for (size_t i = 0; i < n; i++) {
      int x = a[i];
      b[cnt[(((uint32_t)T[x]) ^ 0x80000000u) >> sh & 0xffffu]++] = x;
    }
    a.swap(b);
  }
  return a;
}

void write_ints(const std::filesystem::path &p, const int *src, size_t n) {
  MappedFileSinkMemory f(p, n * sizeof(int));
  if (n) memcpy(f.data(), src, n * sizeof(int));
}

}  // namespace

// ------------------------------------------------------------- EdgeListLoader

bool EdgeListLoader::shouldInit() const {
  return !(std::filesystem::exists(cacheUPath()) && std::filesystem::exists(cacheVPath()) &&
           std::filesystem::exists(cacheTPath()) && std::filesystem::exists(cacheIDSPath()));
}

void EdgeListLoader::createEdgeListCache() {
  auto t0 = std::chrono::steady_clock::now();
  MappedFileSourceMemory file(graphPath_);
  const char *f = (const char *)file.data();
  const char *l = f + file.size();
  const int nt = hw_threads();
  // chunk starts at line boundaries
  std::vector<const char *> starts(nt + 1);
  starts[0] = f;
  for (int i = 1; i < nt; i++) {
    const char *s = f + file.size() * i / nt;
    if (s < starts[i - 1]) s = starts[i - 1];
    if (s > f && s < l && s[-1] != '\n') {
      const char *nl = (const char *)memchr(s, '\n', (size_t)(l - s));
      s = nl ? nl + 1 : l;
    }
    starts[i] = s;
  }
  starts[nt] = l;
  std::vector<std::vector<int>> Us(nt), Vs(nt), Ts(nt);
  std::vector<std::string> errs(nt);
  std::vector<std::thread> ts;
  for (int i = 0; i < nt; i++) {
    ts.emplace_back([&, i] {
      try {
        if (starts[i] < starts[i + 1]) parse_range(starts[i], starts[i + 1], Us[i], Vs[i], Ts[i]);
      } catch (std::exception &e) {
        errs[i] = e.what();
      }
    });
  }
  for (auto &t : ts) t.join();
  for (auto &e : errs)
    if (!e.empty()) throw std::runtime_error(e);
  size_t total = 0;
  for (auto &u : Us) total += u.size();
  if (total > (size_t)INT32_MAX) throw std::runtime_error("edge list longer than 2^31-1 edges");
  std::vector<int> U, V, T;
  U.reserve(total);
  V.reserve(total);
  T.reserve(total);
  for (int i = 0; i < nt; i++) {
    U.insert(U.end(), Us[i].begin(), Us[i].end());
    V.insert(V.end(), Vs[i].begin(), Vs[i].end());
    T.insert(T.end(), Ts[i].begin(), Ts[i].end());
  }
  auto t1 = std::chrono::steady_clock::now();
  std::vector<int> ids = stable_time_order(T.data(), total);
  std::vector<int> Us2(total), Vs2(total), Ts2(total);
  for (size_t i = 0; i < total; i++) {
    Us2[i] = U[ids[i]];
    Vs2[i] = V[ids[i]];
    Ts2[i] = T[ids[i]];
  }
  auto t2 = std::chrono::steady_clock::now();
  write_ints(cacheUPath(), Us2.data(), total);
  write_ints(cacheVPath(), Vs2.data(), total);
  write_ints(cacheTPath(), Ts2.data(), total);
  write_ints(cacheIDSPath(), ids.data(), total);
  auto ms = [](auto a, auto b) { return (long long)std::chrono::duration_cast<std::chrono::milliseconds>(b - a).count(); };
  std::cout << "# [EdgeListLoader] parse " << ms(t0, t1) << " ms, stable time sort " << ms(t1, t2)
            << " ms for " << total << " records of " << name_ << std::endl;
}

void EdgeListLoader::loadEdgeListCache() {
  U_.reset(new MappedFileSourceMemory(cacheUPath()));
  V_.reset(new MappedFileSourceMemory(cacheVPath()));
  T_.reset(new MappedFileSourceMemory(cacheTPath()));
  IDS_.reset(new MappedFileSourceMemory(cacheIDSPath()));
}

EdgeListLoader::EdgeListLoader(std::filesystem::path pathToGraph) : CacheDirectoryAccess(pathToGraph) {
  if (!std::filesystem::exists(pathToGraph) && shouldInit())
    throw std::runtime_error("graph file not found: " + pathToGraph.string());
  if (shouldInit()) createEdgeListCache();
  loadEdgeListCache();
}

EdgeListLoader::EdgeListLoader(const std::string &name, const int *U, const int *V, const int *T, size_t n)
    : CacheDirectoryAccess(std::filesystem::path("/nonexistent") / (name + ".txt"), false), inMemory_(true) {
  if (n > (size_t)INT32_MAX) throw std::runtime_error("edge list longer than 2^31-1 edges");
  std::vector<int> ids = stable_time_order(T, n);
  std::vector<int> Us(n), Vs(n), Ts(n);
  for (size_t i = 0; i < n; i++) {
    Us[i] = U[ids[i]];
    Vs[i] = V[ids[i]];
    Ts[i] = T[ids[i]];
  }
  U_.reset(new VectorSourceMemory(std::move(Us)));
  V_.reset(new VectorSourceMemory(std::move(Vs)));
  T_.reset(new VectorSourceMemory(std::move(Ts)));
  IDS_.reset(new VectorSourceMemory(std::move(ids)));
}

size_t EdgeListLoader::edgeListLength() const { return U_->size() / sizeof(int); }

// ------------------------------------------------------------- FeatureLoader

namespace {
// labels: first `n` values of the label file, or seeded values when absent
std::vector<int> label_values(const std::filesystem::path &rands, size_t n) {
  std::vector<int> out(n);
  if (!rands.empty() && std::filesystem::exists(rands)) {
    MappedFileSourceMemory r(rands);
    size_t have = r.size() / sizeof(int);
    if (have < n) throw std::runtime_error("label file holds fewer labels than needed");
    if (n) memcpy(out.data(), r.data(), n * sizeof(int));
  } else {
    for (size_t i = 0; i < n; i++)
      out[i] = (int)(corelib::util::splitmix64(0x5eedull * 0x100000001b3ull + i) % 1000000001ull);
  }
  return out;
}
}  // namespace

bool FeatureLoader::shouldInit() const {
  return !(std::filesystem::exists(cacheFEPath()) && std::filesystem::exists(cacheFVVPath()) &&
           std::filesystem::exists(cacheFVMPath()));
}

void FeatureLoader::createFeatureCache() {}

void FeatureLoader::loadFeatureCache() {
  edgeFeatures_.reset(new MappedFileSourceMemory(cacheFEPath()));
  verticesFeatures_.reset(new MappedFileSourceMemory(cacheFVVPath()));
  verticesMap_.reset(new MappedFileSourceMemory(cacheFVMPath()));
}

FeatureLoader::FeatureLoader(const EdgeListLoader *eloader, std::filesystem::path pathToRands)
    : CacheDirectoryAccess(eloader->graphPath_, !eloader->inMemory()), pathToRands_(pathToRands), eloader_(eloader) {
  size_t E = eloader->edgeListLength();
  const int *U = (const int *)eloader->getUPtr()->data();
  const int *V = (const int *)eloader->getVPtr()->data();
  auto build = [&](std::vector<int> &fe, std::vector<int> &fvm, std::vector<int> &fvv) {
    fe = label_values(pathToRands_, E);
    // sorted unique vertex ids (dense flags when the id range is small)
    fvm.clear();
    if (E) {
      int lo = INT32_MAX, hi = INT32_MIN;
      for (size_t i = 0; i < E; i++) {
        lo = std::min(lo, std::min(U[i], V[i]));
        hi = std::max(hi, std::max(U[i], V[i]));
      }
      long long range = (long long)hi - lo + 1;
      if (range <= (long long)std::max<size_t>(4 * E, (size_t)1 << 22)) {
        std::vector<uint8_t> seen((size_t)range, 0);
        for (size_t i = 0; i < E; i++) {
          seen[(size_t)((long long)U[i] - lo)] = 1;
          seen[(size_t)((long long)V[i] - lo)] = 1;
        }
        for (long long k = 0; k < range; k++)
          if (seen[(size_t)k]) fvm.push_back((int)(lo + k));
      } else {
        fvm.assign(U, U + E);
        fvm.insert(fvm.end(), V, V + E);
        std::sort(fvm.begin(), fvm.end());
        fvm.erase(std::unique(fvm.begin(), fvm.end()), fvm.end());
      }
    }
    fvv = label_values(pathToRands_, fvm.size());
  };
  if (eloader->inMemory()) {
    std::vector<int> fe, fvm, fvv;
    build(fe, fvm, fvv);
    edgeFeatures_.reset(new VectorSourceMemory(std::move(fe)));
    verticesFeatures_.reset(new VectorSourceMemory(std::move(fvv)));
    verticesMap_.reset(new VectorSourceMemory(std::move(fvm)));
    return;
  }
  if (shouldInit()) {
    std::vector<int> fe, fvm, fvv;
    build(fe, fvm, fvv);
    write_ints(cacheFEPath(), fe.data(), fe.size());
    write_ints(cacheFVMPath(), fvm.data(), fvm.size());
    write_ints(cacheFVVPath(), fvv.data(), fvv.size());
  }
  loadFeatureCache();
}

size_t FeatureLoader::edgeFeaturesLength() const { return edgeFeatures_->size() / sizeof(int); }
size_t FeatureLoader::verticesFeaturesLength() const { return verticesFeatures_->size() / sizeof(int); }

// ------------------------------------------------------------- graph data

size_t calcTotalBytes(size_t E, size_t V) {
  return (sizeof(TemporalEdge) + sizeof(int) * 3) * E + (sizeof(int) * 2) * (V + 1) + (sizeof(int) * V);
}

std::ostream &operator<<(std::ostream &o, const HostGraphData &obj) {
  o << "name: " << obj.name << "\n"
    << "beg: " << obj.beg << ", end: " << obj.end << "\n"
    << "numEdges: " << obj.numEdges << ", numVertices: " << obj.numVertices << ", totalBytes: " << obj.totalBytes;
  return o;
}

GraphDataLoader::GraphDataLoader(const EdgeListLoader *eloader, const FeatureLoader *floader)
    : CacheDirectoryAccess(eloader->graphPath_, !eloader->inMemory()),
      stem_(eloader->name_),
      eloader_(eloader),
      UPtr_(eloader->getUPtr()),
      VPtr_(eloader->getVPtr()),
      TPtr_(eloader->getTPtr()),
      FEPtr_(floader->getFEPtr()),
      FVVPtr_(floader->getFVVPtr()),
      FVMPtr_(floader->getFVMPtr()),
      EuidPtr_(eloader->getIDSPtr()),
      edgeListLength_(eloader->edgeListLength()),
      verticesFeatureLength_(floader->verticesFeaturesLength()) {}

bool GraphDataLoader::isCached(size_t beg, size_t end) const {
  return !eloader_->inMemory() && std::filesystem::exists(cacheGraphDirPath(beg, end) / "Euid");
}

HostGraphData GraphDataLoader::loadGraphData(size_t beg, size_t end) const {
  HostGraphData res;
  auto cpath = cacheGraphDirPath(beg, end);
  res.name = cpath.filename().string();
  res.beg = beg;
  res.end = end;
  res.Eg_h.reset(new MappedFileSinkMemory(cpath / "Eg"));
  res.inEdgesV_h.reset(new MappedFileSinkMemory(cpath / "inEdgesV"));
  res.outEdgesV_h.reset(new MappedFileSinkMemory(cpath / "outEdgesV"));
  res.inEdgesR_h.reset(new MappedFileSinkMemory(cpath / "inEdgesR"));
  res.outEdgesR_h.reset(new MappedFileSinkMemory(cpath / "outEdgesR"));
  res.verticesFeatures.reset(new MappedFileSinkMemory(cpath / "verticesFeatures"));
  res.edgeFeatures.reset(new MappedFileSinkMemory(cpath / "edgeFeatures"));
  res.Euid_h.reset(new MappedFileSinkMemory(cpath / "Euid"));
  res.numEdges = res.Eg_h->size() / sizeof(int) / 3;
  res.numVertices = res.inEdgesR_h->size() / sizeof(int) - 1;
  res.totalBytes = calcTotalBytes(res.numEdges, res.numVertices);
  res.uid = g_graph_uid.fetch_add(1);
  return res;
}

HostGraphData GraphDataLoader::createGraphData(size_t beg, size_t end, bool cache) const {
  if (end == 0) end = edgeListLength_;
  if (beg > end || end > edgeListLength_) throw std::runtime_error("edge range out of bounds");
  if (eloader_->inMemory()) cache = false;
  if (cache && isCached(beg, end)) return loadGraphData(beg, end);
  const size_t n = end - beg;
  const int *U = (const int *)UPtr_->data() + beg;
  const int *V = (const int *)VPtr_->data() + beg;
  const int *T = (const int *)TPtr_->data() + beg;
  const int dev = buildDevice();
  if (dev >= 0) return createGraphDataOnDevice(dev, beg, end, cache);

  // ---- rename by first occurrence (u before v of the same edge)
  std::vector<int> Un(n), Vn(n), raw;
  raw.reserve(n / 4 + 16);
  int lo = INT32_MAX, hi = INT32_MIN;
  for (size_t i = 0; i < n; i++) {
    lo = std::min(lo, std::min(U[i], V[i]));
    hi = std::max(hi, std::max(U[i], V[i]));
  }
  const long long range = n ? (long long)hi - lo + 1 : 0;
  if (n && range <= (long long)std::max<size_t>(4 * n, (size_t)1 << 22)) {
    std::vector<int> map((size_t)range, -1);
    for (size_t i = 0; i < n; i++) {
      int &a = map[(size_t)((long long)U[i] - lo)];
      if (a < 0) { a = (int)raw.size(); raw.push_back(U[i]); }
      Un[i] = a;
      int &b = map[(size_t)((long long)V[i] - lo)];
      if (b < 0) { b = (int)raw.size(); raw.push_back(V[i]); }
      Vn[i] = b;
    }
  } else {
    std::unordered_map<int, int> map;
    map.reserve(n / 2 + 16);
    for (size_t i = 0; i < n; i++) {
      auto a = map.emplace(U[i], (int)raw.size());
      if (a.second) raw.push_back(U[i]);
      Un[i] = a.first->second;
      auto b = map.emplace(V[i], (int)raw.size());
      if (b.second) raw.push_back(V[i]);
      Vn[i] = b.first->second;
    }
  }
  const size_t nv = raw.size();

  HostGraphData res;
  res.name = cacheGraphDirPath(beg, end).filename().string();
  res.beg = beg;
  res.end = end;
  res.numEdges = n;
  res.numVertices = nv;
  res.totalBytes = calcTotalBytes(n, nv);
  res.uid = g_graph_uid.fetch_add(1);
  auto cpath = cacheGraphDirPath(beg, end);
  auto sink = [&](const char *file, size_t bytes) -> SinkMemory * {
    if (cache) return new MappedFileSinkMemory(cpath / file, bytes);
    return new VectorSinkMemory(bytes);
  };
  if (cache) std::filesystem::create_directories(cpath);
  res.Eg_h.reset(sink("Eg", n * sizeof(int) * 3));
  res.inEdgesV_h.reset(sink("inEdgesV", n * sizeof(int)));
  res.outEdgesV_h.reset(sink("outEdgesV", n * sizeof(int)));
  res.inEdgesR_h.reset(sink("inEdgesR", (nv + 1) * sizeof(int)));
  res.outEdgesR_h.reset(sink("outEdgesR", (nv + 1) * sizeof(int)));
  res.edgeFeatures.reset(sink("edgeFeatures", n * sizeof(int)));
  res.verticesFeatures.reset(sink("verticesFeatures", nv * sizeof(int)));
  // Euid last: its presence marks a complete cache directory
  std::unique_ptr<SinkMemory> euid(sink("Euid", n * sizeof(int)));

  int *Eg = (int *)res.Eg_h->data();
  for (size_t i = 0; i < n; i++) {
    Eg[3 * i] = Un[i];
    Eg[3 * i + 1] = Vn[i];
    Eg[3 * i + 2] = T[i];
  }
  // ---- CSR by counting sort (positions ascending within each vertex)
  auto csr = [&](const std::vector<int> &key, int *R, int *P) {
    std::fill(R, R + nv + 1, 0);
    for (size_t i = 0; i < n; i++) R[key[i] + 1]++;
    for (size_t k = 0; k < nv; k++) R[k + 1] += R[k];
    std::vector<int> fill(R, R + nv);
    for (size_t i = 0; i < n; i++) P[fill[key[i]]++] = (int)i;
  };
  csr(Un, (int *)res.outEdgesR_h->data(), (int *)res.outEdgesV_h->data());
  csr(Vn, (int *)res.inEdgesR_h->data(), (int *)res.inEdgesV_h->data());
  // ---- features and Euid
  const int *FE = (const int *)FEPtr_->data();
  if (n) memcpy(res.edgeFeatures->data(), FE + beg, n * sizeof(int));
  const int *FVM = (const int *)FVMPtr_->data();
  const int *FVV = (const int *)FVVPtr_->data();
  int *VF = (int *)res.verticesFeatures->data();
  for (size_t k = 0; k < nv; k++) {
    const int *it = std::lower_bound(FVM, FVM + verticesFeatureLength_, raw[k]);
    size_t idx = (size_t)(it - FVM);
    VF[k] = idx < verticesFeatureLength_ ? FVV[idx] : 0;
  }
  const int *IDS = (const int *)EuidPtr_->data();
  if (n) memcpy(euid->data(), IDS + beg, n * sizeof(int));
  res.Euid_h = std::move(euid);
  return res;
}

// The same arrays built on a GPU (tm_csr_build: stable radix sorts for the
// renaming and both CSR indexes), then placed in the graph's host buffers.
HostGraphData GraphDataLoader::createGraphDataOnDevice(int dev, size_t beg, size_t end, bool cache) const {
  const size_t n = end - beg;
  const int *U = (const int *)UPtr_->data() + beg;
  const int *V = (const int *)VPtr_->data() + beg;
  const int *T = (const int *)TPtr_->data() + beg;
  std::vector<int> Eg3(3 * n), outV(n), inV(n), outR(2 * n + 1), inR(2 * n + 1), raw(2 * n + 1);
  int32_t nv32 = 0;
  if (tm_csr_build(dev, (int64_t)n, U, V, T, Eg3.data(), outV.data(), inV.data(), outR.data(), inR.data(), raw.data(),
                   &nv32) != TM_OK)
    build_failed("graph build");
  const size_t nv = (size_t)nv32;

  HostGraphData res;
  res.name = cacheGraphDirPath(beg, end).filename().string();
  res.beg = beg;
  res.end = end;
  res.numEdges = n;
  res.numVertices = nv;
  res.totalBytes = calcTotalBytes(n, nv);
  res.uid = g_graph_uid.fetch_add(1);
  auto cpath = cacheGraphDirPath(beg, end);
  auto sink = [&](const char *file, size_t bytes) -> SinkMemory * {
    if (cache) return new MappedFileSinkMemory(cpath / file, bytes);
    return new VectorSinkMemory(bytes);
  };
  if (cache) std::filesystem::create_directories(cpath);
  auto put = [&](std::unique_ptr<SinkMemory> &dst, const char *file, const int *src, size_t count) {
    dst.reset(sink(file, count * sizeof(int)));
    if (count) util::memcpyPar(dst->data(), src, count * sizeof(int));
  };
  put(res.Eg_h, "Eg", Eg3.data(), 3 * n);
  put(res.inEdgesV_h, "inEdgesV", inV.data(), n);
  put(res.outEdgesV_h, "outEdgesV", outV.data(), n);
  put(res.inEdgesR_h, "inEdgesR", inR.data(), nv + 1);
  put(res.outEdgesR_h, "outEdgesR", outR.data(), nv + 1);
  res.edgeFeatures.reset(sink("edgeFeatures", n * sizeof(int)));
  res.verticesFeatures.reset(sink("verticesFeatures", nv * sizeof(int)));
  std::unique_ptr<SinkMemory> euid(sink("Euid", n * sizeof(int)));  // last: marks a complete cache directory
  const int *FE = (const int *)FEPtr_->data();
  if (n) memcpy(res.edgeFeatures->data(), FE + beg, n * sizeof(int));
  const int *FVM = (const int *)FVMPtr_->data();
  const int *FVV = (const int *)FVVPtr_->data();
  int *VF = (int *)res.verticesFeatures->data();
  for (size_t k = 0; k < nv; k++) {
    const int *it = std::lower_bound(FVM, FVM + verticesFeatureLength_, raw[k]);
    size_t idx = (size_t)(it - FVM);
    VF[k] = idx < verticesFeatureLength_ ? FVV[idx] : 0;
  }
  const int *IDS = (const int *)EuidPtr_->data();
  if (n) memcpy(euid->data(), IDS + beg, n * sizeof(int));
  res.Euid_h = std::move(euid);
  return res;
}

HostGraphDataList GraphDataLoader::createPartitionsDataMajor(const SubPartitionList &partitions) const {
  HostGraphDataList res;
  size_t major = (partitions.size() + 1) / 2;
  for (size_t i = 0; i < major; i++) res.push_back(createGraphData(partitions[i].beg, partitions[i].end, true));
  return res;
}

HostGraphDataList GraphDataLoader::createPartitionsDataMinor(const SubPartitionList &partitions) const {
  HostGraphDataList res;
  size_t major = (partitions.size() + 1) / 2;
  for (size_t i = major; i < partitions.size(); i++)
    res.push_back(createGraphData(partitions[i].beg, partitions[i].end, false));
  return res;
}

HostGraphDataList GraphDataLoader::createPartitionsData(const SubPartitionList &partitions) const {
  HostGraphDataList major = createPartitionsDataMajor(partitions);
  HostGraphDataList minor = createPartitionsDataMinor(partitions);
  includeMinor(major, minor);
  return major;
}

// Time-range partitions with a delta halo (reference: src/core/cpu/data.cc:345-369).
SubPartitionList divideList(const EdgeListLoader &eloader, int delta, int num) {
  if (num < 1) throw std::runtime_error("divideList needs num >= 1");
  const int *time = (const int *)eloader.getTPtr()->data();
  size_t size = eloader.edgeListLength();
  SubPartitionList res(num * 2 - 1);
  size_t chunk = size / num;
  for (int i = 0; i < num; i++) {
    res[i].beg = (int)(i * chunk);
    res[i].end = (int)((i + 1) * chunk);
  }
  res[num - 1].end = (int)size;
  res[num - 1].n = res[num - 1].end - res[num - 1].beg;
  for (int i = 0; i < num - 1; i++) {
    int p = res[i].end;
    int t = time[p];
    long long lo = (long long)t - delta, hi = (long long)t + delta;
    long long loId = std::lower_bound(time, time + size, lo, [](int a, long long b) { return a < b; }) - time - 1;
    long long hiId = std::lower_bound(time, time + size, hi, [](int a, long long b) { return a < b; }) - time + 1;
    res[i + num].beg = (int)loId;
    res[i + num].end = (int)hiId;
    res[i + num].n = (int)(p - loId);
    res[i].n = (int)(loId - res[i].beg);
  }
  return res;
}

void removeMinor(HostGraphDataList &complete) { complete.resize((complete.size() + 1) / 2); }

void includeMinor(HostGraphDataList &major, HostGraphDataList &minor) {
  std::move(minor.begin(), minor.end(), std::back_inserter(major));
  minor.clear();
}

}  // namespace corelib::data
