// motif.cc -- motif file reader and motif pre-decoding.
//
// A motif file holds M lines "a b k" (reference: inputs/motifs/*.txt).  The
// reference reads it with TemporalGraph::ReadGraph (src/core/cpu/Graph.cc:
// 198-259: stable sort by k, nodes renamed by first appearance, source before
// destination) and pre-decodes it in HostMotifData (src/core/cpu/data.cc:
// 411-449).  Blank lines are skipped here (the reference would read them as
// an edge of indeterminate values); a motif edge with two unmapped endpoints
// is rejected (the reference leaves it undefined).
#include <algorithm>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "everest/data.h"

namespace corelib::data {

HostMotifData HostMotifData::fromEdges(const std::string &name, const std::vector<TemporalEdge> &rawEdges) {
  std::vector<TemporalEdge> raw = rawEdges;
  std::stable_sort(raw.begin(), raw.end(), [](const TemporalEdge &a, const TemporalEdge &b) { return a.t < b.t; });
  std::unordered_map<int, int> ids;
  auto map = [&](int n) {
    auto it = ids.find(n);
    if (it != ids.end()) return it->second;
    int k = (int)ids.size();
    ids.emplace(n, k);
    return k;
  };
  HostMotifData m;
  m.name = name;
  for (auto &e : raw) {
    int u = map(e.u);
    int v = map(e.v);
    m.edges.push_back({u, v, e.t});
  }
  m.numEdges = m.edges.size();
  m.numVertices = ids.size();
  int nodeMax = 1;
  for (size_t i = 1; i < m.numEdges; i++) {
    const TemporalEdge &e = m.edges[i];
    MotifEdgeInfoV1 mi{};
    mi.arrR = nullptr;
    mi.arrV = nullptr;
    if (e.u > nodeMax && e.v > nodeMax)
      throw std::runtime_error("motif edge " + std::to_string(i) + " has two unmapped endpoints");
    if (e.u > nodeMax) {
      mi.baseNode = e.v;
      mi.constraintNode = -e.u;
      mi.io = 0;
      mi.mappedNodes = nodeMax + 1;
      mi.newNode = true;
      nodeMax++;
    } else if (e.v > nodeMax) {
      mi.baseNode = e.u;
      mi.constraintNode = -e.v;
      mi.io = 1;
      mi.mappedNodes = nodeMax + 1;
      mi.newNode = true;
      nodeMax++;
    } else {
      mi.baseNode = e.v;
      mi.constraintNode = e.u;
      mi.io = -1;
      mi.mappedNodes = nodeMax + 1;
      mi.newNode = false;
    }
    m.minfo.push_back(mi);
  }
  return m;
}

HostMotifData::HostMotifData(const std::filesystem::path &pathToMotif) {
  std::ifstream in(pathToMotif);
  if (!in) throw std::runtime_error("cannot open file: " + pathToMotif.string());
  std::vector<TemporalEdge> raw;
  std::string line;
  while (std::getline(in, line)) {
    std::stringstream sl(line);
    TemporalEdge e;
    if (!(sl >> e.u)) continue;
    if (!(sl >> e.v >> e.t)) throw std::runtime_error("malformed motif line: " + line);
    raw.push_back(e);
  }
  *this = fromEdges(pathToMotif.stem().string(), raw);
}

std::ostream &operator<<(std::ostream &o, const HostMotifData &obj) {
  o << "name: " << obj.name << ", numEdges: " << obj.numEdges << ", numVertices: " << obj.numVertices << "\n";
  o << "Edges:\n";
  for (auto &e : obj.edges) o << e.u << " " << e.v << " " << e.t << "\n";
  o << "minfo:";
  for (auto &mi : obj.minfo) o << "\n" << mi;
  return o;
}

}  // namespace corelib::data
