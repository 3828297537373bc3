// Driver for the reference motif reader (test infrastructure only).
// Links against the reference's own src/core/cpu/Graph.cc (compiled where it
// lies by oracle/Makefile) and prints what HostMotifData consumes
// (src/core/cpu/data.cc:411-420): the renamed, time-sorted motif edges.
// Output: "<num_nodes> <num_edges>" then one "u v t" line per edge.
#include <cstdio>
#include "Graph.h"

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: motif_ref <motif-file>\n");
    return 2;
  }
  auto g = TemporalGraph::ReadGraph(argv[1]);
  std::printf("%u %u\n", g->num_nodes(), g->num_edges());
  for (auto &e : g->edges()) std::printf("%d %d %d\n", e.u, e.v, e.t);
  return 0;
}
