// tm_internal.h -- host-side state shared by the library's translation units.
#pragma once
#include <string>
#include <vector>

#include "tm_abi.h"
#include "tm_types.h"

#define TM_LAUNCH_MAXQ 512  // queries per launch (their table and counts live in LDS)

namespace tmh {

// Per-graph staging area for query descriptors and results (device + host mirrors).
struct Scratch {
  int cap_q = 0;
  tm_query_dev *d_queries = nullptr;
  int *d_qtab = nullptr;                 // per query: first-candidate rule of edge 1, delta, root range
  unsigned long long *d_counts = nullptr;
  int *d_stops = nullptr;
  tm_sched *d_sched = nullptr;
  unsigned long long *d_pool = nullptr;  // donated frames: TM_POOL_FIELDS tagged words each
  int pool_cap = 0;
  int gen = 0;                           // launch generation tag of pool words
  std::vector<tm_query_dev> h_queries;
  std::vector<int> h_qtab;
};

tm_status set_error(tm_status s, const std::string &msg);
tm_status build_motif(int M, const int32_t *src, const int32_t *dst, tm_motif_dev *md);

}  // namespace tmh
