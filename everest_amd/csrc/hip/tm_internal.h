// tm_internal.h -- host-side state shared by the library's translation units.
#pragma once
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "tm_abi.h"
#include "tm_types.h"

#define TM_LAUNCH_MAXQ 512  // queries per launch (their table and counts live in LDS)

namespace tmh {

// Per-graph staging area for query descriptors and results (device + host mirrors).
struct Scratch {
  int cap_q = 0;
  tm_query_dev *d_queries = nullptr;
  int *d_qtab = nullptr;                 // per query: delta, root range, start rule of edge 1
  long long *d_prefix = nullptr;         // per query: first chunk of the launch's chunk space
  unsigned long long *d_counts = nullptr;
  int *d_stops = nullptr;
  tm_sched *d_sched = nullptr;
  unsigned long long *d_pool = nullptr;  // donated frames: TM_POOL_FIELDS tagged words each
  unsigned *d_pool_done = nullptr;       // per ring slot: tag of the last record taken
  int pool_cap = 0, pool_shift = 0;
  unsigned long long *d_wstats = nullptr; // per-wavefront counters of the last launch
  int wstats_cap = 0;                    // wavefronts they have room for
  int gen = 0;                           // launch generation tag of pool words
  std::vector<tm_query_dev> h_queries;
  std::vector<int> h_qtab;
  std::vector<long long> h_prefix;
};

// One launch lane: a stream with its own scheduler state, donation pool,
// staging and results, so launches on different slots can run concurrently.
#define TM_SLOTS 4
struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the mining kernel
  Scratch s;
  float last_kernel_ms = 0.f;
  int last_claim = 0, last_grid = 0, last_lds = 0, last_wpe = 0;
  unsigned long long last_stats[16] = {};
};

tm_status set_error(tm_status s, const std::string &msg);
tm_status build_motif(int M, const int32_t *src, const int32_t *dst, tm_motif_dev *md);

}  // namespace tmh
