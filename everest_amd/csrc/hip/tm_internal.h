// tm_internal.h -- host-side state shared by the library's translation units.
#pragma once
#include <mutex>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "tm_abi.h"
#include "tm_sweep.h"
#include "tm_types.h"

#define TM_LAUNCH_MAXQ 512  // queries per launch (their table and counts live in LDS)
#define TM_SB_FLUSH_DEFAULT 32   // counting: survivors buffered before a full-width expansion
#define TM_SB_INLINE_DEFAULT 32  // counting: survivors a step expands itself when nothing is buffered

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
  std::vector<tm_query_dev> h_queries;
  std::vector<int> h_qtab;
  std::vector<long long> h_prefix;
};

// One launch lane: a stream with its own scheduler state, donation pool,
// staging and results, so launches on different slots can run concurrently.
#define TM_SLOTS 4
struct Slot {
  int index = 0;                            // position in tm_graph::slot
  int gen = 0;                              // launches since the ring was last cleared (device sched->gen)
  bool own_ring = false;                    // a view with its own scheduler state and ring: its launches
                                            // are counted here, not on the graph's slot of this index
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the mining kernel
  Scratch s;
  float last_kernel_ms = 0.f;
  int last_claim = 0, last_grid = 0, last_lds = 0, last_wpe = 0, last_sweep = 0;
  unsigned long long last_stats[TM_NSTATS_HOST] = {};
};

tm_status set_error(tm_status s, const std::string &msg);
tm_status build_motif(int M, const int32_t *src, const int32_t *dst, tm_motif_dev *md);

}  // namespace tmh

// A graph resident on one device, with its launch slots.
struct tm_graph {
  int device = 0;
  int E = 0, V = 0;
  double deg2[3] = {};  // sums over nodes of out*out, in*in and out*in degrees
  double tspan = 0;     // last edge time minus first
  int num_cus = 0;
  long long wall_rate_khz = 100000;  // wall_clock64 ticks per millisecond
  std::vector<void *> allocs;
  int64_t bytes = 0;
  tm_graph_view view{};
  tm_graph_view *d_view = nullptr;  // a device copy of view (the kernels read the graph through it)
  tmh::Slot slot[TM_SLOTS];  // slot 0 also carries the upload
  int last_slot = 0;         // slot of the last finished launch (stats, kernel time)
  tm_tuning class_tune[TM_NCLASSES] = {};  // per shape class (autotuner), used when a call brings none
  bool class_set[TM_NCLASSES] = {};
  std::mutex mtx;            // one API call at a time per graph handle
  int *edge_features = nullptr;    // labels (tm_graph_set_features), not read by the kernels
  int *vertex_features = nullptr;
  int num_edge_features = 0, num_vertex_features = 0;
};

namespace tmh {

// Create the slot's stream and reserve its staging space for nq queries.
tm_status reserve(tm_graph *g, Slot &sl, int nq);
// Stage queries idx[0..nq) of the caller's arrays at staging positions
// [q0, q0 + nq) of the slot (motif descriptors, delta, root range).
tm_status stage_queries(tm_graph *g, Slot &sl, int q0, int nq, const int *idx, const int32_t *num_edges,
                        const int32_t *msrc, const int32_t *mdst, int stride, const int32_t *deltas,
                        const int32_t *begs, const int32_t *ends);
// A planned launch: kernel arguments and shape.
struct Plan {
  tm_kernel_args a;
  unsigned long long *d_counts = nullptr;  // the launch's per-query counts
  int *d_stops = nullptr;                  // and stop flags
  int grid = 0, lds = 0, nmap = 4, wpe = 1, q0 = 0, nq = 0;
  bool enumerate = false, multi = false;
  int instr = 0;  // 1: reference work, 2: metered bytes
  tm_launch_fn fn = nullptr;
  void *user = nullptr;
  std::vector<SweepItem> sweep;  // queries counted by the window-sweep kernel (empty root range in the mining launch)
  std::vector<PairItem> pair;    // queries counted by the root-pair sweep kernel (likewise)
};
// Plan a launch of staged queries [q0, q0 + nq) of the slot's staging
// (occupancy, variant, claim size, query table) and upload its tables.
tm_status plan(tm_graph *g, Slot &sl, int q0, int nq, bool enumerate, const tm_tuning *tune, int instr, bool multi,
               tm_launch_fn fn, void *user, Plan *pl);
// Enqueue a planned launch on the slot's stream: ring clear when the
// generation wraps (host side), reset kernel, mining kernel (between the
// slot's timing events when timing), statistics kernel.
tm_status enqueue(tm_graph *g, Slot &sl, const Plan &pl, bool timing, bool count);
// plan + enqueue (multi: the batch kernel; fn: a caller's kernel).
tm_status launch(tm_graph *g, Slot &sl, int q0, int nq, bool enumerate, const tm_tuning *tune, int instr,
                 bool multi, tm_launch_fn fn = nullptr, void *user = nullptr);
// Count n launches on the slot (each advances the device generation by one),
// clearing the ring first on `stream` if the generation would wrap.
tm_status count_launch(tm_graph *g, Slot &sl, hipStream_t stream, int n);
// After the slot's stream has drained: statistics and error bits of its launch.
tm_status collect(tm_graph *g, Slot &sl, const tm_sched &hs);
// Synchronous: wait for the slot's launch and copy out its counts.
tm_status finish(tm_graph *g, Slot &sl, int q0, int nq, uint64_t *counts);
// The tuning table entry of a motif's shape class (nullptr when none is set).
const tm_tuning *class_tuning_of(tm_graph *g, int M, const int32_t *src, const int32_t *dst);
// Shape class of a motif (stack depth x node-map width), and the launch
// parameters of a call: its own unless it brings none, else its class's entry.
int shape_class_of(int M, const int32_t *src, const int32_t *dst);
const tm_tuning *tuning_for_class(tm_graph *g, int cls, const tm_tuning *tune);
// Deterministic rows of an enumeration that reaches its cap (cap > 0 rows of
// M ints into host `rows`): every row of the roots before r*, the root at
// which the cumulative count over [beg, r*] first reaches the cap, and the
// remaining rows from root r*'s rows in increasing lexicographic order of
// their Euid tuples.  *count = the matches of the roots [beg, r*].  With a
// caller's kernel (fn, user as tm_enum_custom's) every count and row comes
// from that kernel.
tm_status enum_cut(tm_graph *g, Slot &sl, int M, const int32_t *src, const int32_t *dst, int32_t delta, int32_t beg,
                   int32_t end, int64_t cap, const tm_tuning *tune, int32_t *rows, uint64_t *count,
                   int64_t *rows_written, tm_launch_fn fn = nullptr, void *user = nullptr);

}  // namespace tmh
