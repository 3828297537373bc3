// tm_sweep.h -- host interface of the sweep counting kernels: the window
// sweep (3-edge stars and a new node's pair twice on one list of a root node)
// and the root-pair sweep (4-edge motifs closing the root's pair).
#pragma once
#include <hip/hip_runtime.h>

#include "tm_abi.h"
#include "tm_types.h"

namespace tmh {

enum { SWEEP_STAR = 1, SWEEP_PAIR = 2 };
#define TM_SWEEP_MAXQ 32  // queries per window-sweep launch (passed by value)

struct SweepItem {
  int q;       // query index in the launch's descriptor array
  int pat;     // SWEEP_STAR or SWEEP_PAIR
  int anchor;  // motif node (0 or 1) whose list holds edges 1 and 2
  int out;     // 1: the anchor's out-list, 0: its in-list
};

struct SweepBatch {
  int n = 0;
  SweepItem item[TM_SWEEP_MAXQ];
};

// The sweep pattern of a motif (0 when it has none); fills *it's shape.
int sweep_shape(const tm_motif_dev &md, SweepItem *it);
// Launch the window sweep for b's queries on `stream`: each adds its count
// to its descriptor's count slot (which must hold 0 or a partial sum).
tm_status sweep_enqueue(const tm_graph_view &g, const tm_query_dev *d_queries, const SweepBatch &b, hipStream_t stream);

// Root-pair sweep (4-edge motifs whose last edge repeats the root's pair,
// edge 1 a new node on a list of a root node, edge 2 a new node on a list of
// edge 1's new node).
struct PairItem {
  int q;   // query index in the launch's descriptor array
  int s1;  // motif node (0 or 1) whose list holds edge 1
  int d1;  // 1: that node's out-list, 0: its in-list
  int d2;  // 1: edge 2 on node 2's out-list, 0: its in-list
};

struct PairBatch {
  int n = 0;
  PairItem item[TM_SWEEP_MAXQ];
};

// 1 when the motif has the root-pair sweep's shape; fills *it's fields.
int pair_shape(const tm_motif_dev &md, PairItem *it);
// Launch the root-pair sweep for b's queries on `stream`: each adds its
// count to its descriptor's count slot.
tm_status pair_enqueue(const tm_graph_view &g, const tm_query_dev *d_queries, const PairBatch &b, hipStream_t stream);

}  // namespace tmh
