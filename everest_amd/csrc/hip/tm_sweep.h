// tm_sweep.h -- the window-sweep counting kernel's host interface (3-edge
// stars and a new node's pair twice on one list of a root node).
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

}  // namespace tmh
