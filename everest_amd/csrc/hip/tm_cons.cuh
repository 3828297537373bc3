// tm_cons.cuh -- counting under constraint features: label predicates on
// motif nodes and edges, fine-grained temporal windows between consecutive
// motif edges, and a filter over the matched edges' and nodes' labels
// (reference: system/tmine/Query.py:136-308, system/tmine/KernelGen.py:
// 467-477,530-534, system/templates/constraints/*) and anti-edges
// (system/templates/common/lookAtMatchRes.snp:1-20, findMatch.snp:1-14,
// nextLevel.snp:1-6, system/tmine/Query.py:39-77, KernelGen.py:479-497).
//
// A predicate can reject any candidate, so none of the mining kernel's
// closed forms (range-arithmetic leaves, two-edge tails, window sweep)
// applies: every candidate of every motif edge is visited, in the
// reference's order (system/templates/common/nextMatchInTask.snp:1-21: the
// list of the edge's mapped endpoint after the previous edge's position,
// until the first candidate past the window).  One thread per root, a
// depth-first search over at most 8 levels; the per-query spec CS (generated
// per query, as the reference generates its worker) supplies the motif
// tables and the constraint functions:
//   kM, kN                     motif edges and nodes
//   kind/a/b/mapped(j)         as the mining kernel's spec
//   node_ok(i, VF, node)       label predicate of motif node i
//   edge_ok(j, EF, p)          label predicate of motif edge j (graph position p)
//   window(j, t_j, t0, tmax)   window end for edge j + 1 after edge j at t_j
//   filter(EF, VF, pos, m)     over a complete match
//   anti(j), prev(j)           edge j is an anti edge; the nearest earlier
//                              edge that is not
// The window of edge j + 1 is min(t_p + temporal[j], t0 + delta) for j >= 1
// and t0 + temporal[0] for edge 1 (KernelGen.py:467-477), t0 + delta
// without a temporal list, where p = prev(j + 1) (j itself without anti
// edges).  An anti edge j is a check, not a match position: the partial
// match survives it when no candidate of edge j (both endpoints mapped; the
// edge's label predicate applied) lies after pos[prev(j)] within its window;
// the next edge's list and window start from pos[prev(j)] as well, and the
// search backtracks to the nearest earlier non-anti edge.
//
// Enumeration (the query's rows pointer set): every complete match claims a
// row index with one atomic and writes its edges' original ids (Euid) when
// the index is below the cap, as the reference's enumOut
// (system/tmine/KernelGen.py:509-525); the claim that reaches the cap sets the
// query's stop flag and no root starts after it.  Counting: one atomic per
// wavefront.
#pragma once
#include <hip/hip_runtime.h>

#include "tm_kernel.cuh"
#include "tm_types.h"

namespace tmc {

// A complete match: counted, or its row claimed and written below the cap.
template <int M>
__device__ __forceinline__ void emit(const tm_query_dev &q, const tm_graph_view &g, const int (&pos)[M],
                                     bool enumerate, unsigned long long &cnt) {
  if (!enumerate) {
    cnt++;
    return;
  }
  const unsigned long long w = atomicAdd(q.count, 1ull);
  if ((long long)w < q.cap) {
    int *row = q.rows + (long long)w * M;
#pragma unroll
    for (int i = 0; i < M; i++) row[i] = g.Euid[pos[i]];
  }
  if ((long long)w + 1 >= q.cap) __hip_atomic_store(q.stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class CS>
__global__ __launch_bounds__(64) void cons_kernel(tm_kernel_args args) {
  const tm_graph_view &g = *args.gv;
  const int *EF = args.edgeFeature;
  const int *VF = args.vertexFeature;
  const int4 qt = *(const int4 *)args.qtab;  // delta, root range
  const long long delta = qt.x;
  const int beg = qt.y, end = qt.z;
  constexpr int M = CS::kM, N = CS::kN;
  const tm_query_dev &q = args.queries[0];
  const bool enumerate = q.rows != nullptr;
  unsigned long long cnt = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long rr = beg + (long long)blockIdx.x * blockDim.x + threadIdx.x; rr < end; rr += stride) {
    const int r = (int)rr;
    if (enumerate && __hip_atomic_load(q.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    const tm_edge e0 = g.Eg[r];
    if (e0.u == e0.v) continue;  // roots with u == v are skipped (processFirstEdge.snp:1)
    if (!CS::edge_ok(0, EF, r) || !CS::node_ok(0, VF, e0.u) || !CS::node_ok(1, VF, e0.v)) continue;
    int m[N], pos[M], cur[M], lim[M];
    long long tl[M];
    m[0] = e0.u;
    m[1] = e0.v;
    pos[0] = r;
    const long long t0 = e0.t, tmax = t0 + delta;
    int j = 1;
    bool open = true;  // level j's candidate list is to be set up
    while (j >= 1) {
      if (open) {
        open = false;
        const int pj = CS::prev(j);
        const int p = pos[pj];
        tl[j] = CS::window(j - 1, pj == 0 ? t0 : (long long)g.T[p], t0, tmax);
        const int k = CS::kind(j);
        if (k == TM_K_NEW_DST) {
          const int x = m[CS::a(j)];
          cur[j] = tmk::upper_bound_i(g.outPos, g.outR[x], g.outR[x + 1], p);
          lim[j] = g.outR[x + 1];
        } else if (k == TM_K_NEW_SRC) {
          const int y = m[CS::b(j)];
          cur[j] = tmk::upper_bound_i(g.inPos, g.inR[y], g.inR[y + 1], p);
          lim[j] = g.inR[y + 1];
        } else {  // both endpoints mapped: the pair's edges in x's out-list
          const int x = m[CS::a(j)], y = m[CS::b(j)];
          const int lo = g.outR[x], hi = g.outR[x + 1];
          const int s0 = tmk::lower_bound_i(g.outPairNbr, lo, hi, y);
          const int s1 = tmk::upper_bound_i(g.outPairNbr, s0, hi, y);
          cur[j] = tmk::upper_bound_i(g.outPairPos, s0, s1, p);
          lim[j] = s1;
        }
      }
      if (CS::anti(j)) {  // an anti edge: any candidate within the window rejects the partial match
        bool hit = false;
        while (!hit && cur[j] < lim[j]) {
          const int p = g.outPairPos[cur[j]++];
          if ((long long)g.T[p] > tl[j]) break;
          hit = CS::edge_ok(j, EF, p);
        }
        const int pj = CS::prev(j);
        if (!hit) {
          pos[j] = pos[pj];
          if (j < M - 1) {
            j++;
            open = true;
            continue;
          }
          if (CS::filter(EF, VF, pos, m)) emit(q, g, pos, enumerate, cnt);
        }
        j = pj;
        continue;
      }
      if (cur[j] >= lim[j]) {
        j = CS::prev(j);
        continue;
      }
      const int s = cur[j]++;
      const int k = CS::kind(j);
      const int p = k == TM_K_NEW_DST ? g.outPos[s] : (k == TM_K_NEW_SRC ? g.inPos[s] : g.outPairPos[s]);
      if ((long long)g.T[p] > tl[j]) {  // lists ascend in position, hence in time: the level is done
        cur[j] = lim[j];
        continue;
      }
      if (k != TM_K_CLOSE) {
        const int node = k == TM_K_NEW_DST ? g.outNbr[s] : g.inNbr[s];
        const int nm = CS::mapped(j);
        bool fresh = true;
        for (int i = 0; i < N; i++)
          if (i < nm && m[i] == node) fresh = false;
        if (!fresh || !CS::node_ok(nm, VF, node)) continue;
        m[nm] = node;
      }
      if (!CS::edge_ok(j, EF, p)) continue;
      pos[j] = p;
      if (j == M - 1) {
        if (CS::filter(EF, VF, pos, m)) emit(q, g, pos, enumerate, cnt);
        continue;
      }
      j++;
      open = true;
    }
  }
  // wavefront sum, one atomic per wavefront; this launch's counter rows zeroed
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if (threadIdx.x == 0 && cnt) atomicAdd(q.count, cnt);
  if (threadIdx.x < TM_NSTATS) args.wstats[(size_t)blockIdx.x * TM_NSTATS + threadIdx.x] = 0ull;
}

}  // namespace tmc
