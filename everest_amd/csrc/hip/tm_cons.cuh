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
// until the first candidate past the window).  A wavefront takes 64 roots
// and deals their edge-1 candidates over its lanes; below edge 1 each lane
// searches depth-first (at most 8 levels).  The per-query spec CS (generated
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

// Candidate list of motif edge j after position p: slots [*lo, *hi) of its
// list (the pair's edges in m[a]'s out-list when both ends are mapped).
template <class CS, int N>
__device__ __forceinline__ void cand_list(const tm_graph_view &g, int j, const int (&m)[N], int p, int *lo, int *hi) {
  const int k = CS::kind(j);
  if (k == TM_K_NEW_DST) {
    const int x = m[CS::a(j)];
    *lo = tmk::upper_bound_i(g.outPos, g.outR[x], g.outR[x + 1], p);
    *hi = g.outR[x + 1];
  } else if (k == TM_K_NEW_SRC) {
    const int y = m[CS::b(j)];
    *lo = tmk::upper_bound_i(g.inPos, g.inR[y], g.inR[y + 1], p);
    *hi = g.inR[y + 1];
  } else {
    const int x = m[CS::a(j)], y = m[CS::b(j)];
    const int lo0 = g.outR[x], hi0 = g.outR[x + 1];
    const int s0 = tmk::lower_bound_i(g.outPairNbr, lo0, hi0, y);
    const int s1 = tmk::upper_bound_i(g.outPairNbr, s0, hi0, y);
    *lo = tmk::upper_bound_i(g.outPairPos, s0, s1, p);
    *hi = s1;
  }
}

// Position and new node of slot s of edge j's list (node -1 when both ends are mapped).
template <class CS>
__device__ __forceinline__ int cand_pos(const tm_graph_view &g, int j, int s, int *node) {
  const int k = CS::kind(j);
  if (k == TM_K_NEW_DST) { *node = g.outNbr[s]; return g.outPos[s]; }
  if (k == TM_K_NEW_SRC) { *node = g.inNbr[s]; return g.inPos[s]; }
  *node = -1;
  return g.outPairPos[s];
}

// Whether new node `node` of edge j passes: not mapped before, its label predicate.
template <class CS, int N>
__device__ __forceinline__ bool fresh_ok(const int *VF, int j, const int (&m)[N], int node) {
  const int nm = CS::mapped(j);
  bool fresh = true;
  for (int i = 0; i < N; i++)
    if (i < nm && m[i] == node) fresh = false;
  return fresh && CS::node_ok(nm, VF, node);
}

// The search below matched edges 0 .. j0-1 (levels j0 .. M-1; j0 < M), the
// match so far in m / pos, the root's time t0: every complete match emitted.
template <class CS>
__device__ void search_from(int j0, const tm_graph_view &g, const int *EF, const int *VF, const tm_query_dev &q,
                            bool enumerate, int (&m)[CS::kN], int (&pos)[CS::kM], long long t0, long long tmax,
                            unsigned long long &cnt) {
  constexpr int M = CS::kM;
  int cur[M], lim[M];
  long long tl[M];
  int j = j0;
  bool open = true;  // level j's candidate list is to be set up
  while (j >= j0) {
    if (open) {
      open = false;
      const int pj = CS::prev(j);
      tl[j] = CS::window(j - 1, pj == 0 ? t0 : (long long)g.T[pos[pj]], t0, tmax);
      cand_list<CS>(g, j, m, pos[pj], &cur[j], &lim[j]);
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
    int node;
    const int p = cand_pos<CS>(g, j, cur[j]++, &node);
    if ((long long)g.T[p] > tl[j]) {  // lists ascend in position, hence in time: the level is done
      cur[j] = lim[j];
      continue;
    }
    if (node >= 0) {
      if (!fresh_ok<CS>(VF, j, m, node)) continue;
      m[CS::mapped(j)] = node;
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

// One wavefront per 64 consecutive roots.  Each lane sets up its root and
// the slots of edge 1's list inside the root's window; the wavefront deals
// the union of those slots over its 64 lanes (a scan of the counts, each
// slot's root found by a 6-step search over the lanes' sums), so a hub
// root's thousands of edge-1 candidates are searched by 64 lanes; each
// batch of 64 edge-1 items deals its edge-2 slots the same way, and below
// edge 2 each lane searches depth-first alone.  An anti edge 1 (a 2-edge
// motif) is decided per root; an anti edge 2 per edge-1 item, and then edge
// 3 is the dealt one.
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
  const int lane = threadIdx.x;
  unsigned long long cnt = 0;
  for (long long base = beg + (long long)blockIdx.x * 64; base < end; base += (long long)gridDim.x * 64) {
    if (enumerate && __hip_atomic_load(q.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    const long long rr = base + lane;
    int lo1 = 0, n1 = 0, u = 0, v = 0;
    if (rr < end) {
      const int r = (int)rr;
      const tm_edge e0 = g.Eg[r];
      u = e0.u;
      v = e0.v;
      // roots with u == v are skipped (processFirstEdge.snp:1)
      if (u != v && CS::edge_ok(0, EF, r) && CS::node_ok(0, VF, u) && CS::node_ok(1, VF, v)) {
        int m[N];
        m[0] = u;
        m[1] = v;
        const long long t0 = e0.t;
        const long long tl1 = CS::window(0, t0, t0, t0 + delta);
        int hi1;
        cand_list<CS>(g, 1, m, r, &lo1, &hi1);
        if (CS::anti(1)) {  // 2-edge motif: the root alone, when no candidate of edge 1 is in the window
          bool hit = false;
          for (int s = lo1; !hit && s < hi1; s++) {
            const int p = g.outPairPos[s];
            if ((long long)g.T[p] > tl1) break;
            hit = CS::edge_ok(1, EF, p);
          }
          int pos[M];
          pos[0] = r;
          pos[M - 1] = r;
          if (!hit && CS::filter(EF, VF, pos, m)) emit(q, g, pos, enumerate, cnt);
        } else {
          // the slots of edge 1 inside the window: positions before the first with a later time
          const int pend = tmk::pend_of(g, tl1);
          const int *P1 = CS::kind(1) == TM_K_NEW_DST ? g.outPos : (CS::kind(1) == TM_K_NEW_SRC ? g.inPos : g.outPairPos);
          n1 = tmk::lower_bound_i(P1, lo1, hi1, pend) - lo1;
        }
      }
    }
    long long incl = n1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    const long long total = __shfl(incl, 63, 64);
    for (long long b = 0; b < total; b += 64) {
      const long long item = b + lane;
      int lo = 0, hi = 63;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (__shfl(incl, mid, 64) > item) hi = mid;
        else lo = mid + 1;
      }
      const long long oincl = __shfl(incl, lo, 64);
      const int on = __shfl(n1, lo, 64), olo1 = __shfl(lo1, lo, 64);
      const int ou = __shfl(u, lo, 64), ov = __shfl(v, lo, 64);
      // this lane's edge-1 item: the root's two nodes, edge 1's slot
      int m[N], pos[M];
#pragma unroll
      for (int i = 0; i < N; i++) m[i] = 0;
#pragma unroll
      for (int i = 0; i < M; i++) pos[i] = 0;
      bool ok = item < total;
      const int r = (int)(base + lo);
      m[0] = ou;
      m[1] = ov;
      pos[0] = r;
      if (ok) {
        int node;
        const int p = cand_pos<CS>(g, 1, olo1 + (int)(item - (oincl - on)), &node);
        if (node >= 0) {
          ok = fresh_ok<CS>(VF, 1, m, node);
          if (ok) m[CS::mapped(1)] = node;
        }
        ok = ok && CS::edge_ok(1, EF, p);
        pos[1] = p;
      }
      const long long t0 = ok ? (long long)g.T[r] : 0;
      if constexpr (M == 2) {
        if (ok && CS::filter(EF, VF, pos, m)) emit(q, g, pos, enumerate, cnt);
      } else {
        // the dealt level: edge 2, or edge 3 after an anti edge 2 (checked per item here)
        int jd = 2;
        if (CS::anti(2)) {
          if (ok) {
            const long long tla = CS::window(1, g.T[pos[1]], t0, t0 + delta);
            int la, ha;
            cand_list<CS>(g, 2, m, pos[1], &la, &ha);
            bool hit = false;
            for (int s2 = la; !hit && s2 < ha; s2++) {
              const int pa = g.outPairPos[s2];
              if ((long long)g.T[pa] > tla) break;
              hit = CS::edge_ok(2, EF, pa);
            }
            ok = !hit;
            pos[2] = pos[1];
          }
          if (M == 3) {
            if (ok && CS::filter(EF, VF, pos, m)) emit(q, g, pos, enumerate, cnt);
            continue;
          }
          jd = 3;
          if (CS::anti(3)) {  // two anti edges in a row: the rest per lane
            if (ok) search_from<CS>(3, g, EF, VF, q, enumerate, m, pos, t0, t0 + delta, cnt);
            continue;
          }
        }
        // level jd dealt like edge 1: this lane's slots of it inside its window
        const int pd = CS::prev(jd);
        int lo2 = 0, n2 = 0;
        if (ok) {
          const long long tl2 = CS::window(jd - 1, g.T[pos[pd]], t0, t0 + delta);
          int hi2;
          cand_list<CS>(g, jd, m, pos[pd], &lo2, &hi2);
          const int kd = CS::kind(jd);
          const int *P2 = kd == TM_K_NEW_DST ? g.outPos : (kd == TM_K_NEW_SRC ? g.inPos : g.outPairPos);
          n2 = tmk::lower_bound_i(P2, lo2, hi2, tmk::pend_of(g, tl2)) - lo2;
        }
        long long incl2 = n2;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const long long t = __shfl_up(incl2, o, 64);
          if (lane >= o) incl2 += t;
        }
        const long long total2 = __shfl(incl2, 63, 64);
        for (long long b2 = 0; b2 < total2; b2 += 64) {
          const long long item2 = b2 + lane;
          int l2 = 0, h2 = 63;
          while (l2 < h2) {
            const int mid = (l2 + h2) >> 1;
            if (__shfl(incl2, mid, 64) > item2) h2 = mid;
            else l2 = mid + 1;
          }
          const long long oincl2 = __shfl(incl2, l2, 64);
          const int on2 = __shfl(n2, l2, 64), olo2 = __shfl(lo2, l2, 64);
          int mm[N], pp[M];
#pragma unroll
          for (int i = 0; i < N; i++) mm[i] = __shfl(m[i], l2, 64);
#pragma unroll
          for (int i = 0; i < M; i++) pp[i] = __shfl(pos[i], l2, 64);
          const long long ot0 = __shfl(t0, l2, 64);
          if (item2 >= total2) continue;
          int node;
          const int p = cand_pos<CS>(g, jd, olo2 + (int)(item2 - (oincl2 - on2)), &node);
          if (node >= 0) {
            if (!fresh_ok<CS>(VF, jd, mm, node)) continue;
            mm[CS::mapped(jd)] = node;
          }
          if (!CS::edge_ok(jd, EF, p)) continue;
          pp[jd] = p;
          if (jd == M - 1) {
            if (CS::filter(EF, VF, pp, mm)) emit(q, g, pp, enumerate, cnt);
          } else {
            search_from<CS>(jd + 1, g, EF, VF, q, enumerate, mm, pp, ot0, ot0 + delta, cnt);
          }
        }
      }
    }
  }
  // wavefront sum, one atomic per wavefront; this launch's counter rows zeroed
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if (threadIdx.x == 0 && cnt) atomicAdd(q.count, cnt);
  if (threadIdx.x < TM_NSTATS) args.wstats[(size_t)blockIdx.x * TM_NSTATS + threadIdx.x] = 0ull;
}

}  // namespace tmc
