// tm_sweep.hip -- window-sweep counting kernel for 3-edge motifs whose last
// two edges take their candidates from one list of a root node.
//
// Shapes (root 0 -> 1, anchor A in {0, 1}, the anchor's out-list or in-list):
//   star  A->2, A->3   or   2->A, 3->A    (two new nodes)
//   pair  A->2, A->2   or   2->A, 2->A    (a new node's pair twice)
// The mining kernel expands every root's window slot by slot (the reference
// search: plugins/GPUWorkerU1B.cu:10-263, candidates by
// system/templates/common/nextMatchInTask.snp:1-21); its cost grows with the
// window length, so stars at deltas of days cost hundreds of slots per root.
// Here, per root r with window A_r (K slots of the anchor's list with
// position > r and time <= t_r + delta):
//   star:  C(K', 2) + sum_x C(n_x, 2) - P_r
//   pair:  P_r - sum_x C(n_x, 2)
// with x over the root's two nodes, n_x = slots of A_r whose other end is x,
// K' = K - sum_x n_x, and P_r = pairs of slots of A_r with the same other end.
// The roots of one anchor node are its root list (out-list for anchor 0,
// in-list for anchor 1) in position order, and both ends of their windows
// only move forward, so P summed over the valid roots telescopes into one
// term per slot c of the anchor's list:
//   W(i_add) * add(c) - W(i_rem) * rem(c)
// c enters the window at root i_add, the first root whose window end passes
// it, and add(c) counts the same-pair slots already in the window then
// (positions between the previous root and c); it leaves at root i_rem, the
// first root at or after it, and rem(c) counts the same-pair slots after it
// inside that root's window; W(i) = valid roots (not self-loops) of the
// anchor at or after root i.  Every term is a few binary searches, so a
// query costs O(E log E) whatever its delta.  Items are edge positions q:
// q is a root (q < end) and a slot of the anchor list of its source (out)
// or destination (in); each thread takes items q, q + stride, ... and the
// wavefront's sum goes to the query's count with one atomic add (mod 2^64,
// like the reference's uint64 counts).  The kernel runs to completion: a
// query's time limit applies to the mining kernel only.
#include <hip/hip_runtime.h>

#include "tm_internal.h"
#include "tm_kernel.cuh"
#include "tm_sweep.h"

namespace {

using tmk::edge_count;
using tmk::lower_bound_i;
using tmk::pend_of;
using tmk::upper_bound_i;

__device__ __forceinline__ unsigned long long c2(long long n) { return (unsigned long long)(n * (n - 1) / 2); }

// slots of x's anchor list (out: x -> y, in: y -> x) with positions in (p, pend)
__device__ __forceinline__ long long pcnt(const tm_graph_view &g, bool out, int x, int y, int p, int pend) {
  return out ? edge_count(g, x, y, p, pend) : edge_count(g, y, x, p, pend);
}

// valid roots of anchor x at slots [i, r1) of its root list: all but its
// self-loops (positions [RP[i], end))
__device__ __forceinline__ long long valid_roots(const tm_graph_view &g, const int *RP, int x, int i, int r1,
                                                 int end) {
  long long n = r1 - i;
  if (n > 0 && !g.no_self_loops) n -= edge_count(g, x, x, RP[i] - 1, end);
  return n;
}

__global__ __launch_bounds__(256) void sweep_kernel(tm_graph_view g, const tm_query_dev *queries, tmh::SweepBatch b) {
  const tmh::SweepItem it = b.item[blockIdx.y];
  const tm_query_dev &qd = queries[it.q];
  const int beg = qd.root_beg, end = qd.root_end, delta = qd.delta;
  unsigned long long acc = 0;
  if (end > beg) {
    const bool out = it.out != 0, star = it.pat == tmh::SWEEP_STAR;
    const bool same = (it.anchor == 0) == out;  // the anchor's root list is its window list
    const int *LR = out ? g.outR : g.inR;
    const int *LP = out ? g.outPos : g.inPos;
    const int *RR = it.anchor == 0 ? g.outR : g.inR;
    const int *RP = it.anchor == 0 ? g.outPos : g.inPos;
    const int *LS = out ? g.outSelf : g.inSelf;
    const int qmax = pend_of(g, (long long)g.T[end - 1] + delta);
    const int qend = max(end, qmax);
    const int stride = gridDim.x * blockDim.x;
    for (long long qq = beg + (long long)blockIdx.x * blockDim.x + threadIdx.x; qq < qend; qq += stride) {
      const int q = (int)qq;
      const tm_edge e = g.Eg[q];
      const int u = e.u, v = e.v;
      if (q < end && u != v) {  // q as a root
        const int x = it.anchor == 0 ? u : v;
        const int pend = pend_of(g, (long long)e.t + delta);
        const int lo = same ? LS[q] + 1 : upper_bound_i(LP, LR[x], LR[x + 1], q);
        const int hi = lower_bound_i(LP, lo, LR[x + 1], pend);
        const long long nu = pcnt(g, out, x, u, q, pend), nv = pcnt(g, out, x, v, q, pend);
        const unsigned long long cc = c2(nu) + c2(nv);
        acc += star ? c2((long long)(hi - lo) - nu - nv) + cc : 0ull - cc;
      }
      if (q > beg) {  // q as a slot of the anchor list of x
        const int x = out ? u : v, w = out ? v : u;
        const int r0 = lower_bound_i(RP, RR[x], RR[x + 1], beg);
        const int r1 = lower_bound_i(RP, r0, RR[x + 1], end);
        if (r0 == r1) continue;
        // first root whose window end reaches q: time >= t_q - delta, i.e.
        // position >= the first position with that time
        const int pq = pend_of(g, (long long)e.t - delta - 1);
        const int ia = lower_bound_i(RP, r0, r1, pq);
        if (ia == r1) continue;
        const int P = RP[ia > r0 ? ia - 1 : r0];
        if (q <= P) continue;  // before the first root's window: in no window
        const long long add = pcnt(g, out, x, w, P, q);
        const int ir = same ? (q < end ? LS[q] : r1) : lower_bound_i(RP, r0, r1, q);
        long long rem = 0;
        if (ir < r1) rem = pcnt(g, out, x, w, q, pend_of(g, (long long)g.T[RP[ir]] + delta));
        unsigned long long s = (unsigned long long)(valid_roots(g, RP, x, ia, r1, end) * add);
        if (rem) s -= (unsigned long long)(valid_roots(g, RP, x, ir, r1, end) * rem);
        acc += star ? 0ull - s : s;
      }
    }
  }
  // wavefront sum, one atomic per wavefront
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(qd.count, acc);
}

}  // namespace

namespace tmh {

int sweep_shape(const tm_motif_dev &md, SweepItem *it) {
  if (md.M != 3) return 0;
  const int k1 = md.kind[1], k2 = md.kind[2];
  const int a1 = md.a[1], b1 = md.b[1], a2 = md.a[2], b2 = md.b[2];
  int pat = 0, anchor = 0, out = 0;
  if (k1 == TM_K_NEW_DST && k2 == TM_K_NEW_DST && a1 == a2) pat = SWEEP_STAR, anchor = a1, out = 1;
  else if (k1 == TM_K_NEW_SRC && k2 == TM_K_NEW_SRC && b1 == b2) pat = SWEEP_STAR, anchor = b1, out = 0;
  else if (k1 == TM_K_NEW_DST && k2 == TM_K_CLOSE && a2 == a1 && b2 == b1) pat = SWEEP_PAIR, anchor = a1, out = 1;
  else if (k1 == TM_K_NEW_SRC && k2 == TM_K_CLOSE && a2 == a1 && b2 == b1) pat = SWEEP_PAIR, anchor = b1, out = 0;
  if (!pat) return 0;
  if (it) {
    it->pat = pat;
    it->anchor = anchor;
    it->out = out;
  }
  return pat;
}

tm_status sweep_enqueue(const tm_graph_view &g, const tm_query_dev *d_queries, const SweepBatch &b, hipStream_t stream) {
  if (b.n <= 0) return TM_OK;
  // enough 256-thread blocks per query to fill the chip a few times over
  const long long per = ((long long)g.E + 255) / 256;
  const unsigned gx = (unsigned)std::max<long long>(1, std::min<long long>(per, 4096));
  hipLaunchKernelGGL(sweep_kernel, dim3(gx, (unsigned)b.n), dim3(256), 0, stream, g, d_queries, b);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(TM_ERR_HIP, std::string("window-sweep launch: ") + hipGetErrorString(e));
  return TM_OK;
}

}  // namespace tmh
