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
        // the window's first slot: after q's own slot, or (the other list of q's
        // own endpoint) the upload's rank arrays
        const int lo = same ? LS[q] + 1 : (it.anchor == 0 ? g.inNext[q] : g.outNext[q]);
        const int hi = lower_bound_i(LP, lo, LR[x + 1], pend);
        const long long nu = pcnt(g, out, x, u, q, pend), nv = pcnt(g, out, x, v, q, pend);
        const unsigned long long cc = c2(nu) + c2(nv);
        acc += star ? c2((long long)(hi - lo) - nu - nv) + cc : 0ull - cc;
      }
      if (q > beg) {  // q as a slot of the anchor list of x
        const int x = out ? u : v, w = out ? v : u;
        const int r0 = beg == 0 ? RR[x] : lower_bound_i(RP, RR[x], RR[x + 1], beg);
        const int r1 = (RR[x + 1] > r0 && RP[RR[x + 1] - 1] < end) ? RR[x + 1] : lower_bound_i(RP, r0, RR[x + 1], end);
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

// Root-pair sweep: 4-edge motifs root 0 -> 1, edge 1 a new node 2 on a list
// of node s1 in {0, 1} (L1: out-list when d1, else in-list), edge 2 a new
// node 3 on a list of node 2 (L2: out when d2, else in), edge 3 the root's
// pair 0 -> 1 again.  Each root r = (u, v) of the mining kernel expands the
// L1 slots in (r, clip_r), clip_r its pair's last edge in its window, and
// the roots of one pair overlap there.  Here the L1 slots of a pair are
// expanded once: root i takes the piece of (p_i, clip_i) past the previous
// root's clip, and an L1 slot a (new node w) counts, per pair edge p_j > a,
// the roots before a whose window reaches p_j times the edge-2 slots between:
//   sum_j (A - lo_j) * n2(a, p_j)  over the j with lo_j < A,
// A = the pair's roots (within [beg, end)) before a, lo_j = its first root
// with time >= t_j - delta (nondecreasing in j, so the sum stops at the first
// j with lo_j >= A), n2 = L2(w) slots in (a, p_j) whose other end is not u,
// v or w.  A wavefront takes 64 consecutive roots, one per lane, and deals
// the union of their pieces over its lanes (a scan of the piece lengths,
// each slot's root found by a binary search over the lanes), so a hub's
// long piece is shared by all 64 lanes.  Runs to completion, like the
// window sweep.
__global__ __launch_bounds__(256) void pair_sweep_kernel(tm_graph_view g, const tm_query_dev *queries,
                                                         tmh::PairBatch b) {
  const tmh::PairItem it = b.item[blockIdx.y];
  const tm_query_dev &qd = queries[it.q];
  const int beg = qd.root_beg, end = qd.root_end, delta = qd.delta;
  const int lane = threadIdx.x & 63;
  const long long groups = ((long long)end - beg + 63) / 64;
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  const int *L1R = it.d1 ? g.outR : g.inR;
  const int *L1P = it.d1 ? g.outPos : g.inPos;
  const int *L1N = it.d1 ? g.outNbr : g.inNbr;
  const int *L2R = it.d2 ? g.outR : g.inR;
  const int *L2P = it.d2 ? g.outPos : g.inPos;
  const int *L2N = it.d2 ? g.outNbr : g.inNbr;
  const int *pp = g.outPairPos, *pn = g.outPairNbr;
  constexpr int kScan = 32;
  unsigned long long acc = 0;
  for (long long gi = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); gi < groups; gi += wstride) {
    // this lane's root: its pair run [s0, s1e), its roots from slot rb, its L1 piece [i0, i0 + n)
    const long long rr = beg + gi * 64 + lane;
    int n = 0, u = 0, v = 0, s0 = 0, s1e = 0, rb = 0, i0 = 0;
    if (rr < end) {
      const int r = (int)rr;
      const tm_edge e = g.Eg[r];
      u = e.u;
      v = e.v;
      if (u != v) {
        const int k = g.outPairSelf[r];
        s0 = lower_bound_i(pn, g.outR[u], k, v);
        s1e = upper_bound_i(pn, k + 1, g.outR[u + 1], v);
        const int clip = pp[lower_bound_i(pp, k, s1e, pend_of(g, (long long)e.t + delta)) - 1];
        int lo = r;
        if (k > s0 && pp[k - 1] >= beg) {
          const int pr = pp[k - 1];
          lo = max(lo, pp[lower_bound_i(pp, k - 1, s1e, pend_of(g, (long long)g.T[pr] + delta)) - 1]);
        }
        if (clip > lo) {
          const int x = it.s1 == 0 ? u : v;
          i0 = upper_bound_i(L1P, L1R[x], L1R[x + 1], lo);
          n = lower_bound_i(L1P, i0, L1R[x + 1], clip) - i0;
          rb = lower_bound_i(pp, s0, s1e, beg);
        }
      }
    }
    long long incl = n;  // 64 pieces of up to E slots each: 64-bit
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    const long long total = __shfl(incl, 63, 64);
    for (long long base = 0; base < total; base += 64) {
      const long long item = base + lane;
      // owner: the first lane whose inclusive sum exceeds the item
      int lo = 0, hi = 63;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (__shfl(incl, mid, 64) > item) hi = mid;
        else lo = mid + 1;
      }
      const int own = lo;
      const long long oincl = __shfl(incl, own, 64);
      const int on = __shfl(n, own, 64), oi0 = __shfl(i0, own, 64);
      const int ou = __shfl(u, own, 64), ov = __shfl(v, own, 64);
      const int os0 = __shfl(s0, own, 64), os1 = __shfl(s1e, own, 64), orb = __shfl(rb, own, 64);
      if (item >= total) continue;
      const int slot = oi0 + (int)(item - (oincl - on));
      const int a = L1P[slot], w = L1N[slot];
      if (w == ou || w == ov) continue;
      // w's edge-2 list after a, from the upload's rank arrays (a lies in
      // in(w) when edge 1 enters w, in out(w) when it leaves w)
      const int l2lo = it.d1 ? (it.d2 ? g.outNext[a] : g.inSelf[a] + 1) : (it.d2 ? g.outSelf[a] + 1 : g.inNext[a]);
      const int l2hi = L2R[w + 1];
      if (l2lo >= l2hi) continue;
      const int first2 = L2P[l2lo];  // no j with p_j <= first2 adds anything
      const int A = lower_bound_i(pp, orb, os1, min(a, end));
      const long long tlim = (long long)g.T[pp[A - 1]] + delta;  // p_j beyond it: no root before a reaches it
      // edge-2 slots of w after a are scanned once for all j (their other
      // ends checked directly) while the scan stays short; past kScan slots
      // each j's count is a rank minus the pair counts of u, v and w
      int cur = l2lo;
      long long valid = 0;
      bool scan = true;
      int loj = -1;
      // the first pair edge after a: A itself when a < end (a is no pair edge and a > r >= beg)
      for (int j = a < end ? A : upper_bound_i(pp, os0, os1, a); j < os1; j++) {
        const int pj = pp[j];
        const long long tpj = g.T[pj];
        if (tpj > tlim) break;
        if (pj <= first2) continue;
        const long long tj = tpj - delta;  // roots from the first with time >= tj
        if (loj < 0) {
          int lo2 = orb, hi2 = A;
          while (lo2 < hi2) {
            const int mid = (lo2 + hi2) >> 1;
            if ((long long)g.T[pp[mid]] < tj) lo2 = mid + 1;
            else hi2 = mid;
          }
          loj = lo2;
        } else {
          while (loj < A && (long long)g.T[pp[loj]] < tj) loj++;
        }
        if (loj >= A) break;
        long long n2;
        if (scan) {
          while (cur < l2hi && cur - l2lo < kScan && L2P[cur] < pj) {
            const int y = L2N[cur++];
            valid += (y != ou) & (y != ov) & (y != w);
          }
          scan = !(cur < l2hi && L2P[cur] < pj);  // every slot before pj scanned
        }
        if (scan) {
          n2 = valid;
        } else {
          n2 = lower_bound_i(L2P, l2lo, l2hi, pj) - l2lo;
          if (it.d2) n2 -= edge_count(g, w, ou, a, pj) + edge_count(g, w, ov, a, pj);
          else n2 -= edge_count(g, ou, w, a, pj) + edge_count(g, ov, w, a, pj);
          if (!g.no_self_loops) n2 -= edge_count(g, w, w, a, pj);
        }
        acc += (unsigned long long)(A - loj) * (unsigned long long)n2;
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0 && acc) atomicAdd(qd.count, acc);
}

}  // namespace

namespace tmh {

int pair_shape(const tm_motif_dev &md, PairItem *it) {
  if (md.M != 4 || md.kind[3] != TM_K_CLOSE || md.a[3] != 0 || md.b[3] != 1) return 0;
  int s1, d1, d2;
  if (md.kind[1] == TM_K_NEW_DST && md.a[1] <= 1) s1 = md.a[1], d1 = 1;
  else if (md.kind[1] == TM_K_NEW_SRC && md.b[1] <= 1) s1 = md.b[1], d1 = 0;
  else return 0;
  if (md.kind[2] == TM_K_NEW_DST && md.a[2] == 2) d2 = 1;
  else if (md.kind[2] == TM_K_NEW_SRC && md.b[2] == 2) d2 = 0;
  else return 0;
  if (it) {
    it->s1 = s1;
    it->d1 = d1;
    it->d2 = d2;
  }
  return 1;
}

tm_status pair_enqueue(const tm_graph_view &g, const tm_query_dev *d_queries, const PairBatch &b, hipStream_t stream) {
  if (b.n <= 0) return TM_OK;
  // one wavefront per 64 roots, at most 4096 blocks of 4 wavefronts per query
  const long long per = ((long long)g.E + 255) / 256;
  const unsigned gx = (unsigned)std::max<long long>(1, std::min<long long>(per, 4096));
  hipLaunchKernelGGL(pair_sweep_kernel, dim3(gx, (unsigned)b.n), dim3(256), 0, stream, g, d_queries, b);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(TM_ERR_HIP, std::string("pair-sweep launch: ") + hipGetErrorString(e));
  return TM_OK;
}

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
