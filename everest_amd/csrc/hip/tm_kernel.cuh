// tm_kernel.cuh -- gfx950 temporal-motif mining kernels (device code).
//
// One 64-lane wavefront per workgroup.  Each wavefront keeps a stack of DFS
// frames in LDS; a frame is a partial match (edges 0..L matched) together
// with the still-unexplored slot range of the adjacency list that feeds motif
// edge L+1.  Every step the wavefront deals the next <=64 candidate slots of
// the topmost frames of one level out to its lanes (ballot + prefix scan),
// each lane checks its candidate, and survivors either push a child frame or,
// at the last enumerated edge, finish the match.  In count mode the final
// motif edge is never enumerated: its matches are counted by range
// arithmetic (binary searches over position-sorted lists plus the per-pair
// index for the injectivity correction).
//
// The time window t(p_k) <= t(p_0) + delta is turned into a position bound
// p_k < pend(p_0) (edges are sorted by time), so every candidate range is an
// interval of positions inside a sorted adjacency list.
//
// Semantics follow the reference DFS (plugins/GPUWorkerU1B.cu:10-263,
// system/templates/common/*.snp): matches are strictly increasing sorted
// positions, injective node map, inclusive window, roots with u == v skipped.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tm_types.h"

namespace tmk {

constexpr int WAVE = 64;

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ int lower_bound_i(const int *__restrict__ a, int lo, int hi, int x) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int upper_bound_i(const int *__restrict__ a, int lo, int hi, int x) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

template <int MAXN>
__device__ __forceinline__ int sel(const int (&r)[MAXN], int idx) {
  int v = r[0];
#pragma unroll
  for (int i = 1; i < MAXN; i++) v = (i == idx) ? r[i] : v;
  return v;
}

template <int MAXN>
__device__ __forceinline__ void put(int (&r)[MAXN], int idx, int v) {
#pragma unroll
  for (int i = 0; i < MAXN; i++) r[i] = (i == idx) ? v : r[i];
}

template <int MAXN>
__device__ __forceinline__ bool in_first(const int (&r)[MAXN], int n, int node) {
  bool hit = false;
#pragma unroll
  for (int i = 0; i < MAXN; i++) hit |= (i < n) && (r[i] == node);
  return hit;
}

__device__ __forceinline__ int wave_incl_scan(int x, int lane) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    int y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_or64(unsigned long long x) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    unsigned lo = __shfl_xor((unsigned)x, d);
    unsigned hi = __shfl_xor((unsigned)(x >> 32), d);
    x |= ((unsigned long long)hi << 32) | lo;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    unsigned lo = __shfl_xor((unsigned)x, d);
    unsigned hi = __shfl_xor((unsigned)(x >> 32), d);
    x += ((unsigned long long)hi << 32) | lo;
  }
  return x;
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// first position whose time is > tl (the window end as a position bound)
__device__ __forceinline__ int pend_of(const tm_graph_view &g, long long tl) {
  long long k = (tl - (long long)g.tmin) >> g.bshift;
  if (k < 0) return 0;
  if (k >= g.nbuckets) return g.E;
  int lo = g.bucket[k], hi = g.bucket[k + 1];
  return upper_bound_i(g.T, lo, hi, (int)tl);
}

// number of entries of pair segment (x -> y) inside one node's pair-sorted list
// [lo, hi) whose positions lie in (p, pend)
__device__ __forceinline__ int pair_count(const int *__restrict__ pnbr, const int *__restrict__ ppos,
                                          int lo, int hi, int y, int p, int pend) {
  int s0 = lower_bound_i(pnbr, lo, hi, y);
  if (s0 >= hi || pnbr[s0] != y) return 0;
  int s1 = upper_bound_i(pnbr, s0 + 1, hi, y);
  int c0 = upper_bound_i(ppos, s0, s1, p);
  int c1 = lower_bound_i(ppos, c0, s1, pend);
  return c1 - c0;
}

// Candidate slot range feeding motif edge j, given the mapping after edge j-1.
// dir (CLOSE only): 0 = slots index the out-pair index of m[a], 1 = the in-pair
// index of m[b].
template <int MAXN>
__device__ __forceinline__ void child_range(const tm_graph_view &g, const tm_motif_dev &md, int j,
                                            const int (&m)[MAXN], int p, int pend,
                                            int &cb, int &ce, int &dir) {
  const int kind = md.kind[j];
  dir = 0;
  if (kind == TM_K_NEW_DST) {
    int x = sel(m, md.a[j]);
    int lo = g.outR[x], hi = g.outR[x + 1];
    cb = upper_bound_i(g.outPos, lo, hi, p);
    ce = lower_bound_i(g.outPos, cb, hi, pend);
  } else if (kind == TM_K_NEW_SRC) {
    int y = sel(m, md.b[j]);
    int lo = g.inR[y], hi = g.inR[y + 1];
    cb = upper_bound_i(g.inPos, lo, hi, p);
    ce = lower_bound_i(g.inPos, cb, hi, pend);
  } else {
    int x = sel(m, md.a[j]), y = sel(m, md.b[j]);
    int olo = g.outR[x], ohi = g.outR[x + 1];
    int ilo = g.inR[y], ihi = g.inR[y + 1];
    const int *pp;
    int s0, s1;
    if (ohi - olo <= ihi - ilo) {
      s0 = lower_bound_i(g.outPairNbr, olo, ohi, y);
      s1 = upper_bound_i(g.outPairNbr, s0, ohi, y);
      pp = g.outPairPos;
    } else {
      s0 = lower_bound_i(g.inPairNbr, ilo, ihi, x);
      s1 = upper_bound_i(g.inPairNbr, s0, ihi, x);
      pp = g.inPairPos;
      dir = 1;
    }
    cb = upper_bound_i(pp, s0, s1, p);
    ce = lower_bound_i(pp, cb, s1, pend);
  }
}

// Matches of the final motif edge j given the mapping through edge j-1,
// counted without enumerating them.
template <int MAXN>
__device__ __forceinline__ unsigned long long leaf_count(const tm_graph_view &g, const tm_motif_dev &md,
                                                         int j, const int (&m)[MAXN], int p, int pend,
                                                         int scan_max) {
  const int kind = md.kind[j];
  if (kind == TM_K_CLOSE) {
    int x = sel(m, md.a[j]), y = sel(m, md.b[j]);
    int olo = g.outR[x], ohi = g.outR[x + 1];
    int ilo = g.inR[y], ihi = g.inR[y + 1];
    if (ohi - olo <= ihi - ilo) return pair_count(g.outPairNbr, g.outPairPos, olo, ohi, y, p, pend);
    return pair_count(g.inPairNbr, g.inPairPos, ilo, ihi, x, p, pend);
  }
  const bool out = (kind == TM_K_NEW_DST);
  const int base = out ? sel(m, md.a[j]) : sel(m, md.b[j]);
  const int *R = out ? g.outR : g.inR;
  const int *pos = out ? g.outPos : g.inPos;
  const int *nbr = out ? g.outNbr : g.inNbr;
  int lo = R[base], hi = R[base + 1];
  int b0 = upper_bound_i(pos, lo, hi, p);
  int b1 = lower_bound_i(pos, b0, hi, pend);
  int c = b1 - b0;
  if (c <= 0) return 0;
  const int nm = md.mapped[j];
  if (c <= scan_max) {
    int bad = 0;
    for (int s = b0; s < b1; s++) bad += in_first(m, nm, nbr[s]) ? 1 : 0;
    return (unsigned long long)(c - bad);
  }
  const int *pn = out ? g.outPairNbr : g.inPairNbr;
  const int *pq = out ? g.outPairPos : g.inPairPos;
#pragma unroll
  for (int i = 0; i < MAXN; i++)
    if (i < nm) c -= pair_count(pn, pq, lo, hi, m[i], p, pend);
  return (unsigned long long)c;
}

// ------------------------------------------------------------------ kernel

// LDS frame stack, structure of arrays, capacity FR frames:
//   cur, end, p (last matched position), pend, lvl (level | dir << 8),
//   m[MAXN][FR] mapping, and in enumeration mode e[MAXE][FR] matched positions.
template <int MAXE, int MAXN, bool ENUM>
__global__ __launch_bounds__(64) void mine_kernel(tm_kernel_args args) {
  extern __shared__ int lds[];
  const int lane = threadIdx.x;
  const int FR = args.frames;
  int *f_cur = lds;
  int *f_end = f_cur + FR;
  int *f_p = f_end + FR;
  int *f_pend = f_p + FR;
  int *f_lvl = f_pend + FR;
  int *f_m = f_lvl + FR;
  int *f_e = f_m + MAXN * FR;

  const tm_graph_view g = args.g;
  tm_sched *sched = args.sched;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  unsigned long long cnt = 0;
  int qcur = -1;
  int top = 0;
  int chunk_left = 0;      // 64-root groups left in the current chunk
  int next_root = 0;       // first root of the next 64-root group (uniform)
  int root_end = 0;

  for (;;) {
    if (top == 0) {
      // ---------------------------------------------------- root dispatch
      if (chunk_left == 0) {
        long long c = 0;
        if (lane == 0) c = (long long)atomicAdd(&sched->next_chunk, 1ull);
        c = __shfl(c, 0);
        if (c >= args.total_chunks) break;
        // locate the query owning chunk c (uniform binary search)
        int lo = 0, hi = args.nq;
        while (hi - lo > 1) {
          int mid = (lo + hi) >> 1;
          if (args.chunk_prefix[mid] <= c) lo = mid; else hi = mid;
        }
        const int qi = uni(lo);
        if (qi != qcur) {
          if (!ENUM && qcur >= 0) {
            unsigned long long s = wave_sum64(cnt);
            if (lane == 0 && s) atomicAdd(args.queries[qcur].count, s);
          }
          cnt = 0;
          qcur = qi;
        }
        const tm_query_dev &q = args.queries[qcur];
        if (ENUM && *((volatile int *)q.stop)) continue;
        long long first = (long long)q.root_beg + (c - args.chunk_prefix[qcur]) * (long long)args.chunk_roots;
        next_root = (int)first;
        root_end = q.root_end;
        long long left = ((long long)root_end - first + WAVE - 1) / WAVE;
        int cap = args.chunk_roots / WAVE;
        chunk_left = (int)(left < cap ? left : cap);
        if (chunk_left <= 0) { chunk_left = 0; continue; }
      }
      const tm_query_dev &q = args.queries[qcur];
      const tm_motif_dev &md = q.m;
      const int M = md.M;
      const int r = next_root + lane;
      next_root += WAVE;
      chunk_left--;
      bool ok = r < root_end;
      int u = 0, v = 0, t = 0;
      if (ok) {
        const tm_edge e = g.Eg[r];
        u = e.u; v = e.v; t = e.t;
        ok = (u != v);
      }
      int pend = 0;
      if (ok) {
        pend = pend_of(g, (long long)t + q.delta);
        ok = (pend - r - 1) >= (M - 1);
      }
      int m[MAXN];
#pragma unroll
      for (int i = 0; i < MAXN; i++) m[i] = 0;
      m[0] = u; m[1] = v;
      const int last = ENUM ? M - 1 : M - 2;  // last motif edge handled per candidate
      if (last == 0) {
        // count mode, M == 2: edge 1 is the final edge
        if (ok) cnt += leaf_count(g, md, 1, m, r, pend, args.scan_max);
        continue;
      }
      int cb = 0, ce = 0, dir = 0;
      if (ok) {
        child_range(g, md, 1, m, r, pend, cb, ce, dir);
        ok = ce > cb;
      }
      unsigned long long bm = __ballot(ok);
      int n = __popcll(bm);
      if (top + n > FR) { if (lane == 0) atomicOr(&sched->error, 1); break; }
      if (ok) {
        int w = top + __popcll(bm & lt_mask);
        f_cur[w] = cb; f_end[w] = ce; f_p[w] = r; f_pend[w] = pend; f_lvl[w] = 0 | (dir << 8);
#pragma unroll
        for (int i = 0; i < MAXN; i++) f_m[i * FR + w] = m[i];
        if constexpr (ENUM) f_e[w] = r;
      }
      top += n;
      __syncthreads();
      continue;
    }

    // ------------------------------------------------------ expansion step
    const tm_query_dev &q = args.queries[qcur];
    const tm_motif_dev &md = q.m;
    const int M = md.M;
    const int L = uni(f_lvl[top - 1] & 0xff);
    const int fi = top - 1 - lane;
    int fcnt = 0;
    if (fi >= 0) {
      int lv = f_lvl[fi];
      if ((lv & 0xff) == L) fcnt = f_end[fi] - f_cur[fi];
    }
    int incl = wave_incl_scan(fcnt, lane);
    int excl = incl - fcnt;
    int total = uni(__shfl(incl, WAVE - 1));
    if (total > WAVE) total = WAVE;
    if (total <= 0) {  // a frame on the stack is never empty; stop rather than spin
      if (lane == 0) atomicOr(&sched->error, 2);
      break;
    }
    const bool part = (fcnt > 0) && (excl < WAVE);
    const unsigned long long pm = __ballot(part);
    const int nparts = __popcll(pm);
    int take = part ? min(fcnt, WAVE - excl) : 0;
    const bool last_partial = __ballot(part && take < fcnt) != 0ull;

    int rank, off;
    if (uni(__shfl(fcnt, 0)) >= WAVE) {
      rank = 0; off = lane;
    } else {
      unsigned long long sm = wave_or64(part ? (1ull << excl) : 0ull);
      unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
      rank = __popcll(sm & le) - 1;
      if (rank < 0) rank = 0;
      off = lane - __shfl(excl, rank);
    }
    const bool has = lane < total;
    const int myf = top - 1 - rank;
    int slot = 0, p = 0, pend = 0, dir = 0;
    int m[MAXN];
    int e[ENUM ? MAXE : 1];
#pragma unroll
    for (int i = 0; i < MAXN; i++) m[i] = 0;
    if (has) {
      slot = f_cur[myf] + off;
      p = f_p[myf];
      pend = f_pend[myf];
      dir = f_lvl[myf] >> 8;
#pragma unroll
      for (int i = 0; i < MAXN; i++) m[i] = f_m[i * FR + myf];
      if constexpr (ENUM) {
#pragma unroll
        for (int i = 0; i < MAXE; i++) e[i] = (i <= L) ? f_e[i * FR + myf] : 0;
      }
    }
    __syncthreads();
    // consume: the deepest participating frame may stay with an advanced cursor
    if (part && take < fcnt) f_cur[fi] += take;
    int ntop = top - nparts + (last_partial ? 1 : 0);

    const int j = L + 1;  // motif edge of the candidates
    const int kind = md.kind[j];
    bool valid = false;
    int pos = 0;
    if (has) {
      int node = 0;
      if (kind == TM_K_NEW_DST) {
        pos = g.outPos[slot]; node = g.outNbr[slot];
        valid = !in_first(m, md.mapped[j], node);
        put(m, md.b[j], node);
      } else if (kind == TM_K_NEW_SRC) {
        pos = g.inPos[slot]; node = g.inNbr[slot];
        valid = !in_first(m, md.mapped[j], node);
        put(m, md.a[j], node);
      } else {
        pos = dir ? g.inPairPos[slot] : g.outPairPos[slot];
        valid = true;
      }
    }
    const int last = ENUM ? M - 1 : M - 2;
    if (j == last) {
      if constexpr (ENUM) {
        if (valid) {
          unsigned long long w = atomicAdd(q.count, 1ull);
          if ((long long)w < q.cap) {
            int *row = q.rows + (long long)w * M;
#pragma unroll
            for (int i = 0; i < MAXE; i++)
              if (i < j) row[i] = g.Euid[e[i]];
            row[j] = g.Euid[pos];
          }
          if ((long long)w + 1 >= q.cap) *((volatile int *)q.stop) = 1;
        }
      } else {
        if (valid) cnt += leaf_count(g, md, j + 1, m, pos, pend, args.scan_max);
      }
      top = ntop;
      __syncthreads();
      continue;
    }
    int cb = 0, ce = 0, cdir = 0;
    if (valid) {
      child_range(g, md, j + 1, m, pos, pend, cb, ce, cdir);
      valid = ce > cb;
    }
    const unsigned long long bm = __ballot(valid);
    const int n = __popcll(bm);
    if (ntop + n > FR) { if (lane == 0) atomicOr(&sched->error, 1); break; }
    if (valid) {
      int w = ntop + __popcll(bm & lt_mask);
      f_cur[w] = cb; f_end[w] = ce; f_p[w] = pos; f_pend[w] = pend; f_lvl[w] = j | (cdir << 8);
#pragma unroll
      for (int i = 0; i < MAXN; i++) f_m[i * FR + w] = m[i];
      if constexpr (ENUM) {
#pragma unroll
        for (int i = 0; i < MAXE; i++)
          if (i < j) f_e[i * FR + w] = e[i];
        f_e[j * FR + w] = pos;
      }
    }
    top = ntop + n;
    __syncthreads();
  }

  if (!ENUM && qcur >= 0) {
    unsigned long long s = wave_sum64(cnt);
    if (lane == 0 && s) atomicAdd(args.queries[qcur].count, s);
  }
}

}  // namespace tmk
