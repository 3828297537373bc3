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
constexpr int TM_BLOCK = 4 * WAVE;  // roots per dispatch block (4 per lane)
constexpr int TM_ROOTS = WAVE + 1;  // window table entries: one per lane's root, one for a taken frame

// ------------------------------------------------------------------ helpers

// Bytes a lane's search reads from memory, kept only by the metering variant
// of the counting kernel (the product kernel passes no meter, and every
// `if (mt)` folds away): 4 per int32 word, 12 per edge record.
struct Meter {
  unsigned long long b;
};

__device__ __forceinline__ int lower_bound_i(const int *__restrict__ a, int lo, int hi, int x, Meter *mt = nullptr) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (mt) mt->b += 4;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int upper_bound_i(const int *__restrict__ a, int lo, int hi, int x, Meter *mt = nullptr) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (mt) mt->b += 4;
    if (a[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Register-array access by a run-time index.  The optimiser would turn a
// select chain over array elements back into an indexed load from private
// (scratch) memory; passing each element through an empty asm makes it an
// opaque register value, so the array stays in VGPRs.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <int MAXN>
__device__ __forceinline__ int sel(const int (&r)[MAXN], int idx) {
  int v = opaque(r[0]);
#pragma unroll
  for (int i = 1; i < MAXN; i++) v = (i == idx) ? opaque(r[i]) : v;
  return v;
}

template <int MAXN>
__device__ __forceinline__ void put(int (&r)[MAXN], int idx, int v) {
#pragma unroll
  for (int i = 0; i < MAXN; i++) r[i] = (i == idx) ? v : opaque(r[i]);
}

template <int MAXN>
__device__ __forceinline__ bool in_first(const int (&r)[MAXN], int n, int node) {
  bool hit = false;
#pragma unroll
  for (int i = 0; i < MAXN; i++) hit |= (i < n) && (r[i] == node);
  return hit;
}

// component i (uniform) of a register quadruple, kept in registers
__device__ __forceinline__ unsigned pick4(uint4 v, int i) {
  const unsigned a = (unsigned)opaque((int)v.x), b = (unsigned)opaque((int)v.y);
  const unsigned c = (unsigned)opaque((int)v.z), d = (unsigned)opaque((int)v.w);
  return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : d));
}

__device__ __forceinline__ unsigned wave_min_u32(unsigned x) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) x = min(x, (unsigned)__shfl_xor((int)x, o));
  return x;
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

// Device-scope (all XCDs) relaxed atomics for the scheduling and donation
// protocol.  Nothing here uses acquire/release orderings or fences: on a
// multi-XCD part those write back / invalidate the XCD's L2, and thousands of
// wavefronts doing that starve the busy ones.  Donated records carry a launch
// generation tag in every word instead (see TM_POOL_FIELDS).
__device__ __forceinline__ int ld_rlx(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_rlx(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_rlx(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ld_rlx(const long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int add_rlx(int *p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_rlx32(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx32(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tag of the words of a donated record: launch generation (18 bits, never 0)
// and lap of the record's ring slot (14 bits).
__device__ __forceinline__ unsigned pool_tag(int gen, unsigned lap) {
  return ((unsigned)gen << 14) | (lap & 0x3fffu);
}

// Motif-shape policy of the kernel.  DynSpec reads the per-query descriptor
// (one kernel serves any mix of queries: batches, serving); a generated spec
// returns compile-time constants for one motif, so the compiler folds the
// per-edge branches away.
struct DynSpec {
  static __device__ __forceinline__ int M(const tm_motif_dev &md) { return md.M; }
  static __device__ __forceinline__ int kind(const tm_motif_dev &md, int j) { return md.kind[j]; }
  static __device__ __forceinline__ int a(const tm_motif_dev &md, int j) { return md.a[j]; }
  static __device__ __forceinline__ int b(const tm_motif_dev &md, int j) { return md.b[j]; }
  static __device__ __forceinline__ int mapped(const tm_motif_dev &md, int j) { return md.mapped[j]; }
  static __device__ __forceinline__ int smode(const tm_motif_dev &md, int j) { return md.smode[j]; }
  static __device__ __forceinline__ int tmode(const tm_motif_dev &md, int j) { return md.tmode[j]; }
  static __device__ __forceinline__ int xmask(const tm_motif_dev &md, int j) { return md.xmask[j]; }
};

// element j of a compile-time table (j uniform: folds to scalar selects)
template <int N>
__device__ __forceinline__ int pick(const int (&t)[N], int j) {
  int v = t[0];
#pragma unroll
  for (int i = 1; i < N; i++) v = (i == j) ? t[i] : v;
  return v;
}

// first position whose time is > tl (the window end as a position bound)
__device__ __forceinline__ int pend_of(const tm_graph_view &g, long long tl, Meter *mt = nullptr) {
  long long k = (tl - (long long)g.tmin) >> g.bshift;
  if (k < 0) return 0;
  if (k >= g.nbuckets) return g.E;
  int lo = g.bucket[k], hi = g.bucket[k + 1];
  if (mt) mt->b += 8;
  return upper_bound_i(g.T, lo, hi, (int)tl, mt);
}

// first index in [lo, hi) with a[i] >= x, probing outward from lo
// (exponential search: cheap when the answer is near lo, log otherwise)
__device__ __forceinline__ int gallop_lb(const int *__restrict__ a, int lo, int hi, int x, Meter *mt = nullptr) {
  if (mt && lo < hi) mt->b += 4;
  if (lo >= hi || a[lo] >= x) return lo;
  int prev = lo, step = 1;
  for (;;) {
    int nxt = prev + step;
    if (nxt >= hi) return lower_bound_i(a, prev + 1, hi, x, mt);
    if (mt) mt->b += 4;
    if (a[nxt] >= x) return lower_bound_i(a, prev + 1, nxt, x, mt);
    prev = nxt;
    step <<= 1;
  }
}

// first index in [lo, hi) with a[i] > x, probing outward from lo
__device__ __forceinline__ int gallop_ub(const int *__restrict__ a, int lo, int hi, int x, Meter *mt = nullptr) {
  if (mt && lo < hi) mt->b += 4;
  if (lo >= hi || a[lo] > x) return lo;
  int prev = lo, step = 1;
  for (;;) {
    int nxt = prev + step;
    if (nxt >= hi) return upper_bound_i(a, prev + 1, hi, x, mt);
    if (mt) mt->b += 4;
    if (a[nxt] > x) return upper_bound_i(a, prev + 1, nxt, x, mt);
    prev = nxt;
    step <<= 1;
  }
}

// Entries of the pair segment key inside one node's (neighbour, position)-
// sorted list [lo, hi) whose positions lie in (p, pend).
__device__ __forceinline__ int pair_count(const int *__restrict__ pnbr, const int *__restrict__ ppos,
                                          int lo, int hi, int key, int p, int pend, Meter *mt = nullptr) {
  int s0 = lower_bound_i(pnbr, lo, hi, key, mt);
  if (mt && s0 < hi) mt->b += 4;
  if (s0 >= hi || pnbr[s0] != key) return 0;
  int s1 = gallop_ub(pnbr, s0 + 1, hi, key, mt);
  int c0 = upper_bound_i(ppos, s0, s1, p, mt);
  int c1 = gallop_lb(ppos, c0, s1, pend, mt);
  return c1 - c0;
}

// The last position in (p, pend) of an edge x -> y, or -1 (the pair's
// positions through the shorter of out(x) and in(y)).
__device__ __forceinline__ int pair_last(const tm_graph_view &g, int x, int y, int p, int pend,
                                         Meter *mt = nullptr) {
  int lo = g.outR[x], hi = g.outR[x + 1], key = y;
  const int *pn = g.outPairNbr, *pp = g.outPairPos;
  const int ilo = g.inR[y], ihi = g.inR[y + 1];
  if (mt) mt->b += 16;
  if (ihi - ilo < hi - lo) { lo = ilo; hi = ihi; key = x; pn = g.inPairNbr; pp = g.inPairPos; }
  const int s0 = lower_bound_i(pn, lo, hi, key, mt);
  if (mt && s0 < hi) mt->b += 4;
  if (s0 >= hi || pn[s0] != key) return -1;
  const int s1 = gallop_ub(pn, s0 + 1, hi, key, mt);
  const int i = lower_bound_i(pp, s0, s1, pend, mt) - 1;
  if (mt && i >= s0) mt->b += 4;
  return (i >= s0 && pp[i] > p) ? pp[i] : -1;
}

// Edges x -> y with positions in (p, pend), searched in the shorter of out(x)
// and in(y).
__device__ __forceinline__ int edge_count(const tm_graph_view &g, int x, int y, int p, int pend,
                                          Meter *mt = nullptr) {
  int olo = g.outR[x], ohi = g.outR[x + 1];
  int ilo = g.inR[y], ihi = g.inR[y + 1];
  if (mt) mt->b += 16;
  if (ohi - olo <= ihi - ilo) return pair_count(g.outPairNbr, g.outPairPos, olo, ohi, y, p, pend, mt);
  return pair_count(g.inPairNbr, g.inPairPos, ilo, ihi, x, p, pend, mt);
}

// Candidate list of motif edge j: slots [lo, hi) of `pos` (and `nbr` for new
// nodes) whose positions are > p, the position matched for edge j-1.
// dir (CLOSE only): 0 = out-pair index of m[a], 1 = in-pair index of m[b].
struct CandList {
  const int *pos;
  const int *nbr;
  int lo, hi, dir;
};

template <class SP, int MAXN>
__device__ __forceinline__ CandList cand_start(const tm_graph_view &g, const tm_motif_dev &md, int j,
                                               const int (&m)[MAXN], int p, Meter *mt = nullptr) {
  CandList L;
  L.dir = 0;
  const int kind = SP::kind(md, j);
  const int mode = SP::smode(md, j);
  if (kind == TM_K_NEW_DST) {
    const int x = sel(m, SP::a(md, j));
    L.pos = g.outPos;
    L.nbr = g.outNbr;
    L.hi = g.outR[x + 1];
    if (mt) mt->b += 8;
    if (mode == TM_S_OUT_SELF) L.lo = g.outSelf[p] + 1;
    else if (mode == TM_S_OUT_NEXT) L.lo = g.outNext[p];
    else L.lo = upper_bound_i(g.outPos, g.outR[x], L.hi, p, mt);
  } else if (kind == TM_K_NEW_SRC) {
    const int y = sel(m, SP::b(md, j));
    L.pos = g.inPos;
    L.nbr = g.inNbr;
    L.hi = g.inR[y + 1];
    if (mt) mt->b += 8;
    if (mode == TM_S_IN_SELF) L.lo = g.inSelf[p] + 1;
    else if (mode == TM_S_IN_NEXT) L.lo = g.inNext[p];
    else L.lo = upper_bound_i(g.inPos, g.inR[y], L.hi, p, mt);
  } else {
    const int x = sel(m, SP::a(md, j)), y = sel(m, SP::b(md, j));
    L.nbr = nullptr;
    if (mode == TM_S_PAIR_SELF) {
      // the same pair as edge j-1: continue right after p inside its segment
      L.pos = g.outPairPos;
      L.lo = g.outPairSelf[p] + 1;
      if (mt) mt->b += 8;
      L.hi = gallop_ub(g.outPairNbr, L.lo, g.outR[x + 1], y, mt);
    } else {
      int olo = g.outR[x], ohi = g.outR[x + 1];
      int ilo = g.inR[y], ihi = g.inR[y + 1];
      if (mt) mt->b += 16;
      int s0, s1;
      if (ohi - olo <= ihi - ilo) {
        s0 = lower_bound_i(g.outPairNbr, olo, ohi, y, mt);
        s1 = gallop_ub(g.outPairNbr, s0, ohi, y, mt);
        L.pos = g.outPairPos;
      } else {
        s0 = lower_bound_i(g.inPairNbr, ilo, ihi, x, mt);
        s1 = gallop_ub(g.inPairNbr, s0, ihi, x, mt);
        L.pos = g.inPairPos;
        L.dir = 1;
      }
      L.lo = upper_bound_i(L.pos, s0, s1, p, mt);
      L.hi = s1;
    }
  }
  return L;
}

// Matches of the final motif edge j given the mapping through edge j-1 and the
// list from cand_start, counted without enumerating them.
template <class SP, int MAXN>
__device__ __forceinline__ unsigned long long leaf_count(const tm_graph_view &g, const tm_motif_dev &md,
                                                         int j, const int (&m)[MAXN], const CandList &L,
                                                         int p, int pend, int scan_max, int skipx = -1,
                                                         int zmask = 0, Meter *mt = nullptr) {
  const int b1 = gallop_lb(L.pos, L.lo, L.hi, pend, mt);
  int c = b1 - L.lo;
  if (c <= 0 || SP::kind(md, j) == TM_K_CLOSE) return c > 0 ? (unsigned long long)c : 0ull;
  const int nm = SP::mapped(md, j);
  if (c <= scan_max) {
    int bad = 0;
    if (mt) mt->b += 4 * (unsigned long long)c;
    for (int s = L.lo; s < b1; s++) bad += in_first(m, nm, L.nbr[s]) ? 1 : 0;
    return (unsigned long long)(c - bad);
  }
  // injectivity: remove candidates whose new endpoint is already mapped
  // (none to the list's own node in a graph without self-loops, none to
  // skipx: the caller knows that pair has no edge left in the window)
  const bool out = SP::kind(md, j) == TM_K_NEW_DST;
  const int base = out ? sel(m, SP::a(md, j)) : sel(m, SP::b(md, j));
#pragma unroll
  for (int i = 0; i < MAXN; i++) {
    if (i >= nm || ((zmask >> i) & 1) || m[i] == skipx || (g.no_self_loops && m[i] == base)) continue;
    c -= out ? edge_count(g, base, m[i], p, pend, mt) : edge_count(g, m[i], base, p, pend, mt);
  }
  return (unsigned long long)c;
}

// ------------------------------------------------ two-edge tail (count mode)
//
// For a frame of the second-to-last motif edge j = M-2 (its candidate list
// A, bounded by the window), the matches of the last two edges are
//   sum over candidates c of A of (matches of edge k = M-1 after c).
// When edge k does not touch the node edge j introduces, its candidate list
// B is the same for every c, so the sum can be taken the other way round:
//   sum over candidates q of B of (candidates c of A before q that are
//   valid: new node not mapped, and different from q's new node).
// The kernel takes that way when B's window part is the shorter list (hub
// nodes: thousands of candidates for edge j against a handful for edge k),
// and deals B's candidates to the lanes instead of A's.
// A join: when edge j brings node w and edge k closes w with a node y mapped
// before, the matches are the pairs (c, q) with c in A, q an edge between w
// and y after c; summed over the edges q at y in the window (B = y's list),
// each counts A's edges to q's other end before it (one pair search).

// lvl word of a frame: level | A's pair direction << 8 | swapped << 9 |
// B's pair direction << 10 | root index << 16
constexpr int LV_DIR = 1 << 8;
constexpr int LV_SWAP = 1 << 9;
constexpr int LV_BDIR = 1 << 10;

enum { T2_NONE = 0, T2_INDEP = 1, T2_JOIN = 2 };

template <class SP>
__device__ __forceinline__ int tail2_mode(const tm_motif_dev &md) {
  const int M = SP::M(md);
  if (M < 3) return T2_NONE;
  if (SP::kind(md, M - 2) == TM_K_CLOSE) return T2_INDEP;
  const int w = SP::mapped(md, M - 2);  // the node edge M-2 introduces
  const int ak = SP::a(md, M - 1), bk = SP::b(md, M - 1);
  if (ak != w && bk != w) return T2_INDEP;
  return SP::kind(md, M - 1) == TM_K_CLOSE ? T2_JOIN : T2_NONE;
}

template <class SP>
__device__ __forceinline__ bool tail2_ok(const tm_motif_dev &md) {
  return tail2_mode<SP>(md) != T2_NONE;
}

// Edge k's candidates (k = M-1) with positions in (pa0, pend), pa0 the
// position of A's first candidate: [*lo, *hi) in the list of *bdir.
template <class SP, int MAXN>
__device__ __forceinline__ void tail2_range(const tm_graph_view &g, const tm_motif_dev &md, const int (&m)[MAXN],
                                            int pa0, int pend, int *lo, int *hi, int *bdir, Meter *mt = nullptr) {
  const int k = SP::M(md) - 1;
  const int kind = SP::kind(md, k);
  *bdir = 0;
  if (tail2_mode<SP>(md) == T2_JOIN) {
    // edges between w and y at y: y's in-list when edge k is w -> y, else its out-list
    const int w = SP::mapped(md, k - 1);
    const bool into_y = SP::a(md, k) == w;
    const int y = sel(m, into_y ? SP::b(md, k) : SP::a(md, k));
    const int *pos = into_y ? g.inPos : g.outPos;
    const int *R = into_y ? g.inR : g.outR;
    const int e = R[y + 1];
    if (mt) mt->b += 8;
    *lo = upper_bound_i(pos, R[y], e, pa0, mt);
    *hi = gallop_lb(pos, *lo, e, pend, mt);
  } else if (kind == TM_K_NEW_DST) {
    const int x = sel(m, SP::a(md, k));
    const int e = g.outR[x + 1];
    if (mt) mt->b += 8;
    *lo = upper_bound_i(g.outPos, g.outR[x], e, pa0, mt);
    *hi = gallop_lb(g.outPos, *lo, e, pend, mt);
  } else if (kind == TM_K_NEW_SRC) {
    const int y = sel(m, SP::b(md, k));
    const int e = g.inR[y + 1];
    if (mt) mt->b += 8;
    *lo = upper_bound_i(g.inPos, g.inR[y], e, pa0, mt);
    *hi = gallop_lb(g.inPos, *lo, e, pend, mt);
  } else {
    const int x = sel(m, SP::a(md, k)), y = sel(m, SP::b(md, k));
    const int olo = g.outR[x], ohi = g.outR[x + 1];
    const int ilo = g.inR[y], ihi = g.inR[y + 1];
    const int *pos;
    int s0, s1;
    if (mt) mt->b += 16;
    if (ohi - olo <= ihi - ilo) {
      s0 = lower_bound_i(g.outPairNbr, olo, ohi, y, mt);
      s1 = gallop_ub(g.outPairNbr, s0, ohi, y, mt);
      pos = g.outPairPos;
    } else {
      s0 = lower_bound_i(g.inPairNbr, ilo, ihi, x, mt);
      s1 = gallop_ub(g.inPairNbr, s0, ihi, x, mt);
      pos = g.inPairPos;
      *bdir = 1;
    }
    *lo = upper_bound_i(pos, s0, s1, pa0, mt);
    *hi = gallop_lb(pos, *lo, s1, pend, mt);
  }
}

// Matches of the last two edges whose last edge is edge k's candidate at
// `slot` (list direction bdir): edge j's candidates in A [alo, ahi) (pair
// direction adir) before it, less those whose new node is mapped or equal to
// the candidate's new node.
template <class SP, int MAXN>
__device__ __forceinline__ unsigned long long tail2_item(const tm_graph_view &g, const tm_motif_dev &md,
                                                         const int (&m)[MAXN], int alo, int ahi, int adir,
                                                         int slot, int bdir, int zmask = 0, Meter *mt = nullptr) {
  const int M = SP::M(md);
  const int j = M - 2, k = M - 1;
  const int nm = SP::mapped(md, j);
  const int kk = SP::kind(md, k);
  if (tail2_mode<SP>(md) == T2_JOIN) {
    // q = an edge between y and z after A's first candidate: the candidates
    // of A whose new node is z (edges between A's node and z) before it
    const int w = SP::mapped(md, j);
    const bool into_y = SP::a(md, k) == w;
    const int qp = into_y ? g.inPos[slot] : g.outPos[slot];
    const int zz = into_y ? g.inNbr[slot] : g.outNbr[slot];
    if (mt) mt->b += 8;
    if (in_first(m, nm, zz)) return 0ull;  // z must be a new node
    const bool aout = SP::kind(md, j) == TM_K_NEW_DST;
    const int *ap = aout ? g.outPos : g.inPos;
    const int base = aout ? sel(m, SP::a(md, j)) : sel(m, SP::b(md, j));
    if (mt) mt->b += 4;
    const int p0 = ap[alo] - 1;
    return (unsigned long long)(aout ? edge_count(g, base, zz, p0, qp, mt) : edge_count(g, zz, base, p0, qp, mt));
  }
  int qpos, z = -1;
  if (kk == TM_K_NEW_DST) {
    qpos = g.outPos[slot];
    z = g.outNbr[slot];
  } else if (kk == TM_K_NEW_SRC) {
    qpos = g.inPos[slot];
    z = g.inNbr[slot];
  } else {
    qpos = bdir ? g.inPairPos[slot] : g.outPairPos[slot];
  }
  if (mt) mt->b += kk == TM_K_CLOSE ? 4 : 8;
  if (kk != TM_K_CLOSE && in_first(m, nm, z)) return 0ull;
  const int kj = SP::kind(md, j);
  const int *apos = kj == TM_K_NEW_DST ? g.outPos : (kj == TM_K_NEW_SRC ? g.inPos : (adir ? g.inPairPos : g.outPairPos));
  int r = lower_bound_i(apos, alo, ahi, qpos, mt) - alo;
  if (r <= 0 || kj == TM_K_CLOSE) return r > 0 ? (unsigned long long)r : 0ull;
  // edge j's candidates before q whose new node is taken: edges between
  // A's node and that node in [first candidate, q)
  const int p0 = apos[alo] - 1;
  if (mt) mt->b += 4;
  const bool out = kj == TM_K_NEW_DST;
  const int base = out ? sel(m, SP::a(md, j)) : sel(m, SP::b(md, j));
#pragma unroll
  for (int i = 0; i < MAXN; i++) {
    if (i >= nm || ((zmask >> i) & 1) || (g.no_self_loops && m[i] == base)) continue;
    r -= out ? edge_count(g, base, m[i], p0, qpos, mt) : edge_count(g, m[i], base, p0, qpos, mt);
  }
  if (kk != TM_K_CLOSE) r -= out ? edge_count(g, base, z, p0, qpos, mt) : edge_count(g, z, base, p0, qpos, mt);
  return (unsigned long long)r;
}

// ------------------------------------- last two edges on one list (count mode)
//
// When the last two motif edges j = M-2 and k = M-1 take their candidates
// from the same list, their matches per frame of edge j (window list A,
// K slots) have closed forms up to one term per slot:
//   same pair, both closing (x->y twice):   C(K, 2)
//   star (both new, same node and side):    C(K', 2) + sum_x C(n_x, 2) - sum_c sp(c)
//   pair (new node w, then w again):        sum_c sp(c) - sum_x C(n_x, 2)
// where x runs over the nodes mapped before edge j, n_x = slots of A whose
// other end is x, K' = K - sum_x n_x, and sp(c) = the later slots of A with
// the same other end as slot c (the same graph pair): its next same-pair
// time (one coalesced load) shows it is 0 almost always.  The frame's terms
// are added when it is pushed; its slots, dealt as usual, add or subtract
// sp(c).  Every candidate of edge j and every leaf count is skipped.
enum { TP_NONE = 0, TP_STAR = 1, TP_PAIR = 2, TP_SAME = 3 };
constexpr int LV_PAT = 1 << 11;  // frame of edge M-2 summed by its pattern

template <class SP>
__device__ __forceinline__ int tail_pattern(const tm_motif_dev &md) {
  const int M = SP::M(md);
  if (M < 3) return TP_NONE;
  const int j = M - 2, k = M - 1;
  const int kj = SP::kind(md, j), kk = SP::kind(md, k);
  const bool samepair = SP::a(md, k) == SP::a(md, j) && SP::b(md, k) == SP::b(md, j);
  if (kj == TM_K_CLOSE) return (kk == TM_K_CLOSE && samepair) ? TP_SAME : TP_NONE;
  // a new node and then its pair again: the leaf-by-leaf path prunes every
  // candidate whose pair has no later edge with one load already, and the
  // pair pattern's per-frame terms measured slower (kept as a closed form
  // for the model and tests: TP_PAIR is not selected here)
  if (kk == TM_K_CLOSE) return TP_NONE;
  if (kk == kj && (kj == TM_K_NEW_DST ? SP::a(md, k) == SP::a(md, j) : SP::b(md, k) == SP::b(md, j))) return TP_STAR;
  return TP_NONE;
}

// The per-frame terms of a pattern (mod 2^64: the pair pattern's is minus a
// sum); p = position matched for edge j - 1, A = (p, pend) window of K slots.
template <class SP, int MAXN>
__device__ __forceinline__ unsigned long long tail_frame_term(const tm_graph_view &g, const tm_motif_dev &md,
                                                              int pat, const int (&m)[MAXN], int K, int p,
                                                              int pend, Meter *mt = nullptr) {
  if (pat == TP_SAME) return (unsigned long long)K * (unsigned long long)(K - 1) / 2ull;
  const int j = SP::M(md) - 2;
  const int nm = SP::mapped(md, j);
  const bool out = SP::kind(md, j) == TM_K_NEW_DST;
  const int base = out ? sel(m, SP::a(md, j)) : sel(m, SP::b(md, j));
  long long kp = K;
  unsigned long long c2 = 0;
#pragma unroll
  for (int i = 0; i < MAXN; i++) {
    if (i >= nm || (g.no_self_loops && m[i] == base)) continue;
    const long long n = out ? edge_count(g, base, m[i], p, pend, mt) : edge_count(g, m[i], base, p, pend, mt);
    kp -= n;
    c2 += (unsigned long long)(n * (n - 1) / 2);
  }
  if (pat == TP_PAIR) return 0ull - c2;
  return (unsigned long long)(kp * (kp - 1) / 2) + c2;
}

// sp(c) of the pattern sums: later slots of A in the window with the same
// other end as the slot (edges of the same graph pair after it).
template <class SP, int MAXN>
__device__ __forceinline__ unsigned long long tail_same_pair(const tm_graph_view &g, const tm_motif_dev &md,
                                                             const int (&m)[MAXN], int slot, int tl, int pend,
                                                             Meter *mt = nullptr) {
  const int j = SP::M(md) - 2;
  const bool out = SP::kind(md, j) == TM_K_NEW_DST;
  const int lst = out ? TM_L_OUT : TM_L_IN;
  if (mt) mt->b += 4;
  const int tn = g.stf[(size_t)(lst * TM_NTFIRST + TM_T_PAIR_SELF) * (size_t)g.E + slot];
  if (tn > tl) return 0ull;
  const int base = out ? sel(m, SP::a(md, j)) : sel(m, SP::b(md, j));
  if (mt) mt->b += 8;
  const int pos = out ? g.outPos[slot] : g.inPos[slot];
  const int w = out ? g.outNbr[slot] : g.inNbr[slot];
  return (unsigned long long)(out ? edge_count(g, base, w, pos, pend, mt) : edge_count(g, w, base, pos, pend, mt));
}

// The last edge k's injectivity term for one mapped node can be decided by
// the candidate c of edge j itself: c is a graph edge m[a_j] -> m[b_j], and
// when edge k's list holds edges of that same pair (or of the reversed pair)
// towards that node, c's next edge of the pair (TM_T_PAIR_SELF) or of the
// reversed pair (TM_T_RECIP) after it, one load, shows whether any is left
// in the window.  Returns rule | motif node << 8, or -1.
template <class SP>
__device__ __forceinline__ int tail_crule(const tm_motif_dev &md) {
  const int M = SP::M(md);
  if (M < 3) return -1;
  const int j = M - 2, k = M - 1;
  const int kk = SP::kind(md, k);
  const int aj = SP::a(md, j), bj = SP::b(md, j);
  if (kk == TM_K_NEW_DST) {  // corrections: edges m[a_k] -> x
    const int ak = SP::a(md, k);
    if (ak == aj) return TM_T_PAIR_SELF | (bj << 8);
    if (ak == bj) return TM_T_RECIP | (aj << 8);
  } else if (kk == TM_K_NEW_SRC) {  // corrections: edges x -> m[b_k]
    const int bk = SP::b(md, k);
    if (bk == bj) return TM_T_PAIR_SELF | (aj << 8);
    if (bk == aj) return TM_T_RECIP | (bj << 8);
  }
  return -1;
}

// Motif nodes mapped before edge j (bit i: node i) whose injectivity terms
// against `base` (edges base -> x when out, x -> base otherwise) are 0 over
// the whole window (p, pend) of a frame of edge j: every leaf or item of the
// frame, whose range lies inside that window, skips them.
template <class SP, int MAXN>
__device__ __forceinline__ int tail_zero_mask(const tm_graph_view &g, const tm_motif_dev &md, const int (&m)[MAXN],
                                              int base, bool out, int p, int pend, Meter *mt = nullptr) {
  const int nm = SP::mapped(md, SP::M(md) - 2);
  int mask = 0;
#pragma unroll
  for (int i = 0; i < MAXN; i++) {
    if (i >= nm) continue;
    if (g.no_self_loops && m[i] == base) { mask |= 1 << i; continue; }
    const int n = out ? edge_count(g, base, m[i], p, pend, mt) : edge_count(g, m[i], base, p, pend, mt);
    if (n == 0) mask |= 1 << i;
  }
  return mask;
}

// The zero mask of a frame of edge M-2 about to be pushed (0: none worth
// computing): for a frame counted leaf by leaf, against the last edge's
// list node when that node is mapped before edge M-2; for a swapped frame
// (independent last edge), against A's node.  Frames of fewer than four
// candidates skip it.
template <class SP, int MAXN>
__device__ __forceinline__ int tail_frame_mask(const tm_graph_view &g, const tm_motif_dev &md, const int (&m)[MAXN],
                                               bool swapped, int K, int p, int pend, Meter *mt = nullptr) {
  if (K < 4) return 0;
  const int M = SP::M(md);
  const int j = M - 2, k = M - 1;
  const int w = SP::mapped(md, j);
  const int kj = SP::kind(md, j), kk = SP::kind(md, k);
  if (swapped) {
    if (kj == TM_K_CLOSE || tail2_mode<SP>(md) != T2_INDEP) return 0;
    const bool out = kj == TM_K_NEW_DST;
    return tail_zero_mask<SP>(g, md, m, out ? sel(m, SP::a(md, j)) : sel(m, SP::b(md, j)), out, p, pend, mt);
  }
  if (kk == TM_K_CLOSE) return 0;
  const int bn = kk == TM_K_NEW_DST ? SP::a(md, k) : SP::b(md, k);
  if (kj != TM_K_CLOSE && bn == w) return 0;  // the list node is edge j's new node: differs per leaf
  return tail_zero_mask<SP>(g, md, m, sel(m, bn), kk == TM_K_NEW_DST, p, pend, mt);
}

// ----------------------------- repeated edges of one graph pair (count mode)
//
// A motif edge that repeats the pair of the edge before it (smode
// TM_S_PAIR_SELF) makes the search quadratic in the number of parallel
// edges of that pair inside the window: every parallel edge chosen for the
// earlier motif edge starts a subtree, and the later ones repeat the same
// choices.  Two rewrites count each subtree once, with a multiplicity:
//   root pair (edge 1 repeats the root's pair, M >= 4): the candidates of
//     edge 2 after the first parallel edge p1 are expanded once, each
//     weighted by the parallel edges before it (its rank among them);
//   pair tail (edge M-3 brings a node w, edge M-2 repeats that pair): only
//     the first edge of each pair in the window expands, and edge M-2's
//     candidate at slot s of the pair list counts s - (first slot) + 1 times.
// Weights are carried by frames (u32) and multiply the counts below them.
constexpr int LV_RW = 1 << 12;   // level-1 frame of the root-pair rewrite: weight = rank in the root's pair list
constexpr int LV_R1 = 1 << 13;   // frame of the pair-tail rewrite: weight = slot - first slot + 1

template <class SP>
__device__ __forceinline__ bool root_pair_ok(const tm_motif_dev &md) {
  return SP::M(md) >= 4 && SP::kind(md, 1) == TM_K_CLOSE && SP::smode(md, 1) == TM_S_PAIR_SELF;
}

// the edge j (= M-3) whose pair edge M-2 repeats, when j brings a node; else -1
template <class SP>
__device__ __forceinline__ int pair_tail_edge(const tm_motif_dev &md) {
  const int M = SP::M(md);
  if (M < 4) return -1;
  const int j = M - 3;
  if (SP::kind(md, j) == TM_K_CLOSE) return -1;
  // under the root-pair rewrite edge 1 is not one position but all of them:
  // "first after p1" has no single p1 for edge 2
  if (j == 2 && root_pair_ok<SP>(md)) return -1;
  return (SP::kind(md, j + 1) == TM_K_CLOSE && SP::smode(md, j + 1) == TM_S_PAIR_SELF) ? j : -1;
}

// ------------------------------------------------------------------ kernel

// LDS frame stack, structure of arrays, capacity FR frames:
//   cur, end, p (last matched position), pend, lvl (level | dir << 8 | swap flags),
//   m[MAXN][FR] mapping, and in enumeration mode e[MAXE][FR] matched positions.
// Reference work of one resolved candidate list (the instrumented variant):
// the reference scans list entries after p until the first one outside the
// window, and that terminating probe counts too.  For a closing edge it scans
// the shorter of in(m[b]) and out(m[a]) (ties to out).
template <class SP, int MAXN>
__device__ __forceinline__ unsigned long long ref_probes(const tm_graph_view &g, const tm_motif_dev &md, int k,
                                                         const int (&m)[MAXN], int p, int pend) {
  const int kind = SP::kind(md, k);
  const int *pos;
  int lo, hi;
  if (kind == TM_K_NEW_DST || (kind == TM_K_CLOSE && !(g.inR[sel(m, SP::b(md, k)) + 1] - g.inR[sel(m, SP::b(md, k))] <
                                                        g.outR[sel(m, SP::a(md, k)) + 1] - g.outR[sel(m, SP::a(md, k))]))) {
    const int x = sel(m, SP::a(md, k));
    pos = g.outPos; lo = g.outR[x]; hi = g.outR[x + 1];
  } else {
    const int y = sel(m, SP::b(md, k));
    pos = g.inPos; lo = g.inR[y]; hi = g.inR[y + 1];
  }
  lo = upper_bound_i(pos, lo, hi, p);
  const int e = gallop_lb(pos, lo, hi, pend);
  return (unsigned long long)(e - lo) + (e < hi ? 1ull : 0ull);
}

// WPE: wavefronts per SIMD the register allocator must leave room for (1: no
// constraint, the compiler's choice; 6: at most 80 VGPRs, a few spills, 24
// wavefronts per CU when LDS allows).
// MULTI: the launch carries several queries (batches); a single-query launch
// skips the search for the query of a root block.
// INSTR: 0 = the counting kernel; 1 = instrumented with the reference's work
// (walks the reference's search tree: no pruning, no two-edge tail); 2 =
// metered: the counting kernel's own search, adding up the bytes it reads.
template <int MAXE, int MAXN, bool ENUM, int INSTR, class SP = DynSpec, int WPE = 1, bool MULTI = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void mine_kernel(tm_kernel_args args) {
  extern __shared__ int lds[];
  const int lane = threadIdx.x;
  const int FR = args.frames;
  // frame f: candidate slots [f_cur, f_end) feeding the next motif edge,
  // level | direction << 8 | root << 16, node map, edge positions (enumeration).
  // The window of a frame (pend, tl) is its root's, kept once per root in
  // r_pend / r_tl: entries 0..63 for the roots of the quarter being
  // expanded (lane order), entry 64 for a frame taken from another wavefront.
  // counting (not the instrumented kernel): two-edge tails, pair rewrites
  constexpr bool T2 = !ENUM && INSTR != 1;
  int *f_cur = lds;
  int *f_end = f_cur + FR;
  int *f_lvl = f_end + FR;
  int *f_m = f_lvl + FR;
  int *f_e = f_m + MAXN * FR;
  // (counting) the frame's weight and the position its candidate list starts after
  int *f_w = f_e + (ENUM ? MAXE * FR : 0);
  int *f_p0 = f_w + (T2 ? FR : 0);
  int *r_pend = f_p0 + (T2 ? FR : 0);
  int *r_tl = r_pend + TM_ROOTS;
  // (counting) the root's pair list [r_alo, r_ahi) of the root-pair rewrite
  int *r_alo = r_tl + TM_ROOTS;
  int *r_ahi = r_alo + (T2 ? TM_ROOTS : 0);
  // survivor buffer (counting, when args.sb_flush > 0): candidates of one
  // level that passed the injectivity and pruning tests, waiting to be
  // expanded up to 64 at a time, with their weights
  const bool COMPACT = !ENUM && INSTR != 1 && args.sb_flush > 0;
  int *sb_pos = r_ahi + (T2 ? TM_ROOTS : 0);
  int *sb_ridx = sb_pos + (COMPACT ? WAVE : 0);
  int *sb_wl = sb_ridx + (COMPACT ? WAVE : 0);  // weight, low and high words
  int *sb_wh = sb_wl + (COMPACT ? WAVE : 0);
  int *sb_m = sb_wh + (COMPACT ? WAVE : 0);
  // two-edge tail (counting): edge M-2's candidate range [sw_lo, sw_hi) of a
  // swapped frame, at the frame's stack index mod 64 (the frames of the
  // deepest level are pushed in one step, at most 64 of them)
  int *sw_lo = sb_m + (COMPACT ? MAXN * WAVE : 0);
  int *sw_hi = sw_lo + (T2 ? WAVE : 0);
  // per-query counts of this wavefront, flushed once at the end
  unsigned long long *q_cnt = (unsigned long long *)(sw_hi + (T2 ? WAVE : 0));
  for (int q = lane; q < args.nq; q += WAVE) q_cnt[q] = 0;
  __syncthreads();

  const tm_graph_view &g = *args.gv;
  tm_sched *sched = args.sched;
  const int gen = ld_rlx(&sched->gen);  // advanced by the reset kernel before this launch
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  unsigned long long cnt = 0;
  unsigned st_steps = 0, st_roots = 0, st_push = 0;  // uniform, per wavefront
  unsigned long long st_items = 0;
  unsigned st_leaves = 0;  // per lane
  unsigned long long rx1 = 0, rx2 = 0, rp = 0;  // reference work (INSTR == 1 only)
  Meter meter{0};                                // bytes read (INSTR == 2 only)
  Meter *const mt = INSTR == 2 ? &meter : nullptr;
  unsigned st_don = 0, st_take = 0, st_capexit = 0;
  unsigned long long st_work = 0;            // wall-clock ticks this wavefront held work
  long long work_t0 = (long long)wall_clock64();
  unsigned iter = 0;
  int qcur = -1;
  int top = 0;
  long long claim_next = 0, claim_end = 0;  // blocks of the current claim not yet started
  long long qb0 = 0, qb1 = 0;               // blocks of the current query
  int blk = -1;            // first root of the block being dispatched (-1: none)
  int quarter = 0;         // quarter of the block: this lane's root is blk + 4 lane + quarter
  int cq = 0;              // query of the current chunk, and its table entry:
  int cdelta = 0, cbeg = 0, cend = 0, crule = 0;
  int cxmask = 0;          // the root's further gap rules (later edges at the root's nodes), bit per rule
  uint4 bgap = make_uint4(0u, 0u, 0u, 0u);  // gaps of this lane's 4 roots of the block
  bool roots_done = false; // the root chunks of every query are handed out
  const int home_stripe = blockIdx.x & (TM_STRIPES - 1);
  int stripe_off = 0;
  long long task_t0 = 0;   // wall clock when the current task (root group or donated frame) began
  bool is_idle = false;    // gave up its busy credit (sched->busy starts at the grid size)
  bool waiter = false;     // counted in sched->waiters
  int sbn = 0;             // survivors buffered (uniform)
  int sb_level = 0;        // level of the frames they came from
  int steal_cd = 1;        // expansion steps until the next donation check

  // copy one frame between the LDS stack (slot w) and a pool record, one field per lane
  auto frame_field = [&](int f, int w) -> int * {
    if (f == 1) return &f_lvl[w];
    if (f == 2) return &f_cur[w];
    if (f == 3) return &f_end[w];
    if (T2 && f == 4) return &f_p0[w];
    if (T2 && f == 7 + TM_MAXN) return &f_w[w];
    if (f == 5) return &r_pend[(f_lvl[w] >> 16) & 0xff];
    if (f == 6) return &r_tl[(f_lvl[w] >> 16) & 0xff];
    if (f >= 7 && f < 7 + MAXN) return &f_m[(f - 7) * FR + w];
    if (ENUM && f >= 7 + TM_MAXN && f < 7 + TM_MAXN + MAXE) return &f_e[(f - 7 - TM_MAXN) * FR + w];
    return nullptr;
  };

  for (;;) {
    if ((++iter & 63u) == 0u) {  // deadline check (wall clock, uniform)
      if (args.limit_ticks > 0 && lane == 0 && (long long)wall_clock64() - ld_rlx(&sched->t0) > args.limit_ticks) {
        atomicOr(&sched->error, 4);
        atomicOr(&sched->abort, 1);
      }
      // abort: the deadline passed (error bit 2) or a protocol wait gave up (bit 3)
      if (uni(ld_rlx(&sched->abort))) break;
      // enumeration cap reached: the subtrees in flight are dropped (their
      // rows could not be written), so the launch ends as soon as every
      // wavefront sees the flag instead of after finishing them
      if (ENUM && top > 0 && uni(ld_rlx(args.queries[qcur].stop))) top = 0;
    }
    if (top == 0 && sbn == 0 && roots_done) {
      // ------------------------------------------ idle: take donated work
      // Ticket protocol: an idle wavefront draws ticket t (one atomic on
      // pool_head) and waits for record t.  Donors fill records in order
      // (pool_tail) while tickets outnumber records.  sched->busy counts the
      // wavefronts holding work plus the records reserved and not yet
      // taken (a donor adds one before it reserves; the taker inherits it),
      // so busy == 0 means no work is left anywhere and a waiting
      // wavefront may leave.
      if (!is_idle) {
        if (lane == 0) add_rlx(&sched->busy, -1);
        is_idle = true;
        st_work += (unsigned long long)((long long)wall_clock64() - work_t0);
      }
      if (!waiter) {
        // at most idle_cap wavefronts of a launch wait for donations (one
        // atomic, so the bound is strict); the others leave, so the
        // wavefronts of this launch or of a concurrent one that have not
        // started yet always find a CU: waiting wavefronts of one launch can
        // never hold every resource while they wait for its unstarted ones
        int w = 0;
        if (lane == 0) w = add_rlx(&sched->waiters, 1);
        if (uni(__shfl(w, 0)) >= args.idle_cap) {
          if (lane == 0) add_rlx(&sched->waiters, -1);
          st_capexit++;
          break;
        }
        waiter = true;
      }
      unsigned ticket = 0;
      if (lane == 0) ticket = atomicAdd(&sched->pool_head, 1u);
      ticket = (unsigned)uni((int)__shfl((int)ticket, 0));
      // record `ticket` lives in ring slot ticket mod pool_cap, lap ticket / pool_cap
      const unsigned slot = ticket & (unsigned)(args.pool_cap - 1);
      const unsigned tag = pool_tag(gen, ticket >> args.pool_shift);
      const unsigned long long *rec = args.pool + (size_t)slot * TM_POOL_FIELDS;
      int fv = 0;
      bool got = false;
      for (unsigned spin = 0;; spin++) {
        unsigned long long word = lane < TM_POOL_FIELDS ? ld_rlx(rec + lane) : 0ull;
        const bool mine = lane >= TM_POOL_FIELDS || (unsigned)(word >> 32) == tag;
        fv = (int)(unsigned)(word & 0xffffffffull);
        if (__ballot(!mine) == 0ull) { got = true; break; }
        if ((spin & 3u) == 3u) {
          int b = 0, ab = 0;
          if (lane == 0) {
            b = ld_rlx(&sched->busy);
            ab = ld_rlx(&sched->abort);
            if (!ab && args.limit_ticks > 0 && (long long)wall_clock64() - ld_rlx(&sched->t0) > args.limit_ticks) {
              atomicOr(&sched->abort, 1);
              atomicOr(&sched->error, 4);
              ab = 1;
            }
          }
          if (uni(__shfl(ab, 0))) break;
          if (uni(__shfl(b, 0)) == 0) break;  // every record is taken and nobody holds work
        }
        __builtin_amdgcn_s_sleep(32);
      }
      if (!got) break;
      // the frame is copied out below (fv): release the slot for its next lap
      if (lane == 0) {
        st_rlx32(args.pool_done + slot, tag);
        add_rlx(&sched->waiters, -1);
      }
      waiter = false;
      st_take++;
      is_idle = false;  // the record's busy credit is now this wavefront's
      work_t0 = (long long)wall_clock64();
      const int fq = uni(__shfl(fv, 0));
      if (fq != qcur) {
        if (!ENUM && qcur >= 0) {
          unsigned long long sm = wave_sum64(cnt);
          if (lane == 0 && sm) q_cnt[qcur] += sm;
        }
        cnt = 0;
        qcur = fq;
      }
      // the taken frame becomes frame 0, its window root entry TM_ROOTS - 1
      if (lane == 1) f_lvl[0] = (fv & 0xffff) | ((TM_ROOTS - 1) << 16);
      __syncthreads();
      if (lane < TM_POOL_FIELDS && lane != 1) {
        int *dst = frame_field(lane, 0);
        if (dst) *dst = fv;
      }
      top = 1;
      task_t0 = (long long)wall_clock64();
      __syncthreads();
      continue;
    }
    if (top == 0 && sbn == 0) {
      // ---------------------------------------------------- root dispatch
      // The launch's roots form one space of 256-root blocks, query after
      // query, handed out in fixed claims over striped counters.
      // In a block, lane l holds roots blk + 4l .. blk + 4l + 3 and their
      // gaps under the query's start rule (one 16-byte load).  A block none
      // of whose gaps is within delta is skipped whole; otherwise its four
      // quarters (one root per lane) are expanded in turn, each only if some
      // lane's root passes.
      int qsel = -1;
      bool pass = false, inq = false;
      while (qsel < 0) {
        if (blk < 0) {
          if (claim_next >= claim_end) {
            // striped counters: stripe k hands out blocks [k B / S, (k+1) B / S)
            // in claims of claim_blocks; a wavefront starts on its own stripe
            // and moves on when it runs dry
            bool got = false;
            while (stripe_off < TM_STRIPES) {
              const int k = (home_stripe + stripe_off) & (TM_STRIPES - 1);
              const long long sb = args.total_blocks * k / TM_STRIPES;
              const long long se = args.total_blocks * (k + 1) / TM_STRIPES;
              const long long n = args.claim_blocks;
              long long o = 0;
              if (lane == 0) o = (long long)atomicAdd(&sched->next_chunk[k][0], (unsigned long long)n);
              o = sb + __shfl(o, 0);
              if (o < se) {
                claim_next = o;
                claim_end = min(o + n, se);
                got = true;
                break;
              }
              stripe_off++;
            }
            if (!got) { roots_done = true; break; }
          }
          const long long b = claim_next++;
          if (b < qb0 || b >= qb1) {
            // the block's query: the last q with block_prefix[q] <= b (uniform search)
            int lo = 0, hi = MULTI ? args.nq : 1;
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (args.block_prefix[mid] <= b) lo = mid; else hi = mid;
            }
            cq = lo;
            qb0 = args.block_prefix[cq];
            qb1 = args.block_prefix[cq + 1];
            const int4 qt = *(const int4 *)(args.qtab + 4 * cq);
            cdelta = qt.x; cbeg = qt.y; cend = qt.z;
            crule = qt.w & ((1 << TM_QT_XRULE_SHIFT) - 1);
            cxmask = qt.w >> TM_QT_XRULE_SHIFT;
          }
          blk = (cbeg & ~3) + (int)(b - qb0) * TM_BLOCK;
          if (ENUM && ld_rlx(args.queries[cq].stop)) { blk = -1; continue; }  // enumeration cap reached
          const int r0 = blk + 4 * lane;
          uint4 x = *(const uint4 *)(g.rgap + (size_t)crule * (size_t)g.epad + r0);
          for (int rr = 0; rr < TM_NTFIRST; rr++) {  // a root passes only if every rule's next edge is in the window
            if (!((cxmask >> rr) & 1)) continue;
            const uint4 y = *(const uint4 *)(g.rgap + (size_t)rr * (size_t)g.epad + r0);
            x.x = max(x.x, y.x); x.y = max(x.y, y.y); x.z = max(x.z, y.z); x.w = max(x.w, y.w);
          }
          if (r0 < cbeg || r0 >= cend) x.x = ~0u;
          if (r0 + 1 < cbeg || r0 + 1 >= cend) x.y = ~0u;
          if (r0 + 2 < cbeg || r0 + 2 >= cend) x.z = ~0u;
          if (r0 + 3 < cbeg || r0 + 3 >= cend) x.w = ~0u;
          bgap = x;
          st_roots += (unsigned)(min(max(cend - blk, 0), TM_BLOCK) - min(max(cbeg - blk, 0), TM_BLOCK));
          const unsigned mn = min(min(x.x, x.y), min(x.z, x.w));
          const bool near = cdelta >= 0 && mn <= (unsigned)cdelta;
          if (mt) meter.b += 16 * (1 + __popc((unsigned)cxmask));  // the gap words of this lane's 4 roots
          if (INSTR != 1 && __ballot(near) == 0ull) { blk = -1; continue; }
          quarter = -1;
        }
        if (++quarter == 4) { blk = -1; continue; }
        const int r = blk + 4 * lane + quarter;
        const unsigned lg = pick4(bgap, quarter);
        inq = r >= cbeg && r < cend;
        pass = inq && cdelta >= 0 && lg <= (unsigned)cdelta;
        if (__ballot(pass) != 0ull || (INSTR == 1 && __ballot(inq) != 0ull)) qsel = cq;
      }
      if (qsel < 0) continue;  // every root chunk is handed out
      if (qsel != qcur) {
        if (!ENUM && qcur >= 0) {
          unsigned long long sm = wave_sum64(cnt);
          if (lane == 0 && sm) q_cnt[qcur] += sm;
        }
        cnt = 0;
        qcur = qsel;
      }
      const tm_query_dev &q = args.queries[qcur];
      const tm_motif_dev &md = q.m;
      const int M = SP::M(md);
      const int r = blk + 4 * lane + quarter;
      int u = 0, v = 0, t = 0;
      bool gvalid = false;
      if (pass || (INSTR == 1 && inq)) {
        if (mt) meter.b += 12;
        const tm_edge e = g.Eg[r];
        u = e.u; v = e.v; t = e.t;
        gvalid = u != v;
      }
      bool ok = pass;
      int m[MAXN];
#pragma unroll
      for (int i = 0; i < MAXN; i++) m[i] = 0;
      m[0] = u; m[1] = v;
      CandList L{g.outPos, g.outNbr, 0, 0, 0};
      const long long tl = (long long)t + q.delta;
      int tli = tl > 0x7fffffffll ? 0x7fffffff : (int)tl;
      if (ok) {
        L = cand_start<SP>(g, md, 1, m, r, mt);
        if (mt && L.lo < L.hi) meter.b += 8;
        ok = L.lo < L.hi && (long long)g.T[L.pos[L.lo]] <= tl;
      }
      int pend = 0;
      if constexpr (INSTR == 1) {
        if (gvalid) {
          const int pe = pend_of(g, tl);
          (SP::kind(md, 1) == TM_K_CLOSE ? rx2 : rx1) += 1;
          rp += ref_probes<SP>(g, md, 1, m, r, pe);
        }
      }
      if (ok) {
        pend = pend_of(g, tl, mt);
        if (INSTR != 1) ok = (pend - r - 1) >= (M - 1);
      }
      // (counting) the last motif edge closes the root's own two nodes: every
      // edge of a match lies at or before that pair's last edge in the
      // window, so the root's window ends there (positions and pruning time)
      if (T2 && ok && M >= 3 && SP::kind(md, M - 1) == TM_K_CLOSE && SP::a(md, M - 1) < 2 &&
          SP::b(md, M - 1) < 2) {
        const int x = SP::a(md, M - 1) == 0 ? u : v, y = SP::b(md, M - 1) == 0 ? u : v;
        const int pl = pair_last(g, x, y, r, pend, mt);
        if (pl < 0) {
          ok = false;
        } else {
          pend = pl + 1;
          if (mt) meter.b += 4;
          tli = g.T[pl];
          ok = (pend - r - 1) >= (M - 1);
        }
      }
      const int last = ENUM ? M - 1 : M - 2;  // last motif edge handled per candidate
      if (last == 0) {
        // count mode, M == 2: edge 1 is the final edge
        if (ok) { cnt += leaf_count<SP>(g, md, 1, m, L, r, pend, args.scan_max, -1, 0, mt); st_leaves++; }
        continue;
      }
      int cb = 0, ce = 0, dir = L.dir;
      if (ok) {
        cb = L.lo;
        ce = gallop_lb(L.pos, L.lo, L.hi, pend, mt);
        ok = ce > cb;
      }
      // root-pair rewrite: edge 1 repeats the root's pair; expand edge 2 once
      // from the first parallel edge, its candidates weighted by their rank
      // among the parallel edges [a1lo, a1hi) of the window
      int lvl0 = 0, rw = 0, p0f = r, a1lo = 0, a1hi = 0;
      if (T2 && ok && root_pair_ok<SP>(md)) {
        a1lo = cb;
        a1hi = ce;
        if (mt) meter.b += 4;
        const int p1 = L.pos[cb];
        const CandList L2 = cand_start<SP>(g, md, 2, m, p1, mt);
        cb = L2.lo;
        dir = L2.dir;
        ce = gallop_lb(L2.pos, L2.lo, L2.hi, pend, mt);
        ok = ce > cb;
        lvl0 = 1;
        rw = LV_RW;
        p0f = p1;
      }
      // two-edge tail of a 3-edge motif: edge 1 is the second-to-last edge
      int swp = 0, t2lo = 0, t2hi = 0;
      const int pat = (T2 && M == 3) ? tail_pattern<SP>(md) : TP_NONE;
      if (T2 && M == 3 && ok && pat != TP_NONE) {
        cnt += tail_frame_term<SP>(g, md, pat, m, ce - cb, r, pend, mt);
        st_leaves++;
        if (pat == TP_SAME) ok = false; else swp = LV_PAT;
      } else if (T2 && M == 3 && ok && ce - cb >= args.t2_min && tail2_ok<SP>(md)) {
        int blo, bhi, bdir;
        if (mt) meter.b += 4;
        tail2_range<SP>(g, md, m, L.pos[cb], pend, &blo, &bhi, &bdir, mt);
        if (2 * (bhi - blo) < ce - cb) {
          swp = LV_SWAP | (bdir ? LV_BDIR : 0);
          t2lo = cb; t2hi = ce;
          cb = blo; ce = bhi;
          ok = ce > cb;
        }
      }
      int zm = 0;  // injectivity terms known to be 0 in this frame's window
      if (T2 && M == 3 && ok && !(swp & LV_PAT)) zm = tail_frame_mask<SP>(g, md, m, (swp & LV_SWAP) != 0, ce - cb, r, pend, mt);
      unsigned long long bm = __ballot(ok);
      int n = __popcll(bm);
      if (top + n > FR) { if (lane == 0) atomicOr(&sched->error, 1); break; }
      if (ok) {
        r_pend[lane] = pend;  // the window of this lane's root (no frame of an older root is left)
        r_tl[lane] = tli;
        int w = top + __popcll(bm & lt_mask);
        f_cur[w] = cb; f_end[w] = ce; f_lvl[w] = lvl0 | (dir << 8) | swp | rw | (lane << 16) | (zm << 24);
        if (T2 && (swp & LV_SWAP)) { sw_lo[w & (WAVE - 1)] = t2lo; sw_hi[w & (WAVE - 1)] = t2hi; }
        if (T2) {
          f_w[w] = 1;
          f_p0[w] = p0f;
          r_alo[lane] = a1lo;
          r_ahi[lane] = a1hi;
        }
#pragma unroll
        for (int i = 0; i < MAXN; i++) f_m[i * FR + w] = m[i];
        if constexpr (ENUM) f_e[w] = r;
      }
      top += n;
      task_t0 = (long long)wall_clock64();
      __syncthreads();
      continue;
    }

    // ------------------------------------------------------ expansion step
    const tm_query_dev &q = args.queries[qcur];
    const tm_motif_dev &md = q.m;
    const int M = SP::M(md);
    const int last = ENUM ? M - 1 : M - 2;  // last motif edge handled per candidate
    const int Ltop = top > 0 ? uni(f_lvl[top - 1] & 0xff) : -1;
    bool flush = COMPACT && sbn > 0 && (top == 0 || Ltop != sb_level || sbn >= args.sb_flush);
    if (T2 && !flush && top > 0) {
      // a pattern frame of 64 slots or more on top: the wave sweeps its slots
      // 64 at a time (one coalesced load of next same-pair times each; the
      // rare slot whose pair repeats in the window counts those edges)
      const int ft = top - 1;
      const int lvt = uni(f_lvl[ft]);
      const int c0 = uni(f_cur[ft]), c1 = uni(f_end[ft]);
      if ((lvt & LV_PAT) && c1 - c0 >= WAVE) {
        const int rt = (lvt >> 16) & 0xff;
        const int tlt = uni(r_tl[rt]), pendt = uni(r_pend[rt]);
        int m[MAXN];
#pragma unroll
        for (int i = 0; i < MAXN; i++) m[i] = uni(f_m[i * FR + ft]);
        const unsigned long long fwt = (unsigned)uni(f_w[ft]);
        unsigned long long acc = 0;
        for (int sl = c0 + lane; sl < c1; sl += WAVE) acc += tail_same_pair<SP>(g, md, m, sl, tlt, pendt, mt);
        acc *= fwt;
        cnt += tail_pattern<SP>(md) == TP_STAR ? 0ull - acc : acc;
        st_steps++;
        st_items += (unsigned long long)(c1 - c0);
        top = ft;
        __syncthreads();
        continue;
      }
    }
    if (!flush) {
    // ---- deal the next candidate slots of the topmost frames of one level
    // to the lanes (at most lim: room left in the survivor buffer)
    const int L = Ltop;
    const int lim = COMPACT ? WAVE - sbn : WAVE;
    const int fi = top - 1 - lane;
    int fcnt = 0;
    if (fi >= 0) {
      int lv = f_lvl[fi];
      if ((lv & 0xff) == L) fcnt = f_end[fi] - f_cur[fi];
    }
    int incl = wave_incl_scan(fcnt, lane);
    int excl = incl - fcnt;
    int total = uni(__shfl(incl, WAVE - 1));
    if (total > lim) total = lim;
    if (total <= 0) {  // a frame on the stack is never empty; stop rather than spin
      if (lane == 0) atomicOr(&sched->error, 2);
      break;
    }
    const bool part = (fcnt > 0) && (excl < lim);
    const unsigned long long pm = __ballot(part);
    const int nparts = __popcll(pm);
    int take = part ? min(fcnt, lim - excl) : 0;
    const bool last_partial = __ballot(part && take < fcnt) != 0ull;

    int rank, off;
    if (uni(__shfl(fcnt, 0)) >= lim) {
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
    st_steps++;
    st_items += total;
    int slot = 0, pend = 0, dir = 0, tl = 0, ridx = 0, lvw = 0;
    unsigned fw = 1;  // the frame's weight (counting)
    int m[MAXN];
    int e[ENUM ? MAXE : 1];
#pragma unroll
    for (int i = 0; i < MAXN; i++) m[i] = 0;
    if (has) {
      slot = f_cur[myf] + off;
      const int lv = f_lvl[myf];
      lvw = lv;
      if (T2) fw = (unsigned)f_w[myf];
      dir = (lv >> 8) & 1;
      ridx = (lv >> 16) & 0xff;
      pend = r_pend[ridx];
      tl = r_tl[ridx];
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
    const int kind = SP::kind(md, j);
    const int tmn = (j + 1 < M) ? SP::tmode(md, j + 1) : -1;  // pruning rule of edge j + 1
    // a swapped frame (two-edge tail): the slot is a candidate of the last
    // edge; its matches with the earlier candidates of edge j are counted here
    const bool swp = T2 && has && (lvw & LV_SWAP);
    if (T2 && swp) {
      cnt += (unsigned long long)fw *
             tail2_item<SP>(g, md, m, sw_lo[myf & (WAVE - 1)], sw_hi[myf & (WAVE - 1)], dir, slot,
                            (lvw & LV_BDIR) ? 1 : 0, (lvw >> 24) & 0xff, mt);
      st_leaves++;
    }
    // a pattern frame (last two edges on one list): the slot adds or removes
    // the later slots of its own graph pair
    const bool patf = T2 && has && (lvw & LV_PAT);
    if (T2 && patf) {
      const unsigned long long sp = (unsigned long long)fw * tail_same_pair<SP>(g, md, m, slot, tl, pend, mt);
      cnt += tail_pattern<SP>(md) == TP_STAR ? 0ull - sp : sp;
      st_leaves++;
    }
    bool valid = false;
    int pos = 0, tfn = 0, node = -1;
    // position, neighbour and the next edge's first-candidate time: three
    // loads at the same slot of one list, issued together
    const int lst = kind == TM_K_NEW_DST ? TM_L_OUT
                    : (kind == TM_K_NEW_SRC ? TM_L_IN : (dir ? TM_L_INPAIR : TM_L_OUTPAIR));
    if (has && !swp && !patf) {
      if (tmn >= 0) tfn = g.stf[(size_t)(lst * TM_NTFIRST + tmn) * (size_t)g.E + slot];
      // further rules: each later edge (k >= j + 2) at a node of this edge
      // needs that node's next edge after this position in the window (not
      // in the instrumented kernel, which follows the reference's search tree)
      const int xm = INSTR == 1 ? 0 : SP::xmask(md, j);
      // position, neighbour (new node), the next edge's time and the further rules' times
      if (mt) meter.b += 4 * (1 + (kind != TM_K_CLOSE ? 1 : 0) + (tmn >= 0 ? 1 : 0) + __popc((unsigned)xm));
      int tx = -0x7fffffff - 1;
#pragma unroll
      for (int rr = 0; rr < TM_NTFIRST; rr++)
        if ((xm >> rr) & 1) tx = max(tx, g.stf[(size_t)(lst * TM_NTFIRST + rr) * (size_t)g.E + slot]);
      if (kind == TM_K_NEW_DST) {
        pos = g.outPos[slot]; node = g.outNbr[slot];
        valid = !in_first(m, SP::mapped(md, j), node);
        put(m, SP::b(md, j), node);
      } else if (kind == TM_K_NEW_SRC) {
        pos = g.inPos[slot]; node = g.inNbr[slot];
        valid = !in_first(m, SP::mapped(md, j), node);
        put(m, SP::a(md, j), node);
      } else {
        pos = dir ? g.inPairPos[slot] : g.outPairPos[slot];
        valid = true;
      }
      if (tx > tl) valid = false;
      // pair tail: only the first edge of its pair after the frame's start expands
      if (T2 && valid && j == pair_tail_edge<SP>(md)) {
        if (mt) meter.b += 4;
        if (g.sprev[(kind == TM_K_NEW_SRC ? (size_t)g.E : 0) + slot] > f_p0[myf]) valid = false;
      }
    }
    // this candidate's weight: the frame's, times its rank among the root's
    // parallel edges (root-pair rewrite) or its offset in the pair list (pair tail)
    unsigned long long wi = fw;
    if (T2 && valid && (lvw & LV_RW)) {
      const int rk = lower_bound_i(g.outPairPos, r_alo[ridx], r_ahi[ridx], pos, mt) - r_alo[ridx];
      wi = (unsigned long long)rk;
      if (rk <= 0) valid = false;  // (its list starts after the first parallel edge)
    }
    if (T2 && has && (lvw & LV_R1)) wi = (unsigned long long)fw * (unsigned long long)(slot - sw_lo[myf & (WAVE - 1)] + 1);
    // survivors are buffered while they are few and more of this level
    // follows; otherwise (nothing buffered, and sb_inline survivors or the
    // level's last candidates in hand) they are expanded in this step
    bool buffer = false;
    if (COMPACT) {
      const bool surv = valid && (tmn < 0 || tfn <= tl);
      const unsigned long long bm = __ballot(surv);
      const int ns = __popcll(bm);
      const bool level_done = ntop == 0 || uni(f_lvl[ntop - 1] & 0xff) != L;
      buffer = sbn > 0 || (ns < args.sb_inline && !level_done);
      if (buffer) {
        if (surv) {
          const int w = sbn + __popcll(bm & lt_mask);
          sb_pos[w] = pos;
          sb_ridx[w] = ridx;
          sb_wl[w] = (int)(unsigned)(wi & 0xffffffffull);
          sb_wh[w] = (int)(unsigned)(wi >> 32);
#pragma unroll
          for (int i = 0; i < MAXN; i++) sb_m[i * WAVE + w] = m[i];
        }
        sbn += ns;
        sb_level = L;
        top = ntop;
        __syncthreads();
        flush = sbn > 0 && (level_done || sbn >= args.sb_flush);
      }
    }
    if (!buffer) {
    if constexpr (INSTR == 1) {
      if (valid && j + 1 < M) {
        (SP::kind(md, j + 1) == TM_K_CLOSE ? rx2 : rx1) += 1;
        rp += ref_probes<SP>(g, md, j + 1, m, pos, pend);
      }
    }
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
          if ((long long)w + 1 >= q.cap) __hip_atomic_store(q.stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      } else {
        if (valid && (tmn < 0 || tfn <= tl)) {
          const CandList LL = cand_start<SP>(g, md, j + 1, m, pos, mt);
          int skipx = -1;
          const int cr = T2 ? tail_crule<SP>(md) : -1;
          if (cr >= 0) {
            if (mt) meter.b += 4;
            if (g.stf[(size_t)(lst * TM_NTFIRST + (cr & 0xff)) * (size_t)g.E + slot] > tl) skipx = sel(m, cr >> 8);
          }
          cnt += wi * leaf_count<SP>(g, md, j + 1, m, LL, pos, pend, args.scan_max, skipx, (lvw >> 24) & 0xff, mt);
          st_leaves++;
        }
      }
      top = ntop;
      __syncthreads();
      continue;
    }
    int cb = 0, ce = 0, cdir = 0;
    int swc = 0, t2lo = 0, t2hi = 0, zm = 0;
    if (valid && tmn >= 0) valid = tfn <= tl;
    if (valid) {
      const CandList LL = cand_start<SP>(g, md, j + 1, m, pos, mt);
      cb = LL.lo;
      cdir = LL.dir;
      ce = gallop_lb(LL.pos, LL.lo, LL.hi, pend, mt);
      valid = ce > cb;
      const int pat = (T2 && j == M - 3) ? tail_pattern<SP>(md) : TP_NONE;
      if (T2 && valid && j == M - 3 && j == pair_tail_edge<SP>(md)) {
        swc = LV_R1;  // pair tail: the pair list after this first edge, weighted by offset
        t2lo = cb;
      } else if (T2 && valid && j == M - 3 && pat != TP_NONE) {
        cnt += wi * tail_frame_term<SP>(g, md, pat, m, ce - cb, pos, pend, mt);
        st_leaves++;
        if (pat == TP_SAME) valid = false; else swc = LV_PAT;
      } else if (T2 && valid && j == M - 3 && ce - cb >= args.t2_min && tail2_ok<SP>(md)) {
        int blo, bhi, bdir;
        if (mt) meter.b += 4;
        tail2_range<SP>(g, md, m, LL.pos[cb], pend, &blo, &bhi, &bdir, mt);
        if (2 * (bhi - blo) < ce - cb) {
          swc = LV_SWAP | (bdir ? LV_BDIR : 0);
          t2lo = cb; t2hi = ce;
          cb = blo; ce = bhi;
          valid = ce > cb;
        }
      }
      if (T2 && valid && j == M - 3 && !(swc & LV_PAT))
        zm = tail_frame_mask<SP>(g, md, m, (swc & LV_SWAP) != 0, ce - cb, pos, pend, mt);
    }
    const unsigned long long bm = __ballot(valid);
    const int n = __popcll(bm);
    if (ntop + n > FR) { if (lane == 0) atomicOr(&sched->error, 1); break; }
    if (valid) {
      int w = ntop + __popcll(bm & lt_mask);
      f_cur[w] = cb; f_end[w] = ce; f_lvl[w] = j | (cdir << 8) | swc | (ridx << 16) | (zm << 24);
      if (T2 && (swc & (LV_SWAP | LV_R1))) { sw_lo[w & (WAVE - 1)] = t2lo; sw_hi[w & (WAVE - 1)] = t2hi; }
      if (T2) {
        f_w[w] = (int)(unsigned)wi;
        f_p0[w] = pos;
      }
#pragma unroll
      for (int i = 0; i < MAXN; i++) f_m[i * FR + w] = m[i];
      if constexpr (ENUM) {
#pragma unroll
        for (int i = 0; i < MAXE; i++)
          if (i < j) f_e[i * FR + w] = e[i];
        f_e[j * FR + w] = pos;
      }
    }
    st_push += n;
    top = ntop + n;
    __syncthreads();
    }  // expanded in this step
    }  // deal and filter
    if (flush) {
      // ---- expand the buffered survivors, one per lane: the long chains of
      // dependent loads (candidate list of the next edge, window end, leaf
      // counts) run with every lane busy instead of once per filter step
      const int j = sb_level + 1;  // motif edge of the buffered candidates
      const bool has = lane < sbn;
      int pos = 0, ridx = 0, pend = 0;
      unsigned long long wi = 1;
      int m[MAXN];
#pragma unroll
      for (int i = 0; i < MAXN; i++) m[i] = 0;
      if (has) {
        pos = sb_pos[lane];
        ridx = sb_ridx[lane];
        wi = ((unsigned long long)(unsigned)sb_wh[lane] << 32) | (unsigned)sb_wl[lane];
        pend = r_pend[ridx];
#pragma unroll
        for (int i = 0; i < MAXN; i++) m[i] = sb_m[i * WAVE + lane];
      }
      st_steps++;
      sbn = 0;
      if (j == last) {
        if (has) {
          const CandList LL = cand_start<SP>(g, md, j + 1, m, pos, mt);
          int skipx = -1;
          const int cr = tail_crule<SP>(md);
          if (cr >= 0) {
            if (mt) meter.b += 4;
            const int rule = cr & 0xff;  // uniform: a branch, not an indexed pointer table
            const int tn = rule == TM_T_PAIR_SELF ? g.tfirst[TM_T_PAIR_SELF][pos] : g.tfirst[TM_T_RECIP][pos];
            if (tn > r_tl[ridx]) skipx = sel(m, cr >> 8);
          }
          cnt += wi * leaf_count<SP>(g, md, j + 1, m, LL, pos, pend, args.scan_max, skipx, 0, mt);
          st_leaves++;
        }
        __syncthreads();
        continue;
      }
      int cb = 0, ce = 0, cdir = 0;
      int swc = 0, t2lo = 0, t2hi = 0, zm = 0;
      bool valid = has;
      if (valid) {
        const CandList LL = cand_start<SP>(g, md, j + 1, m, pos, mt);
        cb = LL.lo;
        cdir = LL.dir;
        ce = gallop_lb(LL.pos, LL.lo, LL.hi, pend, mt);
        valid = ce > cb;
        const int pat = (T2 && j == M - 3) ? tail_pattern<SP>(md) : TP_NONE;
        if (T2 && valid && j == M - 3 && j == pair_tail_edge<SP>(md)) {
          swc = LV_R1;
          t2lo = cb;
        } else if (T2 && valid && j == M - 3 && pat != TP_NONE) {
          cnt += wi * tail_frame_term<SP>(g, md, pat, m, ce - cb, pos, pend, mt);
          st_leaves++;
          if (pat == TP_SAME) valid = false; else swc = LV_PAT;
        } else if (T2 && valid && j == M - 3 && ce - cb >= args.t2_min && tail2_ok<SP>(md)) {
          int blo, bhi, bdir;
          if (mt) meter.b += 4;
          tail2_range<SP>(g, md, m, LL.pos[cb], pend, &blo, &bhi, &bdir, mt);
          if (2 * (bhi - blo) < ce - cb) {
            swc = LV_SWAP | (bdir ? LV_BDIR : 0);
            t2lo = cb; t2hi = ce;
            cb = blo; ce = bhi;
            valid = ce > cb;
          }
        }
        if (T2 && valid && j == M - 3 && !(swc & LV_PAT))
          zm = tail_frame_mask<SP>(g, md, m, (swc & LV_SWAP) != 0, ce - cb, pos, pend, mt);
      }
      const unsigned long long bm = __ballot(valid);
      const int n = __popcll(bm);
      if (top + n > FR) { if (lane == 0) atomicOr(&sched->error, 1); break; }
      if (valid) {
        const int w = top + __popcll(bm & lt_mask);
        f_cur[w] = cb; f_end[w] = ce; f_lvl[w] = j | (cdir << 8) | swc | (ridx << 16) | (zm << 24);
        if (T2 && (swc & (LV_SWAP | LV_R1))) { sw_lo[w & (WAVE - 1)] = t2lo; sw_hi[w & (WAVE - 1)] = t2hi; }
        if (T2) {
          f_w[w] = (int)(unsigned)wi;
          f_p0[w] = pos;
        }
#pragma unroll
        for (int i = 0; i < MAXN; i++) f_m[i * FR + w] = m[i];
      }
      st_push += n;
      top += n;
      __syncthreads();
    }  // flush
    // ---------------------------------------------- donate to idle wavefronts
    // every steal_every steps; only a wavefront that has held its current
    // task for steal_after_ticks donates: short tasks never pay for the exchange
    bool look = false;
    if (top > 0 && --steal_cd <= 0) {
      steal_cd = args.steal_every;
      look = (long long)wall_clock64() - task_t0 > args.steal_after_ticks;
    }
    if (look) {
      // donate only against unmet demand: tickets drawn minus records reserved
      int want = 0;
      if (lane == 0) {
        const int hd = (int)ld_rlx(&sched->pool_head);
        const int tl = (int)ld_rlx(&sched->pool_tail);
        want = min(hd - tl, args.steal_max);
      }
      want = uni(__shfl(want, 0));
      for (int d = 0; d < want; d++) {
        // split the shallowest frame with at least 2 * steal_min candidates left
        // (a swapped frame keeps its edge M-2 range in this wavefront's side table: not donated)
        const bool can = lane < top && !(f_lvl[lane] & (LV_SWAP | LV_RW | LV_R1)) &&
                         (f_end[lane] - f_cur[lane] >= 2 * args.steal_min);
        const unsigned long long cm = __ballot(can);
        if (cm == 0ull) break;
        const int b = __ffsll((long long)cm) - 1;
        unsigned t = 0;
        int ok = 1;
        if (lane == 0) {
          add_rlx(&sched->busy, 1);  // the record's credit, before it can be seen
          t = atomicAdd(&sched->pool_tail, 1u);
          const unsigned lap = t >> args.pool_shift;
          if (lap > 0) {
            // the slot's previous lap must have been copied out by its taker
            // (a resident wavefront holding that ticket, so this wait ends)
            const unsigned prev = pool_tag(gen, lap - 1);
            unsigned w = 0;
            while (ld_rlx32(args.pool_done + (t & (unsigned)(args.pool_cap - 1))) != prev) {
              if (ld_rlx(&sched->abort) || ++w > (1u << 24)) { ok = 0; break; }
              __builtin_amdgcn_s_sleep(8);
            }
            if (!ok) {
              // record t will never be written, so its ticket holder could
              // only leave through the abort flag: raise it, and every
              // wavefront (waiting ones included) leaves; the host reports
              // the error bit as a failed launch
              atomicOr(&sched->error, 8);
              atomicOr(&sched->abort, 1);
            }
          }
        }
        t = (unsigned)uni((int)__shfl((int)t, 0));
        if (!uni(__shfl(ok, 0))) break;  // abort: leave the donation loop; the main loop sees the flag
        const int cur0 = f_cur[b], end0 = f_end[b];
        const int mid = cur0 + (end0 - cur0) / 2;
        const unsigned tag = pool_tag(gen, t >> args.pool_shift);
        unsigned long long *rec = args.pool + (size_t)(t & (unsigned)(args.pool_cap - 1)) * TM_POOL_FIELDS;
        if (lane < TM_POOL_FIELDS) {
          int v = 0;
          if (lane == 0) v = qcur;
          else if (lane == 2) v = mid;
          else {
            int *src = frame_field(lane, b);
            v = src ? *src : 0;
          }
          st_rlx(rec + lane, ((unsigned long long)tag << 32) | (unsigned)v);
        }
        __syncthreads();
        if (lane == 0) f_end[b] = mid;
        st_don++;
        __syncthreads();
      }
    }
  }

  if (!ENUM && qcur >= 0) {
    unsigned long long s = wave_sum64(cnt);
    if (lane == 0 && s) q_cnt[qcur] += s;
  }
  __syncthreads();
  if (!ENUM)
    for (int q = lane; q < args.nq; q += WAVE)
      if (q_cnt[q]) atomicAdd(args.queries[q].count, q_cnt[q]);
  if (!is_idle) st_work += (unsigned long long)((long long)wall_clock64() - work_t0);
  if (lane == 0 && !is_idle) add_rlx(&sched->busy, -1);  // left early holding work (error or deadline)
  if (lane == 0 && waiter) add_rlx(&sched->waiters, -1);
  // work counters: plain stores into this wavefront's own slot, summed by a
  // one-block kernel after the launch (no same-address atomics from every wave)
  {
    unsigned long long v = 0;
    switch (lane) {
      case TM_STAT_STEPS: v = st_steps; break;
      case TM_STAT_ITEMS: v = st_items; break;
      case TM_STAT_ROOTS: v = st_roots; break;
      case TM_STAT_PUSHES: v = st_push; break;
      case TM_STAT_DONATED: v = st_don; break;
      case TM_STAT_TAKEN: v = st_take; break;
      case TM_STAT_CAPEXIT: v = st_capexit; break;
      case TM_STAT_WORK: v = st_work; break;
      default: break;
    }
    const unsigned long long leaves = wave_sum64(st_leaves);
    if (lane == TM_STAT_LEAVES) v = leaves;
    if constexpr (INSTR == 2) {
      const unsigned long long a = wave_sum64(meter.b);
      if (lane == TM_STAT_REF_X1) v = a;  // bytes read by the search (metered variant)
    }
    if constexpr (INSTR == 1) {
      const unsigned long long a = wave_sum64(rx1), b = wave_sum64(rx2), c = wave_sum64(rp);
      if (lane == TM_STAT_REF_X1) v = a;
      if (lane == TM_STAT_REF_X2) v = b;
      if (lane == TM_STAT_REF_P) v = c;
    }
    if (lane < TM_NSTATS) args.wstats[(size_t)blockIdx.x * TM_NSTATS + lane] = v;
  }
}

}  // namespace tmk
