// tm_abi.hip -- host driver of the gfx950 temporal-motif kernels, behind the
// C ABI declared in include/tm_abi.h.
//
// Graph residency: a graph is uploaded once and stays in HBM with the
// reference arrays (Eg, in/out CSR, Euid) plus indexes built on the device at
// upload: packed times T, neighbour arrays aligned with the CSR position
// lists, a per-node (neighbour, position)-sorted copy of each list (the pair
// index), and a time-bucket index for the window end.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tm_abi.h"
#include "tm_kernel.cuh"
#include "tm_internal.h"
#include "tm_variants.h"

namespace {

thread_local std::string g_last_error;

tm_status fail(tm_status s, const std::string &msg) {
  g_last_error = msg;
  return s;
}

#define HIPTRY(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(e_ == hipErrorOutOfMemory ? TM_ERR_NOMEM : TM_ERR_HIP,               \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                  \
  } while (0)

__global__ void aux_kernel(const tm_edge *__restrict__ Eg, const int *__restrict__ outPos,
                           const int *__restrict__ inPos, int E, int *__restrict__ T,
                           int *__restrict__ outNbr, int *__restrict__ inNbr,
                           unsigned long long *__restrict__ okey, unsigned long long *__restrict__ ikey) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E) return;
  T[i] = Eg[i].t;
  int po = outPos[i];
  tm_edge eo = Eg[po];
  outNbr[i] = eo.v;
  okey[i] = ((unsigned long long)(unsigned)eo.u << 32) | (unsigned)eo.v;
  int pi = inPos[i];
  tm_edge ei = Eg[pi];
  inNbr[i] = ei.u;
  ikey[i] = ((unsigned long long)(unsigned)ei.v << 32) | (unsigned)ei.u;
}

// slot of each position inside the list that holds it
__global__ void self_kernel(const int *__restrict__ P, int E, int *__restrict__ self) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < E) self[P[s]] = s;
}

// outNext[p] = first slot of out(dst p) after p; inNext[p] = first slot of in(src p) after p
__global__ void next_kernel(const tm_edge *__restrict__ Eg, int E, const int *__restrict__ outR,
                            const int *__restrict__ outPos, const int *__restrict__ inR,
                            const int *__restrict__ inPos, int *__restrict__ outNext, int *__restrict__ inNext) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  tm_edge e = Eg[p];
  int lo = outR[e.v], hi = outR[e.v + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (outPos[mid] <= p) lo = mid + 1; else hi = mid;
  }
  outNext[p] = lo;
  lo = inR[e.u];
  hi = inR[e.u + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (inPos[mid] <= p) lo = mid + 1; else hi = mid;
  }
  inNext[p] = lo;
}

// Time of the first candidate after p for each start rule (INT_MAX: none).
__global__ void tfirst_kernel(const tm_edge *__restrict__ Eg, const int *__restrict__ T, int E,
                              const int *__restrict__ outR, const int *__restrict__ outPos,
                              const int *__restrict__ inR, const int *__restrict__ inPos,
                              const int *__restrict__ outSelf, const int *__restrict__ inSelf,
                              const int *__restrict__ outNext, const int *__restrict__ inNext,
                              const int *__restrict__ outPairNbr, const int *__restrict__ outPairPos,
                              const int *__restrict__ outPairSelf, int *t0, int *t1, int *t2, int *t3, int *t4,
                              int *t5) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const tm_edge e = Eg[p];
  const int NONE = 0x7fffffff;
  int s = outSelf[p] + 1;
  t0[p] = s < outR[e.u + 1] ? T[outPos[s]] : NONE;
  s = outNext[p];
  t1[p] = s < outR[e.v + 1] ? T[outPos[s]] : NONE;
  s = inSelf[p] + 1;
  t2[p] = s < inR[e.v + 1] ? T[inPos[s]] : NONE;
  s = inNext[p];
  t3[p] = s < inR[e.u + 1] ? T[inPos[s]] : NONE;
  s = outPairSelf[p] + 1;
  t4[p] = (s < outR[e.u + 1] && outPairNbr[s] == e.v) ? T[outPairPos[s]] : NONE;
  // reciprocal: first edge v -> u after p, in the out-pair list of v
  int lo = outR[e.v], hi = outR[e.v + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (outPairNbr[mid] < e.u || (outPairNbr[mid] == e.u && outPairPos[mid] <= p)) lo = mid + 1; else hi = mid;
  }
  t5[p] = (lo < outR[e.v + 1] && outPairNbr[lo] == e.u) ? T[outPairPos[lo]] : NONE;
}

// First-candidate times in the slot order of one candidate list.
__global__ void stf_kernel(const int *__restrict__ pos, int E, const int *t0, const int *t1, const int *t2,
                           const int *t3, const int *t4, const int *t5, int *o0, int *o1, int *o2, int *o3, int *o4,
                           int *o5) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= E) return;
  const int p = pos[s];
  o0[s] = t0[p];
  o1[s] = t1[p];
  o2[s] = t2[p];
  o3[s] = t3[p];
  o4[s] = t4[p];
  o5[s] = t5[p];
}

// Root gaps: first-candidate time minus the root's own time, per rule.
__global__ void rgap_kernel(const tm_edge *__restrict__ Eg, int E, const int *t0, const int *t1, const int *t2,
                            const int *t3, const int *t4, const int *t5, unsigned *g0, unsigned *g1, unsigned *g2,
                            unsigned *g3, unsigned *g4, unsigned *g5) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const tm_edge e = Eg[p];
  const int *tf[TM_NTFIRST] = {t0, t1, t2, t3, t4, t5};
  unsigned *out[TM_NTFIRST] = {g0, g1, g2, g3, g4, g5};
#pragma unroll
  for (int k = 0; k < TM_NTFIRST; k++) {
    const int f = tf[k][p];
    unsigned gap = 0xffffffffu;
    if (e.u != e.v && f != 0x7fffffff) {
      const long long d = (long long)f - e.t;
      gap = d > 0x7fffffffll ? 0x7fffffffu : (unsigned)d;
    }
    out[k][p] = gap;
  }
}

// Previous edge of the same pair, per slot of the out- and in-lists: the
// entry before the edge's own slot in its source's out-pair index.
__global__ void sprev_kernel(const tm_edge *__restrict__ Eg, const int *__restrict__ outPos,
                             const int *__restrict__ inPos, const int *__restrict__ outR,
                             const int *__restrict__ outPairNbr, const int *__restrict__ outPairPos,
                             const int *__restrict__ outPairSelf, int E, int *__restrict__ sprev) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= E) return;
  for (int k = 0; k < 2; k++) {
    const int p = k ? inPos[s] : outPos[s];
    const tm_edge e = Eg[p];
    const int ps = outPairSelf[p];
    int prev = -1;
    if (ps > outR[e.u] && outPairNbr[ps - 1] == e.v) prev = outPairPos[ps - 1];
    sprev[(size_t)k * E + s] = prev;
  }
}

__global__ void low_word_kernel(const unsigned long long *__restrict__ k, int *__restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int)(unsigned)(k[i] & 0xffffffffull);
}

__global__ void bucket_kernel(const int *__restrict__ T, int E, int tmin, int bshift, long long nb,
                              int *__restrict__ bucket) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nb) return;
  long long key = (long long)tmin + (k << bshift);
  int lo = 0, hi = E;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if ((long long)T[mid] < key) lo = mid + 1; else hi = mid;
  }
  bucket[k] = lo;
}

// Every field the mining kernel then reads or updates with device-scope
// atomics is written here with device-scope atomic stores too.
__device__ __forceinline__ void st_dev(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int ld_dev(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the error bits of the launch just finished, kept for the caller (a batch
// of single-query launches checks every launch's bits at the end)
__global__ void error_kernel(const tm_sched *s, int *out) {
  if (threadIdx.x == 0) *out = ld_dev(&s->error);
}

__global__ void reset_kernel(tm_sched *s, unsigned long long *counts, int *stops, int n, int grid) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < TM_STRIPES) st_dev(&s->next_chunk[i][0], 0ull);
  if (i == 0) {
    st_dev(&s->error, 0);
    st_dev(&s->abort, 0);
    st_dev(&s->pool_head, 0u);
    st_dev(&s->pool_tail, 0u);
    st_dev(&s->busy, grid);  // every wavefront starts out holding (root) work
    st_dev(&s->waiters, 0);
    // fresh ring tags for this launch (the host clears the ring before 2^18)
    __hip_atomic_fetch_add(&s->gen, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&s->t0, (long long)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (i < 16) st_dev(&s->stats[i], 0ull);
  if (i < n) { st_dev(&counts[i], 0ull); st_dev(&stops[i], 0); }
}

__global__ void clear_pool_kernel(unsigned long long *pool, size_t nw, unsigned *done, size_t nd, tm_sched *s) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) pool[i] = 0;
  if (i < nd) done[i] = 0;
  if (i == 0) __hip_atomic_store(&s->gen, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // next reset: 1
}

// Sum the per-wavefront counters of a launch into sched->stats (one block).
__global__ void __launch_bounds__(256) stats_kernel(const unsigned long long *__restrict__ w, int grid, tm_sched *s) {
  __shared__ unsigned long long acc[256];
  const int k = threadIdx.x % TM_NSTATS, part = threadIdx.x / TM_NSTATS;  // 16 partial sums per counter
  unsigned long long v = 0;
  for (int b = part; b < grid; b += 256 / TM_NSTATS) v += w[(size_t)b * TM_NSTATS + k];
  acc[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x < TM_NSTATS) {
    unsigned long long t = 0;
    for (int p = 0; p < 256 / TM_NSTATS; p++) t += acc[p * TM_NSTATS + threadIdx.x];
    s->stats[threadIdx.x] = t;
  }
}

int ceil_log2(long long x) {
  int r = 0;
  while ((1ll << r) < x) r++;
  return r;
}

}  // namespace

// ----------------------------------------------------------------- graph

namespace tmh {

tm_status set_error(tm_status s, const std::string &msg) { return fail(s, msg); }

int shape_class_of(int M, const int32_t *src, const int32_t *dst) {
  int n = 0;
  for (int k = 0; k < M; k++) n = std::max(n, 1 + std::max(src[k], dst[k]));
  return std::max(1, M - 2) * 2 + (n <= 4 ? 0 : 1);
}

// The tuning table entry of a motif's shape class, if one is installed.
// The launch parameters of a call: its own, unless it brings none (a null
// pointer or every field 0, both meaning "the library's choice"), in which
// case the tuning installed for the motif's shape class, if any.
const tm_tuning *tuning_for_class(tm_graph *g, int cls, const tm_tuning *tune) {
  static const tm_tuning none{};
  if (tune && memcmp(tune, &none, sizeof(none)) != 0) return tune;
  if (cls >= 0 && cls < TM_NCLASSES && g->class_set[cls]) return &g->class_tune[cls];
  return nullptr;
}

const tm_tuning *class_tuning_of(tm_graph *g, int M, const int32_t *src, const int32_t *dst) {
  const int cls = shape_class_of(M, src, dst);
  return (cls >= 0 && cls < TM_NCLASSES && g->class_set[cls]) ? &g->class_tune[cls] : nullptr;
}

tm_status dev_alloc(tm_graph *g, void **p, size_t bytes) {
  if (bytes == 0) bytes = 4;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? TM_ERR_NOMEM : TM_ERR_HIP,
                std::string("hipMalloc: ") + hipGetErrorString(e));
  g->allocs.push_back(*p);
  g->bytes += (int64_t)bytes;
  return TM_OK;
}

// Build the device descriptor of a motif given as renamed (src, dst) pairs.
// TM_T_* rule that a later motif edge a -> b imposes at the graph edge matched
// for the motif edge pa -> pb (relative to that edge's position), or TM_T_NONE
static int rule_at(int pa, int pb, int a, int b) {
  if (a == pa && b == pb) return TM_T_PAIR_SELF;
  if (a == pb && b == pa) return TM_T_RECIP;
  if (a == pa) return TM_T_OUT_SELF;
  if (a == pb) return TM_T_OUT_NEXT;
  if (b == pb) return TM_T_IN_SELF;
  if (b == pa) return TM_T_IN_NEXT;
  return TM_T_NONE;
}

tm_status build_motif(int M, const int32_t *src, const int32_t *dst, tm_motif_dev *md) {
  if (M < 2 || M > TM_MAXE) return fail(TM_ERR_MOTIF, "motif must have 2..8 edges");
  memset(md, 0, sizeof(*md));
  for (int j = 0; j < TM_MAXE; j++) md->tmode[j] = TM_T_NONE;
  md->M = M;
  if (src[0] != 0 || dst[0] != 1)
    return fail(TM_ERR_MOTIF, "motif nodes must be renamed by first appearance (edge 0 = 0->1)");
  int n = 2;
  for (int j = 1; j < M; j++) {
    int a = src[j], b = dst[j];
    if (a < 0 || b < 0) return fail(TM_ERR_MOTIF, "negative motif node id");
    if (a == b) return fail(TM_ERR_MOTIF, "motif self-loop");
    md->a[j] = a;
    md->b[j] = b;
    md->mapped[j] = n;
    const int pa = src[j - 1], pb = dst[j - 1];
    if (a < n && b < n) {
      md->kind[j] = TM_K_CLOSE;
      md->smode[j] = (a == pa && b == pb) ? TM_S_PAIR_SELF : TM_S_SEARCH;
      // pruning rule: the pair's own next edge when edge j-1 has the same or
      // the reversed pair; otherwise a node edge j shares with edge j-1 must
      // have an edge of the right direction after p (the pair's edges are
      // among them), e.g. a == pb: some out-edge of m[a] after p
      md->tmode[j] = rule_at(pa, pb, a, b);
    } else if (a < n && b == n) {
      md->kind[j] = TM_K_NEW_DST;
      md->smode[j] = a == pa ? TM_S_OUT_SELF : (a == pb ? TM_S_OUT_NEXT : TM_S_SEARCH);
      md->tmode[j] = a == pa ? TM_T_OUT_SELF : (a == pb ? TM_T_OUT_NEXT : TM_T_NONE);
      n++;
    } else if (b < n && a == n) {
      md->kind[j] = TM_K_NEW_SRC;
      md->smode[j] = b == pb ? TM_S_IN_SELF : (b == pa ? TM_S_IN_NEXT : TM_S_SEARCH);
      md->tmode[j] = b == pb ? TM_T_IN_SELF : (b == pa ? TM_T_IN_NEXT : TM_T_NONE);
      n++;
    } else {
      return fail(TM_ERR_MOTIF, "motif is not prefix-connected with first-appearance node ids");
    }
    if (n > TM_MAXN) return fail(TM_ERR_MOTIF, "motif has more than 8 nodes");
  }
  md->n = n;
  // further pruning rules per edge j: every later edge k >= j + 2 that meets
  // a node of edge j needs that node's next edge of its direction after edge
  // j's position inside the window.  Rules implied by another are dropped
  // (the pair's own next edge is an out-edge of its source and an in-edge of
  // its destination; the reversed pair's likewise), and so is edge j + 1's.
  for (int j = 0; j + 2 < M; j++) {
    int mask = 0;
    for (int k = j + 2; k < M; k++) {
      const int r = rule_at(src[j], dst[j], src[k], dst[k]);
      if (r != TM_T_NONE) mask |= 1 << r;
    }
    const int next = md->tmode[j + 1] >= 0 ? 1 << md->tmode[j + 1] : 0;
    const int have = mask | next;
    if (have & (1 << TM_T_PAIR_SELF)) mask &= ~((1 << TM_T_OUT_SELF) | (1 << TM_T_IN_SELF));
    if (have & (1 << TM_T_RECIP)) mask &= ~((1 << TM_T_OUT_NEXT) | (1 << TM_T_IN_NEXT));
    mask &= ~next;
    md->xmask[j] = mask;
  }
  return TM_OK;
}

// The window sweep costs a fixed O(E log E) per query; the mining kernel
// costs about the windows' slots summed over the roots.  Their mean is
// estimated as sum_x deg_root(x) * deg_window(x) * min(1, delta / time span)
// / E over the anchor nodes x; measured on a 64M-edge power-law graph the
// two cost the same near 8 slots per root (mean windows of 0.02 to 3 slots:
// mining 1-30 ms, sweep 15-60 ms; 22 to 350 slots: mining 0.1-3.3 s).
static bool sweep_pays(const tm_graph *g, const SweepItem &it, int delta) {
  const bool root_out = it.anchor == 0, win_out = it.out != 0;
  const double d2 = root_out == win_out ? g->deg2[root_out ? 0 : 1] : g->deg2[2];
  const double frac = g->tspan > 0 ? std::min(1.0, (double)delta / g->tspan) : 1.0;
  return d2 * frac >= 8.0 * (double)g->E;
}

// The root-pair sweep replaces the expansion of every root's edge-1 slots up
// to its pair's last edge in its window by one expansion per pair: it pays
// when those windows are long (the same estimate as the window sweep's, over
// edge 1's list, against PAIR_SLOTS mean slots per root).  Measured on a
// 64M-edge power-law graph: a fixed 18 ms; 0 -> 1, 0 -> 2, 3 -> 2, 0 -> 1 at
// estimates 2 / 12 / 50 / 80 / 274: sweep 20 / 37 / 162 / 287 / 1293 ms,
// mining 4 / 36 / 289 / 614 / 4624 ms; 0 -> 2, 2 -> 3 at 50 / 346: sweep
// 34 / 150, mining 22 / 253.
static const double PAIR_SLOTS = 40.0;
static bool pair_pays(const tm_graph *g, const PairItem &it, int delta) {
  const bool root_out = it.s1 == 0, win_out = it.d1 != 0;
  const double d2 = root_out == win_out ? g->deg2[root_out ? 0 : 1] : g->deg2[2];
  const double frac = g->tspan > 0 ? std::min(1.0, (double)delta / g->tspan) : 1.0;
  return d2 * frac >= PAIR_SLOTS * (double)g->E;
}

// Kernel variants (tm_variants.h): node map of 4 or 8, count / enumerate /
// instrumented count; counting also comes with register caps of 96 and 80
// VGPRs, and in a single-query and a multi-query (batch) form.
// OCC: report the occupancy in *blocks instead of launching.
template <bool OCC>
static tm_status dispatch(Slot &sl, int nmap, bool enumerate, int instr, int wpe, bool multi, int lds, int *blocks,
                          const tm_kernel_args *a, int grid) {
  return tmv::pick(nmap, enumerate, instr, wpe, multi)(OCC, sl.stream, lds, blocks, a, grid);
}

// Plan and launch the mining kernel for staged queries [q0, q0 + nq) of
// scratch.h_queries (nq <= TM_LAUNCH_MAXQ): size the persistent grid from
// occupancy, pick the chunk size over the union of the root ranges, upload,
// reset, launch.
tm_status plan(tm_graph *g, Slot &sl, int q0, int nq, bool enumerate, const tm_tuning *tune, int instr, bool multi,
               tm_launch_fn fn, void *user, Plan *pl) {
  if (nq > 1 && (enumerate || instr || fn || !multi)) return fail(TM_ERR_INTERNAL, "several queries need the batch kernel");
  Scratch &s = sl.s;
  // counting queries of a window-sweep shape go to that kernel (unless the
  // caller's tuning keeps them on the mining kernel); the mining launch
  // sees them with an empty root range
  pl->sweep.clear();
  pl->pair.clear();
  std::vector<char> swept(nq, 0);
  if (!enumerate && instr == 0 && !fn && !(tune && tune->sweep < 0)) {
    for (int i = 0; i < nq; i++) {
      SweepItem it;
      PairItem pi;
      if (sweep_shape(s.h_queries[q0 + i].m, &it) &&
          ((tune && tune->sweep > 0) || sweep_pays(g, it, s.h_queries[q0 + i].delta))) {
        it.q = i;
        pl->sweep.push_back(it);
        swept[i] = 1;
      } else if (pair_shape(s.h_queries[q0 + i].m, &pi) &&
                 ((tune && tune->sweep > 0) || pair_pays(g, pi, s.h_queries[q0 + i].delta))) {
        pi.q = i;
        pl->pair.push_back(pi);
        swept[i] = 1;
      }
    }
  }
  auto range_of = [&](int i, int *b, int *e) {
    const tm_query_dev &qd = s.h_queries[q0 + i];
    *b = qd.root_beg;
    *e = swept[i] ? qd.root_beg : qd.root_end;
  };
  int maxM = 2, maxN = 2;
  long long span = 0;  // roots over all queries (aligned down to 4 per query)
  for (int i = 0; i < nq; i++) {
    const tm_query_dev &qd = s.h_queries[q0 + i];
    if (qd.m.tmode[1] < 0 || qd.m.tmode[1] >= TM_NTFIRST) return fail(TM_ERR_INTERNAL, "edge 1 has no start rule");
    maxM = std::max(maxM, qd.m.M);
    maxN = std::max(maxN, qd.m.n);
    int rb, re;
    range_of(i, &rb, &re);
    if (re > rb) span += re - (rb & ~3);
  }
  const int lvls = enumerate ? std::max(1, maxM - 1) : std::max(1, maxM - 2);
  const int frames = 64 * lvls;
  const int nmap = maxN <= 4 ? 4 : 8;
  // cur, end, level word, node map, positions (enumeration), weight and list start (counting)
  const bool counting = !enumerate && instr != 1;
  const int per_frame = 3 + nmap + (enumerate ? TM_MAXE : 0) + (counting ? 2 : 0);
  // survivor compaction of the counting kernel: survivors buffered before a
  // full-width expansion (0: expanded in the step that filtered them), and
  // the survivors with which a step expands its own when nothing is buffered
  int sb_flush = 0, sb_inline = 0;
  bool sb_auto = false;  // the library's choice: on unless its buffer costs resident wavefronts
  if (!enumerate && instr != 1) {  // the counting and the metered kernels
    sb_flush = TM_SB_FLUSH_DEFAULT;
    sb_inline = TM_SB_INLINE_DEFAULT;
    sb_auto = !(tune && tune->sb_flush != 0);
    if (!sb_auto) sb_flush = tune->sb_flush < 0 ? 0 : std::min(tune->sb_flush, tmk::WAVE);
    if (tune && tune->sb_inline > 0) sb_inline = std::min(tune->sb_inline, tmk::WAVE);
  }
  // frame stack (an even number of ints), the roots' window table (2 x 65
  // ints), the survivor buffer (counting), then the launch's per-query counts
  auto lds_of = [&](int flush) {
    const int sb = flush > 0 ? (4 + nmap) * tmk::WAVE : 0;  // position, root, weight (2), node map
    const int t2 = counting ? 2 * tmk::WAVE : 0;             // two-edge tail side table
    const int roots = (counting ? 4 : 2) * tmk::TM_ROOTS;    // window, and the root-pair list (counting)
    return (frames * per_frame + roots + sb + t2) * (int)sizeof(int) + nq * (int)sizeof(unsigned long long);
  };
  int wpe = 0;  // register cap of the counting kernel: 1 (none), 4 (128 VGPRs), 5 (96) or 6 (80)
  if (tune && tune->waves_per_simd_hint > 0) {
    const int h = tune->waves_per_simd_hint;
    wpe = h >= 6 ? 6 : h == 5 ? 5 : h == 4 ? 4 : 1;
  }
  if (!fn && wpe == 0) {
    wpe = 5;  // measured best on the 64M-edge set over 15 heavy queries (5: 7.00 s, 6: 7.81 s, 1: 7.68 s);
              // the autotuner picks 6 per shape where it wins
  }
  auto blocks_of = [&](int lds_bytes, int *blocks) -> tm_status {
    if (fn) {
      int32_t b = 0;
      if (fn(0, nullptr, 0, lds_bytes, sl.stream, &b, user) != 0)
        return fail(TM_ERR_HIP, "custom kernel occupancy query failed");
      *blocks = b;
      return TM_OK;
    }
    return dispatch<true>(sl, nmap, enumerate, instr, wpe, multi, lds_bytes, blocks, nullptr, 0);
  };
  int lds = lds_of(sb_flush);
  int blocks = 0;
  tm_status st;
  if ((st = blocks_of(lds, &blocks)) != TM_OK) return st;
  if (sb_auto && sb_flush > 0) {
    int b0 = 0;
    if ((st = blocks_of(lds_of(0), &b0)) != TM_OK) return st;
    const int cap = (tune && tune->waves_per_cu > 0) ? tune->waves_per_cu : 32;
    if (std::min(b0, cap) > std::min(blocks, cap)) {
      // the buffer would cost resident wavefronts: measured to cost more than it saves
      sb_flush = 0;
      lds = lds_of(0);
      blocks = b0;
    }
  }
  int wpc = (tune && tune->waves_per_cu > 0) ? tune->waves_per_cu : 32;
  if (blocks > 0) wpc = std::min(wpc, blocks);
  if (wpc < 1) wpc = 1;
  int grid = std::min(g->num_cus * wpc, s.wstats_cap);

  long long total_blocks = 0, max_blocks = 0;
  for (int i = 0; i < nq; i++) {
    const tm_query_dev &qd = s.h_queries[q0 + i];
    int *qt = &s.h_qtab[4 * (size_t)(q0 + i)];
    int rb, re;
    range_of(i, &rb, &re);
    qt[0] = qd.delta;
    qt[1] = rb;
    qt[2] = re;
    // start rule of edge 1 and, outside the instrumented kernel (which
    // follows the reference's search tree), the root's second rule
    qt[3] = qd.m.tmode[1] | (instr == 1 ? 0 : qd.m.xmask[0] << TM_QT_XRULE_SHIFT);
    s.h_prefix[q0 + i] = total_blocks;
    if (re > rb) {
      const long long nb = (re - (rb & ~3) + 255) / 256;
      total_blocks += nb;
      max_blocks = std::max(max_blocks, nb);
    }
  }
  s.h_prefix[q0 + nq] = total_blocks;
  if (total_blocks < grid) grid = (int)std::max<long long>(1, total_blocks);
  // claims of about 1/8 of one query's roots per wavefront: as fine-grained
  // as a launch of that query alone, whatever the number of queries
  int claim_blocks;
  if (tune && tune->claim_blocks > 0) {
    claim_blocks = tune->claim_blocks;
  } else {
    const long long c = (max_blocks + (long long)grid * 8 - 1) / ((long long)grid * 8);
    claim_blocks = (int)std::min<long long>(std::max<long long>(c, 4), 64);
  }

  HIPTRY(hipMemcpyAsync(s.d_queries + q0, s.h_queries.data() + q0, sizeof(tm_query_dev) * nq,
                        hipMemcpyHostToDevice, sl.stream));
  HIPTRY(hipMemcpyAsync(s.d_qtab + 4 * (size_t)q0, s.h_qtab.data() + 4 * (size_t)q0, sizeof(int) * 4 * nq,
                        hipMemcpyHostToDevice, sl.stream));
  HIPTRY(hipMemcpyAsync(s.d_prefix + q0, s.h_prefix.data() + q0, sizeof(long long) * (nq + 1),
                        hipMemcpyHostToDevice, sl.stream));

  tm_kernel_args a;
  memset(&a, 0, sizeof(a));
  a.g = g->view;
  a.gv = g->d_view;
  a.edgeFeature = g->edge_features;
  a.vertexFeature = g->vertex_features;
  a.queries = s.d_queries + q0;
  a.qtab = s.d_qtab + 4 * (size_t)q0;
  a.block_prefix = s.d_prefix + q0;
  a.nq = nq;
  a.total_blocks = total_blocks;
  a.claim_blocks = claim_blocks;
  a.frames = frames;
  a.sb_flush = sb_flush;
  a.sb_inline = sb_inline;
  a.scan_max = (tune && tune->scan_max > 0) ? tune->scan_max : 16;
  a.limit_ticks = 0;
  if (tune && tune->time_limit_ms > 0) a.limit_ticks = (long long)tune->time_limit_ms * (g->wall_rate_khz);
  a.sched = s.d_sched;
  a.wstats = s.d_wstats;
  a.pool = s.d_pool;
  a.pool_done = s.d_pool_done;
  a.pool_cap = s.pool_cap;
  a.pool_shift = s.pool_shift;
  a.steal_every = (tune && tune->steal_every > 0) ? tune->steal_every : 16;
  a.steal_max = (tune && tune->steal_max > 0) ? tune->steal_max : 4;
  a.idle_cap = std::max(64, g->num_cus / 2);
  a.steal_min = (tune && tune->steal_min > 0) ? tune->steal_min : 64;
  {
    const long long us = (tune && tune->steal_after_us > 0) ? tune->steal_after_us : 1000;
    a.steal_after_ticks = us * g->wall_rate_khz / 1000;
  }
  // two-edge tail: frames of edge M-2 with at least this many window candidates
  a.t2_min = (tune && tune->tail_swap < 0) ? 0x7fffffff : ((tune && tune->tail_swap > 0) ? tune->tail_swap : 32);
  sl.last_claim = claim_blocks;
  sl.last_grid = grid;
  sl.last_lds = lds;
  sl.last_wpe = fn ? 0 : wpe;
  sl.last_sweep = (int)(pl->sweep.size() + pl->pair.size());
  pl->a = a;
  pl->d_counts = s.d_counts + q0;
  pl->d_stops = s.d_stops + q0;
  pl->grid = grid;
  pl->lds = lds;
  pl->nmap = nmap;
  pl->wpe = wpe;
  pl->q0 = q0;
  pl->nq = nq;
  pl->enumerate = enumerate;
  pl->instr = instr;
  pl->multi = multi;
  pl->fn = fn;
  pl->user = user;
  return TM_OK;
}

tm_status count_launch(tm_graph *g, Slot &sl, hipStream_t stream, int n) {
  (void)g;
  Scratch &s = sl.s;
  if (sl.gen + n >= (1 << TM_GEN_BITS)) {
    // generations wrap: clear the ring so no stale tag can match
    const size_t nw = (size_t)TM_POOL_FIELDS * s.pool_cap;
    hipLaunchKernelGGL(clear_pool_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, stream, s.d_pool, nw,
                       s.d_pool_done, (size_t)s.pool_cap, s.d_sched);
    HIPTRY(hipGetLastError());
    sl.gen = 0;
  }
  sl.gen += n;
  return TM_OK;
}

tm_status enqueue(tm_graph *g, Slot &sl, const Plan &pl, bool timing, bool count) {
  tm_status st;
  // the generation count belongs to the graph's slot (sl may be a view of it
  // with other staging arrays); a captured launch is counted at each replay
  if (count && (st = count_launch(g, sl.own_ring ? sl : g->slot[sl.index], sl.stream, 1)) != TM_OK) return st;
  hipLaunchKernelGGL(reset_kernel, dim3((pl.nq + 255) / 256), dim3(256), 0, sl.stream, pl.a.sched, pl.d_counts,
                     pl.d_stops, pl.nq, pl.grid);
  HIPTRY(hipGetLastError());
  if (timing) HIPTRY(hipEventRecord(sl.ev0, sl.stream));
  if (pl.fn) {
    if (pl.fn(1, &pl.a, pl.grid, pl.lds, sl.stream, nullptr, pl.user) != 0)
      return fail(TM_ERR_HIP, "custom kernel launch failed");
  } else {
    int blocks = 0;
    if ((st = dispatch<false>(sl, pl.nmap, pl.enumerate, pl.instr, pl.wpe, pl.multi, pl.lds, &blocks, &pl.a,
                              pl.grid)) != TM_OK)
      return st;
  }
  // the window sweep after the mining kernel on the same stream: it adds
  // its queries' counts to the slots the reset kernel cleared
  for (size_t i = 0; i < pl.sweep.size(); i += TM_SWEEP_MAXQ) {
    SweepBatch b;
    b.n = (int)std::min<size_t>(TM_SWEEP_MAXQ, pl.sweep.size() - i);
    for (int k = 0; k < b.n; k++) b.item[k] = pl.sweep[i + k];
    if ((st = sweep_enqueue(g->view, pl.a.queries, b, sl.stream)) != TM_OK) return st;
  }
  for (size_t i = 0; i < pl.pair.size(); i += TM_SWEEP_MAXQ) {
    PairBatch b;
    b.n = (int)std::min<size_t>(TM_SWEEP_MAXQ, pl.pair.size() - i);
    for (int k = 0; k < b.n; k++) b.item[k] = pl.pair[i + k];
    if ((st = pair_enqueue(g->view, pl.a.queries, b, sl.stream)) != TM_OK) return st;
  }
  if (timing) HIPTRY(hipEventRecord(sl.ev1, sl.stream));
  hipLaunchKernelGGL(stats_kernel, dim3(1), dim3(256), 0, sl.stream, pl.a.wstats, pl.grid, pl.a.sched);
  HIPTRY(hipGetLastError());
  return TM_OK;
}

tm_status launch(tm_graph *g, Slot &sl, int q0, int nq, bool enumerate, const tm_tuning *tune, int instr,
                 bool multi, tm_launch_fn fn, void *user) {
  Plan pl;
  tm_status st = plan(g, sl, q0, nq, enumerate, tune, instr, multi, fn, user, &pl);
  if (st != TM_OK) return st;
  return enqueue(g, sl, pl, true, true);
}

// Create the slot's stream and reserve its staging space for nq queries
// (grows, never shrinks).
tm_status reserve(tm_graph *g, Slot &sl, int nq) {
  Scratch &s = sl.s;
  if (!sl.stream) {
    HIPTRY(hipStreamCreate(&sl.stream));
    HIPTRY(hipEventCreate(&sl.ev0));
    HIPTRY(hipEventCreate(&sl.ev1));
  }
  if (nq <= s.cap_q) return TM_OK;
  int cap = std::max(nq, 256);
  void *p;
  tm_status st;
  if ((st = dev_alloc(g, &p, sizeof(tm_query_dev) * cap)) != TM_OK) return st;
  s.d_queries = (tm_query_dev *)p;
  if ((st = dev_alloc(g, &p, sizeof(int) * 4 * cap)) != TM_OK) return st;
  s.d_qtab = (int *)p;
  if ((st = dev_alloc(g, &p, sizeof(long long) * (cap + 1))) != TM_OK) return st;
  s.d_prefix = (long long *)p;
  if ((st = dev_alloc(g, &p, sizeof(unsigned long long) * cap)) != TM_OK) return st;
  s.d_counts = (unsigned long long *)p;
  if ((st = dev_alloc(g, &p, sizeof(int) * cap)) != TM_OK) return st;
  s.d_stops = (int *)p;
  if (!s.d_sched) {
    if ((st = dev_alloc(g, &p, sizeof(tm_sched))) != TM_OK) return st;
    s.d_sched = (tm_sched *)p;
    HIPTRY(hipMemsetAsync(s.d_sched, 0, sizeof(tm_sched), sl.stream));  // generation 0: no launch yet
    s.pool_shift = 18;
    s.pool_cap = 1 << s.pool_shift;
    const size_t pool_bytes = sizeof(unsigned long long) * TM_POOL_FIELDS * (size_t)s.pool_cap;
    if ((st = dev_alloc(g, &p, pool_bytes)) != TM_OK) return st;
    s.d_pool = (unsigned long long *)p;
    if ((st = dev_alloc(g, &p, sizeof(unsigned) * (size_t)s.pool_cap)) != TM_OK) return st;
    s.d_pool_done = (unsigned *)p;
    HIPTRY(hipMemsetAsync(s.d_pool_done, 0, sizeof(unsigned) * (size_t)s.pool_cap, sl.stream));
    s.wstats_cap = g->num_cus * 32;  // the most resident one-wave workgroups
    if ((st = dev_alloc(g, &p, sizeof(unsigned long long) * TM_NSTATS * (size_t)s.wstats_cap)) != TM_OK) return st;
    s.d_wstats = (unsigned long long *)p;
    HIPTRY(hipMemsetAsync(s.d_pool, 0, pool_bytes, sl.stream));  // generation 0 is never used
    HIPTRY(hipStreamSynchronize(sl.stream));
  }
  s.cap_q = cap;
  s.h_queries.resize(cap);
  s.h_qtab.resize(4 * (size_t)cap);
  s.h_prefix.resize(cap + 1);
  return TM_OK;
}

// Stage queries idx[0..nq) (indexes into the caller's arrays) at staging
// positions [q0, q0 + nq) of the slot.
tm_status stage_queries(tm_graph *g, Slot &sl, int q0, int nq, const int *idx, const int32_t *num_edges,
                               const int32_t *msrc, const int32_t *mdst, int stride, const int32_t *deltas,
                               const int32_t *begs, const int32_t *ends) {
  tm_status st;
  if ((st = reserve(g, sl, q0 + nq)) != TM_OK) return st;
  Scratch &s = sl.s;
  for (int i = 0; i < nq; i++) {
    const int q = idx ? idx[i] : i;
    tm_query_dev &qd = s.h_queries[q0 + i];
    memset(&qd, 0, sizeof(qd));
    if ((st = build_motif(num_edges[q], msrc + (size_t)q * stride, mdst + (size_t)q * stride, &qd.m)) != TM_OK)
      return st;
    if (deltas[q] < 0) return fail(TM_ERR_ARG, "delta must be >= 0");
    qd.delta = deltas[q];
    int beg = begs ? begs[q] : 0;
    int end = ends ? ends[q] : -1;
    if (end < 0) end = g->E - qd.m.M + 1;
    if (beg < 0) beg = 0;
    if (end > g->E) end = g->E;
    if (end < beg) end = beg;
    qd.root_beg = beg;
    qd.root_end = end;
    qd.count = s.d_counts + q0 + i;
    qd.stop = s.d_stops + q0 + i;
  }
  return TM_OK;
}

// Check a completed launch of the slot: its scheduler state hs (copied back)
// gives the launch statistics and error bits.
tm_status collect(tm_graph *g, Slot &sl, const tm_sched &hs) {
  // the statistics land in the graph's slot of that index (sl may be a view
  // of it with other device buffers)
  Slot &dst = g->slot[sl.index];
  memcpy(dst.last_stats, hs.stats, sizeof(hs.stats));
  dst.last_stats[TM_STAT_SWEEP] = (unsigned long long)sl.last_sweep;
  dst.last_stats[TM_STAT_GRID] = (unsigned long long)sl.last_grid;
  dst.last_stats[TM_STAT_CLAIM] = (unsigned long long)sl.last_claim;
  dst.last_stats[TM_STAT_LDS] = (unsigned long long)sl.last_lds;
  dst.last_stats[TM_STAT_WPE] = (unsigned long long)sl.last_wpe;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, sl.ev0, sl.ev1) == hipSuccess) dst.last_kernel_ms = ms;
  g->last_slot = sl.index;
  if (hs.error & ~4) return fail(TM_ERR_INTERNAL, "kernel invariant violated (error bits " + std::to_string(hs.error) + ")");
  if (hs.error & 4) return fail(TM_ERR_TIMEOUT, "query exceeded its time limit; counts discarded");
  return TM_OK;
}

// Wait for the slot's launch of staged queries [q0, q0 + nq); copy their
// counts out and check the launch's error bits.
tm_status finish(tm_graph *g, Slot &sl, int q0, int nq, uint64_t *counts) {
  Scratch &s = sl.s;
  tm_sched hs;
  HIPTRY(hipMemcpyAsync(&hs, s.d_sched, sizeof(hs), hipMemcpyDeviceToHost, sl.stream));
  HIPTRY(hipMemcpyAsync(counts, s.d_counts + q0, sizeof(uint64_t) * nq, hipMemcpyDeviceToHost, sl.stream));
  HIPTRY(hipStreamSynchronize(sl.stream));
  return collect(g, sl, hs);
}

// Counts of one motif over consecutive root ranges [cuts[i], cuts[i + 1])
// (i < n <= TM_LAUNCH_MAXQ), in one launch of the batch kernel; with a
// caller's kernel (fn), one launch of it per range.
static tm_status count_ranges(tm_graph *g, Slot &sl, int M, const int32_t *src, const int32_t *dst, int32_t delta,
                              const int32_t *cuts, int n, const tm_tuning *tune, uint64_t *counts,
                              tm_launch_fn fn = nullptr, void *user = nullptr) {
  if (fn) {
    for (int i = 0; i < n; i++) {
      int32_t b = cuts[i], e = cuts[i + 1];
      tm_status st = stage_queries(g, sl, 0, 1, nullptr, &M, src, dst, M, &delta, &b, &e);
      if (st == TM_OK) st = launch(g, sl, 0, 1, false, tune, false, false, fn, user);
      if (st == TM_OK) st = finish(g, sl, 0, 1, counts + i);
      if (st != TM_OK) return st;
    }
    return TM_OK;
  }
  std::vector<int32_t> ne(n, M), dl(n, delta), bg(cuts, cuts + n), en(cuts + 1, cuts + n + 1);
  // query i: the same motif (stride 0), roots [cuts[i], cuts[i + 1])
  tm_status st = stage_queries(g, sl, 0, n, nullptr, ne.data(), src, dst, 0, dl.data(), bg.data(), en.data());
  if (st == TM_OK) st = launch(g, sl, 0, n, false, tune, false, true);
  if (st == TM_OK) st = finish(g, sl, 0, n, counts);
  return st;
}

// Every row of the roots [b, e) (fewer than `room` of them) into host rows.
static tm_status enum_all(tm_graph *g, Slot &sl, int M, const int32_t *src, const int32_t *dst, int32_t delta,
                          int32_t b, int32_t e, int64_t room, std::vector<int32_t> *rows, tm_launch_fn fn,
                          void *user) {
  int32_t *d_rows = nullptr;
  HIPTRY(hipMalloc(&d_rows, sizeof(int32_t) * (size_t)std::max<int64_t>(room, 1) * M));
  tm_status st = stage_queries(g, sl, 0, 1, nullptr, &M, src, dst, M, &delta, &b, &e);
  uint64_t c = 0;
  if (st == TM_OK) {
    sl.s.h_queries[0].rows = d_rows;
    sl.s.h_queries[0].cap = room;
    st = launch(g, sl, 0, 1, true, nullptr, false, false, fn, user);
  }
  if (st == TM_OK) st = finish(g, sl, 0, 1, &c);
  if (st == TM_OK && (int64_t)c >= room) st = fail(TM_ERR_INTERNAL, "enumeration found more rows than counted");
  if (st == TM_OK) {
    rows->resize((size_t)c * M);
    if (c) {
      hipError_t he = hipMemcpy(rows->data(), d_rows, sizeof(int32_t) * rows->size(), hipMemcpyDeviceToHost);
      if (he != hipSuccess) st = fail(TM_ERR_HIP, hipGetErrorString(he));
    }
  }
  (void)hipFree(d_rows);
  return st;
}

tm_status enum_cut(tm_graph *g, Slot &sl, int M, const int32_t *src, const int32_t *dst, int32_t delta, int32_t beg,
                   int32_t end, int64_t cap, const tm_tuning *tune, int32_t *rows, uint64_t *count,
                   int64_t *rows_written, tm_launch_fn fn, void *user) {
  // the cut is found with the counting kernel's own launch parameters (its
  // class's), and the caller's choice of counting path
  *rows_written = 0;
  if (cap == 0) { *count = 0; return TM_OK; }
  tm_tuning cut_tune;
  memset(&cut_tune, 0, sizeof(cut_tune));
  if (const tm_tuning *cls = class_tuning_of(g, M, src, dst)) cut_tune = *cls;
  if (tune && tune->sweep) cut_tune.sweep = tune->sweep;
  const tm_tuning *ct = &cut_tune;
  // Find the root r* at which the cumulative count over [beg, r] first
  // reaches the cap: ranges of growing width after beg, then the range that
  // holds the crossing cut into TM_CUT_PARTS parts, down to one root.
  const int P = 64;
  uint64_t before = 0;       // matches of the roots [beg, lo)
  int32_t lo = beg;
  long long width = 1 << 14;
  int32_t rstar = -1;
  uint64_t c_star = 0;       // matches of root r*
  std::vector<int32_t> cuts(P + 1);
  std::vector<uint64_t> cnt(P);
  tm_status st;
  for (;;) {
    const int32_t hi = (int32_t)std::min<long long>(end, (long long)lo + width);
    if (hi <= lo) return fail(TM_ERR_INTERNAL, "the cap was reached but the counts never reach it");
    const long long span = hi - lo;
    int n = (int)std::min<long long>(P, span);
    for (int i = 0; i <= n; i++) cuts[i] = lo + (int32_t)(span * i / n);
    if ((st = count_ranges(g, sl, M, src, dst, delta, cuts.data(), n, fn ? tune : ct, cnt.data(), fn, user)) != TM_OK)
      return st;
    int k = -1;
    uint64_t acc = before;
    for (int i = 0; i < n; i++) {
      if (acc + cnt[i] >= (uint64_t)cap) { k = i; break; }
      acc += cnt[i];
    }
    if (k < 0) {  // not within [lo, hi): move on, twice as wide
      before = acc;
      lo = hi;
      width *= 2;
      continue;
    }
    before = acc;
    lo = cuts[k];
    if (cuts[k + 1] - cuts[k] == 1) {
      rstar = cuts[k];
      c_star = cnt[k];
      break;
    }
    width = cuts[k + 1] - cuts[k];  // refine inside the part that holds the crossing
  }
  // rows of [beg, r*): all of them (fewer than cap); rows of r*: the
  // k = cap - before smallest by their Euid tuples
  std::vector<int32_t> head, tail;
  if ((st = enum_all(g, sl, M, src, dst, delta, beg, rstar, (int64_t)before + 1, &head, fn, user)) != TM_OK) return st;
  if ((int64_t)head.size() != (int64_t)before * M) return fail(TM_ERR_INTERNAL, "enumeration and count disagree");
  const int64_t kMaxRootRows = (int64_t)1 << 27;
  if ((int64_t)c_star > kMaxRootRows)
    return fail(TM_ERR_NOMEM, "the root at the cap holds more than 2^27 matches; its rows cannot be ordered");
  if ((st = enum_all(g, sl, M, src, dst, delta, rstar, rstar + 1, (int64_t)c_star + 1, &tail, fn, user)) != TM_OK)
    return st;
  if ((int64_t)tail.size() != (int64_t)c_star * M) return fail(TM_ERR_INTERNAL, "enumeration and count disagree");
  const int64_t k = cap - (int64_t)before;
  std::vector<int64_t> idx((size_t)c_star);
  for (int64_t i = 0; i < (int64_t)c_star; i++) idx[(size_t)i] = i;
  auto less = [&](int64_t a, int64_t b) {
    return std::lexicographical_compare(tail.begin() + a * M, tail.begin() + (a + 1) * M, tail.begin() + b * M,
                                        tail.begin() + (b + 1) * M);
  };
  if (k < (int64_t)c_star) std::nth_element(idx.begin(), idx.begin() + k, idx.end(), less);
  std::sort(idx.begin(), idx.begin() + k, less);
  memcpy(rows, head.data(), sizeof(int32_t) * head.size());
  for (int64_t i = 0; i < k; i++)
    memcpy(rows + (size_t)(before + i) * M, tail.data() + (size_t)idx[(size_t)i] * M, sizeof(int32_t) * M);
  *rows_written = cap;
  *count = before + c_star;
  return TM_OK;
}

}  // namespace tmh

using namespace tmh;

extern "C" {

int32_t tm_abi_version(void) { return TM_ABI_VERSION; }

const char *tm_last_error(void) { return g_last_error.c_str(); }

int32_t tm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

tm_status tm_graph_upload(int32_t device, const tm_host_graph *hg, tm_graph **out) {
  if (!hg || !out) return fail(TM_ERR_ARG, "null argument");
  *out = nullptr;
  const int E = hg->num_edges, V = hg->num_vertices;
  if (E < 1 || V < 1) return fail(TM_ERR_ARG, "graph needs at least one edge and vertex");
  if (!hg->Eg || !hg->inEdgesV || !hg->outEdgesV || !hg->inEdgesR || !hg->outEdgesR || !hg->Euid)
    return fail(TM_ERR_ARG, "null graph array");
  // Host-side validation: every index the kernels follow must be in range.
  const tm_edge *Eg = (const tm_edge *)hg->Eg;
  bool self_loops = false;
  for (int p = 0; p < E; p++) {
    if (Eg[p].u < 0 || Eg[p].u >= V || Eg[p].v < 0 || Eg[p].v >= V)
      return fail(TM_ERR_ARG, "edge endpoint out of [0, V)");
    if (p && Eg[p].t < Eg[p - 1].t) return fail(TM_ERR_ARG, "edges are not sorted by time");
    self_loops |= Eg[p].u == Eg[p].v;
  }
  for (int k = 0; k < 2; k++) {
    const int32_t *R = k ? hg->inEdgesR : hg->outEdgesR;
    const int32_t *P = k ? hg->inEdgesV : hg->outEdgesV;
    if (R[0] != 0 || R[V] != E) return fail(TM_ERR_ARG, "CSR offsets must span [0, E]");
    for (int x = 0; x < V; x++) {
      if (R[x + 1] < R[x]) return fail(TM_ERR_ARG, "CSR offsets not monotone");
      for (int s = R[x]; s < R[x + 1]; s++) {
        int p = P[s];
        if (p < 0 || p >= E) return fail(TM_ERR_ARG, "CSR position out of range");
        if ((k ? Eg[p].v : Eg[p].u) != x) return fail(TM_ERR_ARG, "CSR position on the wrong vertex");
        if (s > R[x] && P[s - 1] >= p) return fail(TM_ERR_ARG, "CSR positions not ascending");
      }
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(TM_ERR_NODEVICE, "no GPU visible");
  if (device < 0 || device >= ndev) return fail(TM_ERR_ARG, "device out of range");
  HIPTRY(hipSetDevice(device));

  tm_graph *g = new tm_graph();
  g->device = device;
  g->E = E;
  g->V = V;
  // degree products per node (root list x window list) and the time span:
  // the window sweep's break-even estimate
  for (int x = 0; x < V; x++) {
    const double o = hg->outEdgesR[x + 1] - hg->outEdgesR[x], i = hg->inEdgesR[x + 1] - hg->inEdgesR[x];
    g->deg2[0] += o * o;
    g->deg2[1] += i * i;
    g->deg2[2] += o * i;
  }
  g->tspan = (double)Eg[E - 1].t - (double)Eg[0].t;
  hipDeviceProp_t prop;
  hipError_t he = hipGetDeviceProperties(&prop, device);
  if (he != hipSuccess) { delete g; return fail(TM_ERR_HIP, hipGetErrorString(he)); }
  g->num_cus = prop.multiProcessorCount;
  tm_status st0;
  {
    int rate = 0;
    if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, device) == hipSuccess && rate > 0)
      g->wall_rate_khz = rate;
  }
  for (int k = 0; k < TM_SLOTS; k++) g->slot[k].index = k;
  for (Slot &sl : g->slot) {  // every slot's stream, scheduler state and pool, up front
    if ((st0 = reserve(g, sl, 1)) != TM_OK) {
      tm_graph_free(g);
      return st0;
    }
  }

  auto cleanup = [&](tm_status s) {
    tm_graph_free(g);
    return s;
  };
  const size_t e4 = sizeof(int) * (size_t)E, v4 = sizeof(int) * (size_t)(V + 1);
  void *p;
  tm_graph_view &vw = g->view;
  vw.E = E;
  vw.V = V;
  vw.no_self_loops = self_loops ? 0 : 1;
  tm_status st;
#define ALLOC(field, bytes)                                            \
  if ((st = dev_alloc(g, &p, (bytes))) != TM_OK) return cleanup(st); \
  field = (decltype(field))p;
  ALLOC(vw.Eg, 3 * e4);
  ALLOC(vw.T, e4);
  ALLOC(vw.outR, v4);
  ALLOC(vw.inR, v4);
  ALLOC(vw.outPos, e4);
  ALLOC(vw.inPos, e4);
  ALLOC(vw.outNbr, e4);
  ALLOC(vw.inNbr, e4);
  ALLOC(vw.outPairNbr, e4);
  ALLOC(vw.outPairPos, e4);
  ALLOC(vw.inPairNbr, e4);
  ALLOC(vw.inPairPos, e4);
  ALLOC(vw.Euid, e4);
  ALLOC(vw.outSelf, e4);
  ALLOC(vw.inSelf, e4);
  ALLOC(vw.outNext, e4);
  ALLOC(vw.inNext, e4);
  ALLOC(vw.outPairSelf, e4);
  for (int k = 0; k < TM_NTFIRST; k++) {
    ALLOC(vw.tfirst[k], e4);
  }
  ALLOC(vw.stf, e4 * 4 * TM_NTFIRST);
  ALLOC(vw.sprev, e4 * 2);
  vw.epad = (int)(((long long)E + 255) / 256 * 256 + 256);
  ALLOC(vw.rgap, sizeof(unsigned) * (size_t)vw.epad * TM_NTFIRST);
  const int tmin = Eg[0].t, tmax = Eg[E - 1].t;
  const long long span = (long long)tmax - tmin + 1;
  int bshift = ceil_log2(std::max(1ll, (span * 16) / std::max(1, E)));
  if (bshift > 30) bshift = 30;
  const long long nb = (span + (1ll << bshift) - 1) >> bshift;
  vw.tmin = tmin;
  vw.bshift = bshift;
  vw.nbuckets = nb;
  ALLOC(vw.bucket, sizeof(int) * (size_t)(nb + 1));
#undef ALLOC
  hipStream_t s = g->slot[0].stream;
#define H2D(dst, src, bytes) \
  if ((he = hipMemcpyAsync((void *)(dst), (src), (bytes), hipMemcpyHostToDevice, s)) != hipSuccess) \
    return cleanup(fail(TM_ERR_HIP, hipGetErrorString(he)));
  H2D(vw.Eg, hg->Eg, 3 * e4);
  H2D(vw.outR, hg->outEdgesR, v4);
  H2D(vw.inR, hg->inEdgesR, v4);
  H2D(vw.outPos, hg->outEdgesV, e4);
  H2D(vw.inPos, hg->inEdgesV, e4);
  H2D(vw.Euid, hg->Euid, e4);
#undef H2D
  // temporaries for the pair index
  unsigned long long *okey = nullptr, *ikey = nullptr, *skey = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  auto free_tmp = [&]() {
    if (okey) (void)hipFree(okey);
    if (ikey) (void)hipFree(ikey);
    if (skey) (void)hipFree(skey);
    if (tmp) (void)hipFree(tmp);
  };
#define TMPCHK(expr) \
  if ((he = (expr)) != hipSuccess) { free_tmp(); return cleanup(fail(he == hipErrorOutOfMemory ? TM_ERR_NOMEM : TM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(he))); }
  TMPCHK(hipMalloc(&okey, sizeof(unsigned long long) * (size_t)E));
  TMPCHK(hipMalloc(&ikey, sizeof(unsigned long long) * (size_t)E));
  TMPCHK(hipMalloc(&skey, sizeof(unsigned long long) * (size_t)E));
  hipLaunchKernelGGL(aux_kernel, dim3((E + 255) / 256), dim3(256), 0, s, vw.Eg, vw.outPos, vw.inPos, E,
                     (int *)vw.T, (int *)vw.outNbr, (int *)vw.inNbr, okey, ikey);
  TMPCHK(hipGetLastError());
  const int end_bit = std::min(64, 32 + std::max(1, ceil_log2((long long)V + 1)));
  TMPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, okey, skey, vw.outPos, (int *)vw.outPairPos,
                                            E, 0, end_bit, s));
  TMPCHK(hipMalloc(&tmp, tmp_bytes));
  TMPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, okey, skey, vw.outPos, (int *)vw.outPairPos, E,
                                            0, end_bit, s));
  hipLaunchKernelGGL(low_word_kernel, dim3((E + 255) / 256), dim3(256), 0, s, skey, (int *)vw.outPairNbr, E);
  TMPCHK(hipGetLastError());
  TMPCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, ikey, skey, vw.inPos, (int *)vw.inPairPos, E, 0,
                                            end_bit, s));
  hipLaunchKernelGGL(low_word_kernel, dim3((E + 255) / 256), dim3(256), 0, s, skey, (int *)vw.inPairNbr, E);
  TMPCHK(hipGetLastError());
  {
    const dim3 gr((E + 255) / 256), bl(256);
    hipLaunchKernelGGL(self_kernel, gr, bl, 0, s, vw.outPos, E, (int *)vw.outSelf);
    hipLaunchKernelGGL(self_kernel, gr, bl, 0, s, vw.inPos, E, (int *)vw.inSelf);
    hipLaunchKernelGGL(self_kernel, gr, bl, 0, s, vw.outPairPos, E, (int *)vw.outPairSelf);
    hipLaunchKernelGGL(next_kernel, gr, bl, 0, s, vw.Eg, E, vw.outR, vw.outPos, vw.inR, vw.inPos,
                       (int *)vw.outNext, (int *)vw.inNext);
    TMPCHK(hipGetLastError());
    hipLaunchKernelGGL(tfirst_kernel, gr, bl, 0, s, vw.Eg, vw.T, E, vw.outR, vw.outPos, vw.inR, vw.inPos,
                       vw.outSelf, vw.inSelf, vw.outNext, vw.inNext, vw.outPairNbr, vw.outPairPos,
                       vw.outPairSelf, (int *)vw.tfirst[0], (int *)vw.tfirst[1], (int *)vw.tfirst[2],
                       (int *)vw.tfirst[3], (int *)vw.tfirst[4], (int *)vw.tfirst[5]);
    TMPCHK(hipGetLastError());
    {
      const int *lists[4] = {vw.outPos, vw.inPos, vw.outPairPos, vw.inPairPos};  // TM_L_* order
      for (int l = 0; l < 4; l++) {
        int *o = (int *)vw.stf + (size_t)l * TM_NTFIRST * E;
        hipLaunchKernelGGL(stf_kernel, gr, bl, 0, s, lists[l], E, vw.tfirst[0], vw.tfirst[1], vw.tfirst[2],
                           vw.tfirst[3], vw.tfirst[4], vw.tfirst[5], o, o + (size_t)E, o + 2 * (size_t)E,
                           o + 3 * (size_t)E, o + 4 * (size_t)E, o + 5 * (size_t)E);
      }
      TMPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(sprev_kernel, gr, bl, 0, s, vw.Eg, vw.outPos, vw.inPos, vw.outR, vw.outPairNbr, vw.outPairPos,
                       vw.outPairSelf, E, (int *)vw.sprev);
    TMPCHK(hipGetLastError());
    TMPCHK(hipMemsetAsync((void *)vw.rgap, 0xff, sizeof(unsigned) * (size_t)vw.epad * TM_NTFIRST, s));
    {
      unsigned *rg = (unsigned *)vw.rgap;
      const size_t ep = (size_t)vw.epad;
      hipLaunchKernelGGL(rgap_kernel, gr, bl, 0, s, vw.Eg, E, vw.tfirst[0], vw.tfirst[1], vw.tfirst[2],
                         vw.tfirst[3], vw.tfirst[4], vw.tfirst[5], rg, rg + ep, rg + 2 * ep, rg + 3 * ep,
                         rg + 4 * ep, rg + 5 * ep);
    }
    TMPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(bucket_kernel, dim3((unsigned)((nb + 1 + 255) / 256)), dim3(256), 0, s, vw.T, E, tmin,
                     bshift, nb, (int *)vw.bucket);
  TMPCHK(hipGetLastError());
  TMPCHK(hipStreamSynchronize(s));
#undef TMPCHK
  free_tmp();
  // the view itself in device memory, for the kernels to read the graph through
  if ((st = dev_alloc(g, &p, sizeof(tm_graph_view))) != TM_OK) return cleanup(st);
  g->d_view = (tm_graph_view *)p;
  if ((he = hipMemcpy(g->d_view, &g->view, sizeof(tm_graph_view), hipMemcpyHostToDevice)) != hipSuccess)
    return cleanup(fail(TM_ERR_HIP, std::string("view upload: ") + hipGetErrorString(he)));
  *out = g;
  return TM_OK;
}

tm_status tm_graph_free(tm_graph *g) {
  if (!g) return TM_OK;
  (void)hipSetDevice(g->device);
  for (Slot &sl : g->slot)
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  for (void *p : g->allocs) (void)hipFree(p);
  if (g->edge_features) (void)hipFree(g->edge_features);
  if (g->vertex_features) (void)hipFree(g->vertex_features);
  for (Slot &sl : g->slot) {
    if (sl.ev0) (void)hipEventDestroy(sl.ev0);
    if (sl.ev1) (void)hipEventDestroy(sl.ev1);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  delete g;
  return TM_OK;
}

tm_status tm_graph_arrays(const tm_graph *g, tm_device_arrays *out) {
  if (!g || !out) return fail(TM_ERR_ARG, "null argument");
  out->num_edges = g->E;
  out->num_vertices = g->V;
  out->Eg = g->view.Eg;
  out->inEdgesV = g->view.inPos;
  out->outEdgesV = g->view.outPos;
  out->inEdgesR = g->view.inR;
  out->outEdgesR = g->view.outR;
  out->Euid = g->view.Euid;
  out->edgeFeatures = g->edge_features;
  out->verticesFeatures = g->vertex_features;
  out->num_vertex_features = g->num_vertex_features;
  return TM_OK;
}

tm_status tm_graph_set_features(tm_graph *g, const int32_t *edge_features, int32_t num_edge_features,
                                const int32_t *vertex_features, int32_t num_vertex_features) {
  if (!g || num_edge_features < 0 || num_vertex_features < 0 || (num_edge_features && !edge_features) ||
      (num_vertex_features && !vertex_features))
    return fail(TM_ERR_ARG, "bad argument");
  if (num_edge_features && num_edge_features != g->E) return fail(TM_ERR_ARG, "one edge label per edge");
  std::lock_guard<std::mutex> lk(g->mtx);
  HIPTRY(hipSetDevice(g->device));
  auto replace = [&](int **dst, int *n, const int32_t *src, int32_t cnt) -> tm_status {
    if (*dst) {
      HIPTRY(hipFree(*dst));
      *dst = nullptr;
      *n = 0;
    }
    if (cnt == 0) return TM_OK;
    HIPTRY(hipMalloc((void **)dst, sizeof(int32_t) * (size_t)cnt));
    HIPTRY(hipMemcpy(*dst, src, sizeof(int32_t) * (size_t)cnt, hipMemcpyHostToDevice));
    *n = cnt;
    return TM_OK;
  };
  tm_status st = replace(&g->edge_features, &g->num_edge_features, edge_features, num_edge_features);
  if (st == TM_OK) st = replace(&g->vertex_features, &g->num_vertex_features, vertex_features, num_vertex_features);
  return st;
}

tm_status tm_device_copy(tm_graph *g, const void *host, int64_t bytes, void **device_ptr) {
  if (!g || !device_ptr || bytes < 0 || (bytes > 0 && !host)) return fail(TM_ERR_ARG, "bad argument");
  HIPTRY(hipSetDevice(g->device));
  void *p = nullptr;
  HIPTRY(hipMalloc(&p, (size_t)std::max<int64_t>(bytes, 4)));
  if (bytes > 0) {
    hipError_t e = hipMemcpy(p, host, (size_t)bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return fail(TM_ERR_HIP, std::string("tm_device_copy: ") + hipGetErrorString(e));
    }
  }
  *device_ptr = p;
  return TM_OK;
}

tm_status tm_device_free(tm_graph *g, void *device_ptr) {
  if (!g) return fail(TM_ERR_ARG, "null graph");
  if (!device_ptr) return TM_OK;
  HIPTRY(hipSetDevice(g->device));
  HIPTRY(hipFree(device_ptr));
  return TM_OK;
}

tm_status tm_last_stats(const tm_graph *g, uint64_t *stats, int32_t n) {
  if (!g || !stats || n < 0) return fail(TM_ERR_ARG, "bad argument");
  for (int i = 0; i < n && i < TM_NSTATS_HOST; i++) stats[i] = g->slot[g->last_slot].last_stats[i];
  return TM_OK;
}

tm_status tm_last_kernel_ms(const tm_graph *g, double *ms) {
  if (!g || !ms) return fail(TM_ERR_ARG, "bad argument");
  *ms = g->slot[g->last_slot].last_kernel_ms;
  return TM_OK;
}

tm_status tm_graph_info(const tm_graph *g, int64_t *device_bytes, int32_t *num_edges,
                        int32_t *num_vertices, int32_t *device) {
  if (!g) return fail(TM_ERR_ARG, "null graph");
  if (device_bytes) *device_bytes = g->bytes;
  if (num_edges) *num_edges = g->E;
  if (num_vertices) *num_vertices = g->V;
  if (device) *device = g->device;
  return TM_OK;
}

static tm_status count_impl(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                            int32_t beg, int32_t end, const tm_tuning *tune, uint64_t *count, tm_launch_fn fn,
                            void *user) {
  if (!g || !msrc || !mdst || !count) return fail(TM_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(g->mtx);
  HIPTRY(hipSetDevice(g->device));
  Slot &sl = g->slot[0];
  tm_status st = stage_queries(g, sl, 0, 1, nullptr, &num_edges, msrc, mdst, num_edges, &delta, &beg, &end);
  if (st != TM_OK) return st;
  if ((st = launch(g, sl, 0, 1, false, tune, false, false, fn, user)) != TM_OK) return st;
  return finish(g, sl, 0, 1, count);
}

static int shape_class(int M, int n);
static const tm_tuning *class_tuning(tm_graph *g, int cls, const tm_tuning *tune);
static int motif_class(int32_t num_edges, const int32_t *msrc, const int32_t *mdst);

tm_status tm_count(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                   int32_t beg, int32_t end, const tm_tuning *tune, uint64_t *count) {
  if (g && msrc && mdst && num_edges >= 2 && num_edges <= TM_MAX_MOTIF_EDGES)
    tune = class_tuning(g, motif_class(num_edges, msrc, mdst), tune);
  return count_impl(g, num_edges, msrc, mdst, delta, beg, end, tune, count, nullptr, nullptr);
}

tm_status tm_shape_class(int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t *cls) {
  if (!msrc || !mdst || !cls || num_edges < 2 || num_edges > TM_MAX_MOTIF_EDGES) return fail(TM_ERR_ARG, "bad argument");
  *cls = motif_class(num_edges, msrc, mdst);
  return TM_OK;
}

tm_status tm_set_class_tuning(tm_graph *g, int32_t cls, const tm_tuning *tune) {
  if (!g || cls < 0 || cls >= TM_NCLASSES) return fail(TM_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(g->mtx);
  if (tune) {
    g->class_tune[cls] = *tune;
    g->class_set[cls] = true;
  } else {
    g->class_set[cls] = false;
  }
  return TM_OK;
}

tm_status tm_count_custom(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                          int32_t beg, int32_t end, const tm_tuning *tune, tm_launch_fn launcher, void *user,
                          uint64_t *count) {
  if (!launcher) return fail(TM_ERR_ARG, "null launcher");
  return count_impl(g, num_edges, msrc, mdst, delta, beg, end, tune, count, launcher, user);
}

tm_status tm_count_work(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                        int32_t beg, int32_t end, const tm_tuning *tune, uint64_t *count, uint64_t *work) {
  if (!g || !msrc || !mdst || !count || !work) return fail(TM_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(g->mtx);
  HIPTRY(hipSetDevice(g->device));
  Slot &sl = g->slot[0];
  tm_status st = stage_queries(g, sl, 0, 1, nullptr, &num_edges, msrc, mdst, num_edges, &delta, &beg, &end);
  if (st != TM_OK) return st;
  if ((st = launch(g, sl, 0, 1, false, tune, true, false)) != TM_OK) return st;
  if ((st = finish(g, sl, 0, 1, count)) != TM_OK) return st;
  const tm_query_dev &q = sl.s.h_queries[0];
  work[0] = (uint64_t)(q.root_end - q.root_beg);
  work[1] = sl.last_stats[TM_STAT_REF_X1];
  work[2] = sl.last_stats[TM_STAT_REF_X2];
  work[3] = sl.last_stats[TM_STAT_REF_P];
  return TM_OK;
}

tm_status tm_count_bytes(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                         int32_t beg, int32_t end, const tm_tuning *tune, uint64_t *count, uint64_t *bytes) {
  if (!g || !msrc || !mdst || !count || !bytes) return fail(TM_ERR_ARG, "null argument");
  if (num_edges >= 2 && num_edges <= TM_MAX_MOTIF_EDGES)
    tune = class_tuning(g, motif_class(num_edges, msrc, mdst), tune);
  std::lock_guard<std::mutex> lk(g->mtx);
  HIPTRY(hipSetDevice(g->device));
  Slot &sl = g->slot[0];
  tm_status st = stage_queries(g, sl, 0, 1, nullptr, &num_edges, msrc, mdst, num_edges, &delta, &beg, &end);
  if (st != TM_OK) return st;
  if ((st = launch(g, sl, 0, 1, false, tune, 2, false)) != TM_OK) return st;
  if ((st = finish(g, sl, 0, 1, count)) != TM_OK) return st;
  *bytes = sl.last_stats[TM_STAT_REF_X1];
  return TM_OK;
}

// Launch shape class of a count query: frame-stack levels and node-map width
// (they set the LDS a wavefront needs, hence the occupancy of its launch).
static int shape_class(int M, int n) { return std::max(1, M - 2) * 2 + (n <= 4 ? 0 : 1); }

// The graph's tuning table: a call without launch parameters of its own uses
// the entry of its query's shape class, if one was set (tuning_for_class).
static const tm_tuning *class_tuning(tm_graph *g, int cls, const tm_tuning *tune) {
  return tuning_for_class(g, cls, tune);
}


static int motif_class(int32_t num_edges, const int32_t *msrc, const int32_t *mdst) {
  int n = 0;
  for (int k = 0; k < num_edges; k++) n = std::max(n, 1 + std::max(msrc[k], mdst[k]));
  return shape_class(num_edges, n);
}

tm_status tm_batch_count(tm_graph *g, int32_t nq, const int32_t *num_edges, const int32_t *msrc,
                         const int32_t *mdst, const int32_t *deltas, const tm_tuning *tune, uint64_t *counts) {
  if (!g || nq < 1 || !num_edges || !msrc || !mdst || !deltas || !counts) return fail(TM_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(g->mtx);
  HIPTRY(hipSetDevice(g->device));
  // Queries are grouped by shape class, so no launch sizes its LDS stack for
  // a deeper motif than its own; groups (at most TM_LAUNCH_MAXQ queries) run
  // one after another, each with the whole GPU (measured on the 64M-edge
  // set, 238 queries in 4 classes: 11.95 s one after another, 13.3 s with
  // the classes' launches concurrent on the slot streams, where the
  // waiting wavefronts and the tail of one launch hold CUs the others need).
  std::vector<int> key(nq), order(nq);
  for (int q = 0; q < nq; q++) {
    if (num_edges[q] < 2 || num_edges[q] > TM_MAX_MOTIF_EDGES) return fail(TM_ERR_MOTIF, "motif needs 2..8 edges");
    int n = 0;
    for (int k = 0; k < num_edges[q]; k++)
      n = std::max(n, 1 + std::max(msrc[(size_t)q * TM_MAX_MOTIF_EDGES + k], mdst[(size_t)q * TM_MAX_MOTIF_EDGES + k]));
    key[q] = shape_class(num_edges[q], n);
    order[q] = q;
  }
  const int mode = tune ? tune->batch_mode : 0;
  std::vector<uint64_t> out(nq);
  Slot &sl = g->slot[0];
  if (mode <= 0) {
    // one single-query launch per query, in call order, alternating between
    // two slots (each its own stream, scheduler state and ring), so the tail
    // of one launch overlaps the start of the next; one synchronization per
    // slot for the call.  The single-query kernel runs 6-11% faster per
    // query than the multi-query one (a query index known at compile time),
    // and on a work-dominated set the launches of light queries are a few
    // percent of the time anyway.
    const int ns = nq > 1 ? 2 : 1;
    tm_status st = TM_OK;
    for (int k = 0; k < ns && st == TM_OK; k++) st = reserve(g, g->slot[k], nq);
    for (int q = 0; q < nq && st == TM_OK; q++) {
      Slot &sq = g->slot[q % ns];
      st = stage_queries(g, sq, q, 1, &q, num_edges, msrc, mdst, TM_MAX_MOTIF_EDGES, deltas, nullptr, nullptr);
      if (st == TM_OK) st = launch(g, sq, q, 1, false, class_tuning(g, key[q], tune), 0, false);
      if (st == TM_OK) {
        hipLaunchKernelGGL(error_kernel, dim3(1), dim3(64), 0, sq.stream, sq.s.d_sched, sq.s.d_stops + q);
        HIPTRY(hipGetLastError());
      }
    }
    std::vector<int> err(nq);
    std::vector<uint64_t> part(nq);
    std::vector<int> perr(nq);
    for (int k = 0; k < ns && st == TM_OK; k++) {
      Slot &sk = g->slot[k];
      st = finish(g, sk, 0, nq, part.data());
      if (st == TM_OK)
        HIPTRY(hipMemcpy(perr.data(), sk.s.d_stops, sizeof(int) * nq, hipMemcpyDeviceToHost));
      for (int q = k; q < nq && st == TM_OK; q += ns) {
        counts[q] = part[q];
        err[q] = perr[q];
      }
    }
    if (st != TM_OK) return st;
    int bits = 0;
    for (int q = 0; q < nq; q++) bits |= err[q];
    if (bits & ~4) return fail(TM_ERR_INTERNAL, "kernel invariant violated (error bits " + std::to_string(bits) + ")");
    if (bits & 4) return fail(TM_ERR_TIMEOUT, "a query exceeded its time limit; counts discarded");
    return TM_OK;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  // fused: every query in one launch (its stack sized for the deepest class,
  // whose tuning entry it takes), at most TM_LAUNCH_MAXQ per launch
  const bool fuse = mode >= 2;
  struct Group { int first, n; };
  std::vector<Group> groups;
  for (int i = 0; i < nq;) {
    int j = i;
    while (j < nq && (fuse || key[order[j]] == key[order[i]]) && j - i < TM_LAUNCH_MAXQ) j++;
    groups.push_back({i, j - i});
    i = j;
  }
  for (const Group &gr : groups) {
    tm_status st = stage_queries(g, sl, 0, gr.n, order.data() + gr.first, num_edges, msrc, mdst, TM_MAX_MOTIF_EDGES,
                                 deltas, nullptr, nullptr);
    const int cls = key[order[gr.first + gr.n - 1]];  // the deepest class of the launch
    tm_tuning merged;
    const tm_tuning *t = class_tuning(g, cls, tune);
    if (fuse) {
      // the deepest class's table entry, with the call's own nonzero fields on top
      merged = (cls >= 0 && cls < TM_NCLASSES && g->class_set[cls]) ? g->class_tune[cls] : tm_tuning{};
      const int32_t *src = (const int32_t *)tune;
      int32_t *dst = (int32_t *)&merged;
      for (size_t f = 0; f < sizeof(tm_tuning) / sizeof(int32_t); f++)
        if (src[f]) dst[f] = src[f];
      t = &merged;
    }
    if (st == TM_OK) st = launch(g, sl, 0, gr.n, false, t, false, true);
    if (st == TM_OK) st = finish(g, sl, 0, gr.n, out.data() + gr.first);
    if (st != TM_OK) return st;
  }
  for (int i = 0; i < nq; i++) counts[order[i]] = out[i];
  return TM_OK;
}

static tm_status enum_impl(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                           int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune, int32_t *rows,
                           uint64_t *count, int64_t *rows_written, tm_launch_fn fn, void *user, bool cut_with_fn);

tm_status tm_enum(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                  int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune, int32_t *rows, uint64_t *count,
                  int64_t *rows_written) {
  return enum_impl(g, num_edges, msrc, mdst, delta, beg, end, cap, tune, rows, count, rows_written, nullptr, nullptr,
                   false);
}

tm_status tm_enum_custom(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                         int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune, tm_launch_fn launcher,
                         void *user, int32_t *rows, uint64_t *count, int64_t *rows_written) {
  if (!launcher) return fail(TM_ERR_ARG, "null launcher");
  return enum_impl(g, num_edges, msrc, mdst, delta, beg, end, cap, tune, rows, count, rows_written, launcher, user,
                   false);
}

tm_status tm_enum_custom_counting(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                                  int32_t delta, int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune,
                                  tm_launch_fn launcher, void *user, int32_t *rows, uint64_t *count,
                                  int64_t *rows_written) {
  if (!launcher) return fail(TM_ERR_ARG, "null launcher");
  return enum_impl(g, num_edges, msrc, mdst, delta, beg, end, cap, tune, rows, count, rows_written, launcher, user,
                   true);
}

static tm_status enum_impl(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                           int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune, int32_t *rows,
                           uint64_t *count, int64_t *rows_written, tm_launch_fn fn, void *user, bool cut_with_fn) {
  if (!g || !msrc || !mdst || !count || cap < 0 || (cap > 0 && !rows)) return fail(TM_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(g->mtx);
  HIPTRY(hipSetDevice(g->device));
  Slot &sl = g->slot[0];
  tm_status st = stage_queries(g, sl, 0, 1, nullptr, &num_edges, msrc, mdst, num_edges, &delta, &beg, &end);
  if (st != TM_OK) return st;
  Scratch &s = sl.s;
  const int32_t rbeg = s.h_queries[0].root_beg, rend = s.h_queries[0].root_end;
  int32_t *d_rows = nullptr;
  const size_t row_bytes = sizeof(int32_t) * (size_t)std::max<int64_t>(cap, 1) * num_edges;
  HIPTRY(hipMalloc(&d_rows, row_bytes));
  s.h_queries[0].rows = d_rows;
  s.h_queries[0].cap = cap;
  if ((st = launch(g, sl, 0, 1, true, tune, false, false, fn, user)) != TM_OK) { (void)hipFree(d_rows); return st; }
  st = finish(g, sl, 0, 1, count);
  if (st == TM_OK && *count < (uint64_t)cap) {
    // below the cap no root was held back: every match is written
    if (*count > 0) {
      hipError_t he = hipMemcpy(rows, d_rows, sizeof(int32_t) * (size_t)*count * num_edges, hipMemcpyDeviceToHost);
      if (he != hipSuccess) st = fail(TM_ERR_HIP, hipGetErrorString(he));
    }
    if (rows_written) *rows_written = (int64_t)*count;
  } else if (st == TM_OK) {
    // at the cap: replace the schedule-dependent subset by the deterministic one
    (void)hipFree(d_rows);
    d_rows = nullptr;
    int64_t nw = 0;
    // the cut counts with the caller's kernel only when it counts (else the
    // library's kernels, whose match semantics the caller's kernel shares)
    st = enum_cut(g, sl, num_edges, msrc, mdst, delta, rbeg, rend, cap, tune, rows, count, &nw,
                  cut_with_fn ? fn : nullptr, cut_with_fn ? user : nullptr);
    if (rows_written) *rows_written = st == TM_OK ? nw : 0;
  }
  if (d_rows) (void)hipFree(d_rows);
  return st;
}

}  // extern "C"
