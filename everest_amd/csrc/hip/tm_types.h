// tm_types.h -- layouts shared by the gfx950 kernels and their host driver.
// Internal to the library: the public boundary is include/tm_abi.h, which
// passes only plain pointers and sizes.
#pragma once
#include <stdint.h>

#define TM_MAXE 8  // most motif edges handled (the reference register stack: 8-bit stackio)
#define TM_MAXN 8  // most motif nodes handled

// How motif edge j (j >= 1) extends a partial match.
enum {
  TM_K_NEW_DST = 0,  // src a mapped, dst b new: candidates = out-list of m[a]
  TM_K_NEW_SRC = 1,  // dst b mapped, src a new: candidates = in-list of m[b]
  TM_K_CLOSE = 2     // both mapped: candidates = the (m[a] -> m[b]) pair segment
};

// Where the candidate list of motif edge j starts, given the position p matched
// for edge j-1 (fixed per motif edge by which node the list hangs off).
enum {
  TM_S_SEARCH = 0,     // binary search for the first position > p
  TM_S_OUT_SELF = 1,   // list out(src of p):  outSelf[p] + 1
  TM_S_OUT_NEXT = 2,   // list out(dst of p):  outNext[p]
  TM_S_IN_SELF = 3,    // list in(dst of p):   inSelf[p] + 1
  TM_S_IN_NEXT = 4,    // list in(src of p):   inNext[p]
  TM_S_PAIR_SELF = 5   // same pair as edge j-1: outPairSelf[p] + 1
};

// Time of the first candidate of edge j after position p, per start rule
// (INT_MAX when there is none): one load decides that a list has nothing in
// the window.
enum {
  TM_T_NONE = -1,
  TM_T_OUT_SELF = 0,  // next out-edge of src p
  TM_T_OUT_NEXT = 1,  // next out-edge of dst p
  TM_T_IN_SELF = 2,   // next in-edge of dst p
  TM_T_IN_NEXT = 3,   // next in-edge of src p
  TM_T_PAIR_SELF = 4, // next edge src p -> dst p
  TM_T_RECIP = 5      // next edge dst p -> src p
};
#define TM_NTFIRST 6
#define TM_QT_XRULE_SHIFT 8  // query table rule word: edge 1's rule | the root's further rules << TM_QT_XRULE_SHIFT

typedef struct { int u, v, t; } tm_edge;  // reference TemporalEdge (include/data.h:27-29)

typedef struct {
  int M;  // motif edges
  int n;  // motif nodes
  int kind[TM_MAXE];
  int a[TM_MAXE];       // motif source node of edge j
  int b[TM_MAXE];       // motif destination node of edge j
  int mapped[TM_MAXE];  // nodes mapped before edge j (a new node gets this index)
  int smode[TM_MAXE];   // TM_S_* start rule of edge j's candidate list
  int tmode[TM_MAXE];   // TM_T_* first-candidate time array of edge j (or TM_T_NONE)
  int xmask[TM_MAXE];   // TM_T_* rules (bit r) that later edges k >= j + 2 impose at the nodes of
                        // edge j's match: each rule's next edge after that position must be in the window
} tm_motif_dev;

// Device-resident graph: reference arrays plus the indexes built at upload.
typedef struct {
  int E, V;
  const tm_edge *Eg;      // sorted by time, renamed ids
  const int *T;           // Eg[p].t, packed
  const int *outR, *outPos, *outNbr;  // out CSR: offsets, positions (outEdgesV), destinations
  const int *inR, *inPos, *inNbr;     // in CSR: offsets, positions (inEdgesV), sources
  const int *outPairNbr, *outPairPos; // per node, out-list re-sorted by (dst, position)
  const int *inPairNbr, *inPairPos;   // per node, in-list re-sorted by (src, position)
  const int *outSelf;     // slot of position p in out(src of p)
  const int *inSelf;      // slot of position p in in(dst of p)
  const int *outNext;     // first slot of out(dst of p) with position > p
  const int *inNext;      // first slot of in(src of p) with position > p
  const int *outPairSelf; // slot of position p in the out-pair index of src of p
  const int *tfirst[TM_NTFIRST];  // TM_T_* arrays, one entry per position
  // The same first-candidate times in the slot order of each candidate list
  // (TM_L_*), one allocation: stf[(l * TM_NTFIRST + k) * E + s] =
  // tfirst[k][position of slot s of list l], so the pruning test of a
  // candidate is a coalesced load next to its position and neighbour instead
  // of a gather (and the array is addressed by arithmetic, not by indexing
  // a pointer table).
  const int *stf;
  // Root gaps, one array per TM_T_* rule: tfirst - T[p] clamped to INT_MAX,
  // or 0xffffffff when p is a self-loop or has no candidate; a root passes a
  // query with that rule for edge 1 iff gap <= delta.  Padded with
  // 0xffffffff to a multiple of 256 entries plus 256, so a 256-root block
  // read as one 16-byte load per lane never leaves the array.
  const unsigned *rgap;    // TM_NTFIRST arrays of epad entries, one allocation
  int epad;               // entries per rgap array
  // Per slot of the out-lists (entries 0..E-1) and of the in-lists (E..2E-1):
  // the position of the previous edge of the same graph pair, or -1
  const int *sprev;
  int no_self_loops;      // 1 when no edge has u == v: corrections for a list's own node are 0
  const int *bucket;      // bucket[k] = first position with T >= tmin + (k << bshift)
  int tmin, bshift;
  long long nbuckets;
  const int *Euid;        // original line number of each position
} tm_graph_view;

// Candidate lists, in the order of stf
enum { TM_L_OUT = 0, TM_L_IN = 1, TM_L_OUTPAIR = 2, TM_L_INPAIR = 3 };

typedef struct {
  tm_motif_dev m;
  int delta;
  int root_beg, root_end;
  int pad0;
  unsigned long long *count;  // result slot (device)
  int *rows;                  // enumeration rows (device), cap * M ints
  long long cap;
  int *stop;                  // set once cap rows are claimed: no more roots of this query
} tm_query_dev;

#define TM_STRIPES 8  // root-block counters, one 128-byte line each (one per XCD slot)

typedef struct {
  unsigned long long next_chunk[TM_STRIPES][16];  // blocks claimed per stripe
  int error;        // bit 0: frame stack overflow, bit 1: empty frame, bit 2: deadline passed,
                    // bit 3: a donated frame never became ready
  int abort;        // set once the deadline passed: every wavefront stops
  long long t0;     // wall clock at launch (set by the reset kernel)
  unsigned long long stats[16];  // work counters (see TM_STAT_*)
  // each on its own 128-byte line: donors, takers and waiters do not queue
  // behind each other's atomics
  unsigned int pool_head;  // tickets drawn by idle wavefronts
  int pad0[31];
  unsigned int pool_tail;  // donated records reserved so far
  int pad1[31];
  int busy;         // wavefronts holding work + records reserved and not yet taken
  int pad2[31];
  int waiters;      // wavefronts waiting for donations (at most idle_cap)
  int pad4[31];
  int gen;          // launch generation of the donation ring: the reset kernel advances it,
                    // so replays of a captured launch get fresh tags
  int pad3[31];
} tm_sched;

// A frame donated by a busy wavefront to an idle one: fields q, lvl, cur,
// end, (unused), pend, tl, m[TM_MAXN], e[TM_MAXE], each stored as a 64-bit word
// (tag << 32 | value) with device-scope atomic stores, so a taker knows the
// record is complete when every word carries the tag it expects: no fences,
// no cache invalidation.  Records live in a ring of pool_cap slots; record r
// goes to slot r mod pool_cap with tag (launch generation, lap r / pool_cap),
// and its taker stores that tag into pool_done[slot] once the frame is
// copied out, which lets the slot's next lap be written.
#define TM_POOL_FIELDS 23
#define TM_GEN_BITS 18  // launch generations cycle through 1 .. 2^18 - 1

enum {
  TM_STAT_STEPS = 0,   // wavefront expansion steps
  TM_STAT_ITEMS = 1,   // candidate slots examined
  TM_STAT_ROOTS = 2,   // roots dispatched
  TM_STAT_LEAVES = 3,  // last-edge range computations
  TM_STAT_PUSHES = 4,  // frames pushed
  TM_STAT_REF_X1 = 5,  // reference list resolutions, new-node edges (instrumented run)
  TM_STAT_REF_X2 = 6,  // reference list resolutions, closing edges (instrumented run)
  TM_STAT_REF_P = 7,   // reference candidate probes (instrumented run)
  TM_STAT_DONATED = 8, // frames donated to the pool
  TM_STAT_TAKEN = 9,   // donated frames taken
  TM_STAT_CAPEXIT = 10,// wavefronts that left because enough were waiting
  TM_STAT_WORK = 11,   // wall-clock ticks wavefronts held work (summed)
  TM_STAT_GRID = 12,   // launch shape, filled in by the host: wavefronts,
  TM_STAT_CLAIM = 13,  //   256-root blocks per claim,
  TM_STAT_LDS = 14,    //   LDS bytes per wavefront,
  TM_STAT_WPE = 15     //   register-cap variant (1: none, 6: 80 VGPRs, 0: caller's kernel)
};
#define TM_NSTATS 16  // per-wavefront counter slot (one 128-byte line)
#define TM_STAT_SWEEP 16      // host only: queries of the launch counted by the window-sweep kernel
#define TM_NSTATS_HOST 17

typedef struct {
  tm_graph_view g;
  const tm_query_dev *queries;
  // 4 ints per query: (delta, root_beg, root_end, start rule of edge 1)
  const int *qtab;
  // The launch's roots as one space of 256-root blocks, query after query:
  // query q owns blocks [block_prefix[q], block_prefix[q + 1]), its block i
  // starting at root (root_beg & ~3) + 256 i (nq + 1 entries).
  const long long *block_prefix;
  int nq;
  long long total_blocks;
  int claim_blocks;           // blocks per claim (the same for every query of the launch)
  int frames;                 // LDS frame capacity per wavefront
  int scan_max;               // last-edge ranges up to this size are scanned
  int sb_flush;               // counting: survivors buffered before a full-width expansion (0: none buffered)
  int sb_inline;              // counting: with nothing buffered, a step with this many survivors expands them itself
  long long limit_ticks;      // wall-clock budget of the launch (0 = none)
  tm_sched *sched;
  unsigned long long *wstats; // TM_NSTATS counters per wavefront (workgroup)
  unsigned long long *pool;   // donated frames, TM_POOL_FIELDS tagged words each
  unsigned *pool_done;        // per ring slot: tag of the last record copied out
  int pool_cap;               // ring slots, a power of two
  int pool_shift;             // log2(pool_cap)
  int steal_every;            // busy wavefronts look for idle ones every this many steps
  int steal_max;              // frames donated per look
  int idle_cap;               // most wavefronts kept waiting on tickets without records
  int steal_min;              // a donated half holds at least this many candidates
  long long steal_after_ticks; // a task donates only after running this long (wall clock)
  int t2_min;                 // counting: frames of edge M-2 with at least this many candidates are summed
                              // over the last edge's candidates when that list is shorter (two-edge tail)
  int pad_t2;
  // The same view as g, in device memory (written at upload).  The mining
  // kernel reads the graph through it: a view held in the kernel's own
  // variables ends up in per-lane scratch memory once the compiler selects
  // between two of its fields' addresses (sunk loads of out- vs in-lists).
  const tm_graph_view *gv;
  // labels set by tm_graph_set_features (null when none): per edge position
  // and per vertex, read by kernels of queries with label constraints
  const int *edgeFeature;
  const int *vertexFeature;
} tm_kernel_args;
