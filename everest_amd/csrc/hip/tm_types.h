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

typedef struct { int u, v, t; } tm_edge;  // reference TemporalEdge (include/data.h:27-29)

typedef struct {
  int M;  // motif edges
  int n;  // motif nodes
  int kind[TM_MAXE];
  int a[TM_MAXE];       // motif source node of edge j
  int b[TM_MAXE];       // motif destination node of edge j
  int mapped[TM_MAXE];  // nodes mapped before edge j (a new node gets this index)
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
  const int *bucket;      // bucket[k] = first position with T >= tmin + (k << bshift)
  int tmin, bshift;
  long long nbuckets;
  const int *Euid;        // original line number of each position
} tm_graph_view;

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

typedef struct {
  unsigned long long next_chunk;
  int error;
  int pad[13];
} tm_sched;

typedef struct {
  tm_graph_view g;
  const tm_query_dev *queries;
  int nq;
  int chunk_roots;            // roots per dispatched chunk (multiple of 64)
  const long long *chunk_prefix;  // nq + 1 entries: first chunk of each query
  long long total_chunks;
  int frames;                 // LDS frame capacity per wavefront
  int scan_max;               // last-edge ranges up to this size are scanned
  tm_sched *sched;
} tm_kernel_args;
