/*
 * tm_abi.h -- C ABI of the MI355X temporal-motif query library (libtm_hip.so).
 *
 * Plain C types only: int32/int64/uint64 scalars, pointers to host arrays and
 * an opaque graph handle.  Every entry point returns a tm_status; on failure
 * tm_last_error() holds a message for the calling thread.  Device work runs
 * on the library's own HIP stream for the graph's device and is complete when
 * the call returns (the reference's launches are synchronous from the
 * caller's view: a D2H copy ends each run, system/templates/Impl/result.snp).
 *
 * Semantics of a match (reference: plugins/GPUWorkerU1B.cu:10-263 and
 * system/templates/common/ snippets): a strictly increasing tuple of positions
 * p_0 < ... < p_{M-1} in the stably time-sorted edge list, an injective motif
 * node -> graph node map, t(p_{M-1}) - t(p_0) <= delta (inclusive), root
 * p_0 in [beg, end) with u != v.  Counts are exact uint64.
 */
#ifndef TM_ABI_H_
#define TM_ABI_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TM_ABI_VERSION 7
#define TM_MAX_MOTIF_EDGES 8
#define TM_MAX_MOTIF_NODES 8

typedef enum tm_status {
  TM_OK = 0,
  TM_ERR_ARG = 1,       /* bad argument (null pointer, range, size) */
  TM_ERR_MOTIF = 2,     /* motif not prefix-connected, self-loop, too large */
  TM_ERR_HIP = 3,       /* a HIP runtime call failed */
  TM_ERR_NOMEM = 4,     /* device allocation failed or budget exceeded */
  TM_ERR_INTERNAL = 5,  /* kernel-side invariant violated (result discarded) */
  TM_ERR_NODEVICE = 6,  /* no usable GPU */
  TM_ERR_TIMEOUT = 7    /* the launch passed tm_tuning.time_limit_ms; results discarded */
} tm_status;

/* Opaque device-resident graph (one device). */
typedef struct tm_graph tm_graph;

/*
 * Reference-layout host graph: the arrays of corelib::data::HostGraphData
 * (include/data.h:174-186, filled by GraphDataLoader::createGraphData,
 * src/core/gpu/data.cu:124-348).  Eg holds 3 ints (u, v, t) per edge, edges
 * stably sorted by t, vertex ids renamed 0..V-1.
 */
typedef struct tm_host_graph {
  int32_t num_edges;
  int32_t num_vertices;
  const int32_t *Eg;         /* 3 * num_edges */
  const int32_t *inEdgesV;   /* num_edges: positions grouped by destination */
  const int32_t *outEdgesV;  /* num_edges: positions grouped by source */
  const int32_t *inEdgesR;   /* num_vertices + 1 */
  const int32_t *outEdgesR;  /* num_vertices + 1 */
  const int32_t *Euid;       /* num_edges: original line numbers */
} tm_host_graph;

/* Launch parameters; zero fields take the library defaults (occupancy-limited
 * waves up to 32 per CU, scan_max 16, no time
 * limit, steal_every 16, steal_max 4, steal_min 64, steal_after_us 1000,
 * claim_blocks: about 8 claims of one query's roots per wavefront). */
typedef struct tm_tuning {
  int32_t waves_per_cu;  /* resident wavefronts per CU in the persistent grid */
  int32_t claim_blocks;  /* 256-root blocks per claim of the root schedule */
  int32_t scan_max;      /* last-edge ranges up to this size are scanned, larger ones use the pair index */
  int32_t time_limit_ms; /* abort the launch after this wall time on the device (0 = no limit) */
  int32_t steal_every;   /* busy wavefronts look for idle ones every this many steps */
  int32_t steal_max;     /* frames a busy wavefront donates per look */
  int32_t steal_min;     /* a donated half-frame holds at least this many candidates */
  int32_t steal_after_us; /* a wavefront donates only after holding one task this long */
  int32_t waves_per_simd_hint; /* counting kernel variant: 1 = compiler's register allocation,
                                  4 = capped at 128 VGPRs (16 wavefronts per CU, no spills),
                                  5 = capped at 96 VGPRs (20 per CU), 6 = capped at 80 VGPRs
                                  (24 per CU); 0 = 5 */
  int32_t sb_flush;      /* counting: candidates that pass the filters of one level are buffered in
                            LDS and expanded together once this many (1..64) are held; -1 = each
                            step expands its own survivors; 0 = the library's choice */
  int32_t sb_inline;     /* counting: with nothing buffered, a step with at least this many survivors
                            (1..64) expands them itself; 0 = the library's choice */
  int32_t tail_swap;     /* counting, two-edge tail: when the last motif edge does not touch the node the
                            second-to-last one brings, a partial match whose second-to-last edge has at
                            least this many candidates in the window is counted over the last edge's
                            candidates instead, if those are fewer; 0 = library default (32), < 0 = never */
  int32_t batch_mode;    /* tm_batch_count: 0 = one single-query launch per query, back to back on one
                            stream, one synchronization; 1 = one multi-query launch per shape class;
                            2 = every query of the call in one multi-query launch (its stack sized for
                            the deepest motif) */
  int32_t sweep;         /* counting, 3-edge motifs whose last two edges take candidates from one list of a
                            root node (a star of two new nodes, or a new node's pair twice): > 0 = counted by
                            the window-sweep kernel (O(E log E) per query whatever delta), < 0 = by the
                            mining kernel like every other motif, 0 = the sweep when the roots' windows
                            are estimated (node degrees x delta / time span) at 8 or more slots on average.
                            The same choice for 4-edge motifs whose last edge repeats the root's pair
                            (edge 1 a new node on a root node's list, edge 2 a new node on edge 1's new
                            node's list): counted by the root-pair sweep kernel, which expands each
                            pair's edge-1 slots once for all of the pair's roots */
} tm_tuning;

/* Library version (TM_ABI_VERSION) and last error of the calling thread. */
int32_t tm_abi_version(void);
const char *tm_last_error(void);

/* Number of visible GPUs (0 when none; never fails). */
int32_t tm_device_count(void);

/*
 * Graph arrays built on the device (no graph handle needed).
 *
 * tm_time_order: `order` (n positions) = the stable order of the edges by
 * time T, ties in input order.  Replaces the host-side ranges::stable_sort of
 * EdgeListLoader::createEdgeListCache (src/core/cpu/data.cc:246-248).
 *
 * tm_csr_build: for a time-sorted edge list (U, V, T of n edges), the
 * reference's graph arrays: vertices renamed by first appearance (a source
 * before the destination of the same edge), Eg (3n ints: u, v, t), the
 * positions grouped by renamed source / destination in ascending order
 * (outEdgesV / inEdgesV, n each) with their offsets (outEdgesR / inEdgesR,
 * numVertices + 1 each), and vertexIds[k] = the raw id renamed to k.  The
 * caller sizes outEdgesR, inEdgesR and vertexIds for 2n + 1 entries (at most
 * 2n vertices).  Replaces GraphDataLoader::createGraphData's renaming and CSR
 * build (src/core/gpu/data.cu:124-348).
 */
tm_status tm_time_order(int32_t device, int64_t n, const int32_t *T, int32_t *order);
tm_status tm_csr_build(int32_t device, int64_t n, const int32_t *U, const int32_t *V, const int32_t *T, int32_t *Eg,
                       int32_t *outEdgesV, int32_t *inEdgesV, int32_t *outEdgesR, int32_t *inEdgesR,
                       int32_t *vertexIds, int32_t *numVertices);

/*
 * Upload a host graph to `device` and build the device-side indexes
 * (packed times, neighbour arrays, per-pair index, time buckets).
 * Replaces GraphDataManager::alloc + moveAsync + waitMove
 * (src/core/gpu/data.cu:384-472) and the pointer binding of DeviceJobData
 * (src/core/gpu/data.cu:496-529).
 */
tm_status tm_graph_upload(int32_t device, const tm_host_graph *hg, tm_graph **out);

/* Release a graph's device memory (reference: DevicePtrDeleter, data.cu:364-370). */
tm_status tm_graph_free(tm_graph *g);

/*
 * Device pointers of a resident graph's reference arrays, the fields of
 * DeviceGraphData / DeviceJobData a hand-written worker reads
 * (include/data.h:195-209 and 304-320; bound in
 * src/core/gpu/data.cu:496-529): Eg_d (TemporalEdge {u, v, t} per edge),
 * in/outEdgesV_d, in/outEdgesR_d, Euid_d, edgefeatures_d and
 * nodefeatures_d (null until tm_graph_set_features).  Valid until
 * tm_graph_free.
 */
typedef struct tm_device_arrays {
  int32_t num_edges;
  int32_t num_vertices;
  const void *Eg;
  const int32_t *inEdgesV;
  const int32_t *outEdgesV;
  const int32_t *inEdgesR;
  const int32_t *outEdgesR;
  const int32_t *Euid;
  const int32_t *edgeFeatures;     /* num_edges labels, or null */
  const int32_t *verticesFeatures; /* num_vertex_features labels, or null */
  int32_t num_vertex_features;
} tm_device_arrays;
tm_status tm_graph_arrays(const tm_graph *g, tm_device_arrays *out);

/*
 * Copy edge and vertex labels (HostGraphData::edgeFeatures and
 * verticesFeatures, include/data.h:174-186) next to a resident graph, as
 * GraphDataManager::moveAsync does (src/core/gpu/data.cu:440-472).  Replaces
 * labels set before.  The counting and enumeration kernels do not read them.
 */
tm_status tm_graph_set_features(tm_graph *g, const int32_t *edge_features, int32_t num_edge_features,
                                const int32_t *vertex_features, int32_t num_vertex_features);

/*
 * A device copy of `bytes` bytes of host memory on the graph's device (the
 * motif edges and MotifEdgeInfoV1 records of a DeviceJobData,
 * src/core/gpu/data.cu:510-528), and its release.
 */
tm_status tm_device_copy(tm_graph *g, const void *host, int64_t bytes, void **device_ptr);
tm_status tm_device_free(tm_graph *g, void *device_ptr);

/*
 * Work counters of the graph's last launch: [0] wavefront steps, [1] candidate
 * items examined, [2] roots dispatched, [3] last-edge range computations,
 * [4] frames pushed, [5..7] reference work of an instrumented run (X1, X2, P),
 * [8] frames donated between wavefronts, [9] donated frames taken, [10]
 * wavefronts that left early, [11] wall-clock ticks (100 MHz) wavefronts held
 * work, summed, and the launch shape: [12]
 * wavefronts in the grid, [13] blocks per claim, [14] LDS bytes per
 * wavefront, [15] register-cap variant, [16] queries of the launch counted
 * by the window-sweep or root-pair sweep kernels (their work is not in [0..11]).  Up to 17
 * entries.
 * No reference counterpart (instrumentation).
 */
tm_status tm_last_stats(const tm_graph *g, uint64_t *stats, int32_t n);

/* Device bytes held by a graph, and its edge / vertex counts. */
tm_status tm_graph_info(const tm_graph *g, int64_t *device_bytes, int32_t *num_edges,
                        int32_t *num_vertices, int32_t *device);

/*
 * Count the matches of one motif.  The motif is num_edges (src, dst) pairs
 * over renamed node ids (first appearance, src before dst) in time order, the
 * form HostMotifData holds (src/core/cpu/data.cc:411-420).  Roots are
 * [beg, end); end < 0 means num_edges_graph - num_edges + 1
 * (MineJob, src/core/cpu/data.cc:452-457).
 * Replaces GPUWorker::run of a generated worker (system/templates/GPUWorker/run.snp,
 * system/templates/Impl/top.snp; plugins/GPUWorkerU1B.cu:230-263).
 */
tm_status tm_count(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                   int32_t delta, int32_t beg, int32_t end, const tm_tuning *tune,
                   uint64_t *count);

/*
 * tm_count plus the reference's own work for the same query, counted exactly
 * by an instrumented kernel: work[0] = R roots examined (end - beg),
 * work[1] = X1 candidate-list resolutions of new-node edges, work[2] = X2
 * resolutions of closing edges, work[3] = P candidate probes, including the
 * probe that ends each scan at the window (the reference's scan loop,
 * system/templates/common/nextMatchInTask.snp:1-21, resolveMi.snp:1-29).
 * Algorithmic bytes of the reference: 12 R + 8 X1 + 16 X2 + 16 P.
 * No reference counterpart (measurement).
 */
tm_status tm_count_work(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                        int32_t delta, int32_t beg, int32_t end, const tm_tuning *tune,
                        uint64_t *count, uint64_t *work);

/*
 * tm_count plus the algorithmic bytes of this library's own search for the
 * same query: a metered copy of the counting kernel follows the same search
 * (root gap filter, pruning rules, two-edge tail, last-edge range
 * arithmetic) and adds up the bytes each of its memory reads needs: 16 per
 * lane and gap rule for a block of 4 roots, 12 per root edge record, 4 per
 * int32 word of a candidate (position, neighbour, first-candidate times) and
 * per probe of every binary or galloping search.  Launch parameters as
 * tm_count (register caps do not apply: one variant).  This is the unit
 * behind the roofline of the counting kernel: bytes / kernel time.
 * No reference counterpart (measurement).
 */
tm_status tm_count_bytes(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                         int32_t delta, int32_t beg, int32_t end, const tm_tuning *tune,
                         uint64_t *count, uint64_t *bytes);

/* Duration in ms of the graph's last mining kernel, from HIP events recorded
 * around it on the library's stream.  No reference counterpart (measurement). */
tm_status tm_last_kernel_ms(const tm_graph *g, double *ms);

/*
 * Enumerate matches: writes up to `cap` rows of num_edges int32 original line
 * numbers (Euid of p_0 .. p_{M-1}) to `rows` (host, cap * num_edges ints).
 * When the query has fewer than `cap` matches every match is written
 * (*count = *rows_written = the number of matches).  At the cap the
 * reference stops handing out roots and returns whichever rows the schedule
 * produced (system/tmine/KernelGen.py:509-528,
 * system/templates/enumerate/func.snp); here the rows are deterministic
 * instead: with r* the first root at which the cumulative count over
 * [beg, r*] reaches `cap`, every row of the roots [beg, r*) and, from root
 * r*'s rows, the smallest in lexicographic order of their Euid tuples, cap
 * rows in all; *count = the matches of the roots [beg, r*] (>= cap).
 * *rows_written = min(matches, cap).
 */
tm_status tm_enum(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                  int32_t delta, int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune,
                  int32_t *rows, uint64_t *count, int64_t *rows_written);

/*
 * Kernel launcher of a generated worker plugin (a motif-specialised
 * instantiation of the library's kernel template, compiled into the plugin).
 * mode 0: write the kernel's resident workgroups per CU for lds_bytes of
 * dynamic LDS to *blocks; mode 1: launch it with grid x 64 threads, lds_bytes
 * of LDS, on `stream` (a hipStream_t), passing *kernel_args (the library's
 * launch descriptor) by value.  Returns 0 on success.
 */
typedef int32_t (*tm_launch_fn)(int32_t mode, const void *kernel_args, int32_t grid, int32_t lds_bytes,
                                void *stream, int32_t *blocks, void *user);

/*
 * tm_count / tm_enum with the plugin's own kernel instead of the library's
 * table-driven one; staging, scheduling state, timing and result handling
 * are the library's.  Replaces TMotifMatchingGPUImpl of a generated worker
 * (system/templates/Impl/top.snp:1-10).
 */
tm_status tm_count_custom(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                          int32_t delta, int32_t beg, int32_t end, const tm_tuning *tune,
                          tm_launch_fn launcher, void *user, uint64_t *count);
tm_status tm_enum_custom(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                         int32_t delta, int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune,
                         tm_launch_fn launcher, void *user, int32_t *rows, uint64_t *count,
                         int64_t *rows_written);

/*
 * tm_enum_custom for a kernel that also counts: launched with the staged
 * query's rows pointer null, it counts the query's matches.  At the cap the
 * deterministic cut (tm_enum) is then found with the plugin's kernel too
 * (tm_enum_custom finds it with the library's own kernels, which is right
 * only for a kernel with the library's match semantics).  For a generated
 * worker under constraint features (label predicates, temporal gaps, a
 * filter: system/tmine/KernelGen.py:467-534), whose matches only its own
 * kernel knows.
 */
tm_status tm_enum_custom_counting(tm_graph *g, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                                  int32_t delta, int32_t beg, int32_t end, int64_t cap, const tm_tuning *tune,
                                  tm_launch_fn launcher, void *user, int32_t *rows, uint64_t *count,
                                  int64_t *rows_written);

/*
 * Shape classes and the graph's tuning table (the autotuner's output).  A
 * motif's class is 2 * max(1, M - 2) + (nodes > 4): the LDS stack depth and
 * node-map width of its launch, which set its occupancy.  tm_count,
 * tm_batch_count and the server use the class's entry when they are called
 * without launch parameters (tune == NULL or every field 0).  tune == NULL in
 * tm_set_class_tuning clears the entry.  No reference counterpart.
 */
#define TM_NCLASSES 16
tm_status tm_shape_class(int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t *cls);
tm_status tm_set_class_tuning(tm_graph *g, int32_t cls, const tm_tuning *tune);

/*
 * Count many motif queries on one graph in one call.  Query q uses edges
 * msrc/mdst[q * TM_MAX_MOTIF_EDGES ...], num_edges[q], deltas[q] and roots
 * [0, E - num_edges[q] + 1).  counts[q] receives the exact count, equal to
 * tm_count of the same query.  The queries are grouped by stack depth; each
 * group is one persistent launch of the batch kernel with the whole GPU, the
 * groups one after another.  No
 * reference counterpart: the reference runs one generated worker per query
 * (system/batch-query.py:12-19).
 */
tm_status tm_batch_count(tm_graph *g, int32_t nq, const int32_t *num_edges,
                         const int32_t *msrc, const int32_t *mdst, const int32_t *deltas,
                         const tm_tuning *tune, uint64_t *counts);

/*
 * tm_batch_count for callers that cannot allocate device memory: code under
 * HIP graph capture, serving loops and harnesses that own every device
 * allocation.  No counterpart in the reference.  Same queries, launches and
 * results as tm_batch_count, for at most TM_NOALLOC_MAXQ queries per call
 * (TM_ERR_ARG beyond).  Every device buffer of its launches (scheduler
 * state, query tables, counts, per-wavefront counters, a donation ring of
 * 2^15 records) lives in one static device array of
 * TM_NOALLOC_SCRATCH_BYTES, one copy per GPU the library runs on, held
 * until the process ends; the call makes no hipMalloc / hipFree.  Calls on
 * one GPU share that array: this library's calls take turns (one lock per
 * device), the array keeps the last call's contents, and launches that
 * overlap on one GPU through another copy of the library would race.
 */
#define TM_NOALLOC_MAXQ 256
#define TM_NOALLOC_SCRATCH_BYTES 7292160
tm_status tm_batch_count_noalloc(tm_graph *g, int32_t nq, const int32_t *num_edges, const int32_t *msrc,
                                 const int32_t *mdst, const int32_t *deltas, const tm_tuning *tune,
                                 uint64_t *counts);

/*
 * A prepared batch: the queries of tm_batch_count planned once, their tables
 * uploaded, and the whole set of per-class launches (reset, mining kernel,
 * statistics, result copies, on the graph's slot streams) captured into one
 * HIP graph; tm_batch_run replays it.  The graph records no memset: every
 * clear is a kernel.  Counts equal tm_batch_count's.  While a prepared batch
 * exists, other calls on its graph may run between its replays (each launch
 * resets the scheduler state it shares with them).  No reference counterpart.
 */
typedef struct tm_batch tm_batch;
tm_status tm_batch_prepare(tm_graph *g, int32_t nq, const int32_t *num_edges, const int32_t *msrc,
                           const int32_t *mdst, const int32_t *deltas, const tm_tuning *tune, tm_batch **out);
tm_status tm_batch_run(tm_batch *b, uint64_t *counts);
tm_status tm_batch_free(tm_batch *b);

/*
 * Serving: a request loop over a resident graph.  No reference counterpart
 * (the reference answers one YAML query per process run,
 * system/tmine/frontend.py:138-143).
 *
 * Any host thread submits count or enumerate requests.  A dispatcher thread
 * forms launches from whatever is queued whenever one of the server's launch
 * slots is free (continuous dispatch).  With max_batch 1 (the default) each
 * request is one single-query launch, the waiting request with the highest
 * response ratio (wait + S) / S first, S the service time of that query's
 * last run (1 ms for a query not seen yet): light requests overtake heavy
 * ones and heavy ones age to the front.  With max_batch > 1 the queued count
 * requests of the oldest request's stack-depth class go together into one
 * multi-query launch (at most max_batch).  An enumerate request runs alone
 * and is admitted only
 * once its row buffer (enum_cap * num_edges int32) fits in the enumeration
 * arena.  Launches on different slots run concurrently on the device, so
 * requests that arrive while others run overlap with them.  Each result is
 * the same as tm_count / tm_enum of the request alone (rows at the cap
 * included: they are chosen as tm_enum chooses them).  While a server
 * exists, other calls on its graph wait.
 */
typedef struct tm_server tm_server;

typedef struct tm_server_config {
  int32_t slots;        /* launch slots used, 1 .. 4 (0: 2) */
  int32_t max_batch;    /* count requests per launch, 1 .. 512 (0: 1) */
  int64_t arena_bytes;  /* device memory for enumeration rows in flight (0: 1 GiB) */
  tm_tuning tune;       /* launch parameters of every launch */
} tm_server_config;

typedef struct tm_server_result {
  int32_t status;       /* TM_OK or the request's error */
  int32_t batch;        /* requests in the launch that served it */
  uint64_t count;       /* exact match count */
  int64_t rows;         /* enumerated rows held for tm_server_rows */
  double latency_ms;    /* submit -> result on the host */
  double queue_ms;      /* submit -> launch */
} tm_server_result;

tm_status tm_server_create(tm_graph *g, const tm_server_config *cfg, tm_server **out);
/* enum_cap < 0: count only; >= 0: enumerate up to enum_cap rows. */
tm_status tm_server_submit(tm_server *s, int32_t num_edges, const int32_t *msrc, const int32_t *mdst,
                           int32_t delta, int64_t enum_cap, uint64_t *ticket);
/* Wait up to timeout_ms (< 0: no limit) for a request; TM_ERR_TIMEOUT if it
 * is still running (nothing is lost: wait again). */
tm_status tm_server_wait(tm_server *s, uint64_t ticket, int32_t timeout_ms, tm_server_result *res);
/* Copy a finished request's rows (at most max_rows) and forget the request. */
tm_status tm_server_take_rows(tm_server *s, uint64_t ticket, int32_t *rows, int64_t max_rows,
                              int64_t *rows_written);
/* Forget a finished request. */
tm_status tm_server_release(tm_server *s, uint64_t ticket);
/* [0] launches, [1] count requests served, [2] enumerate requests served,
 * [3] largest batch, [4] requests waiting, [5] requests running. */
tm_status tm_server_stats(tm_server *s, uint64_t *stats, int32_t n);
/* Finish every submitted request, stop the dispatcher and free the server. */
tm_status tm_server_destroy(tm_server *s);

#ifdef __cplusplus
}
#endif

#endif /* TM_ABI_H_ */
