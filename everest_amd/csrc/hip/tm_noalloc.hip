// tm_noalloc.hip -- tm_batch_count without device allocation
// (tm_batch_count_noalloc in include/tm_abi.h).
//
// Every device buffer the batch launch needs beyond the graph and the
// caller's results -- scheduler state, query tables, per-query counts and
// stop flags, per-wavefront counters, the donation ring -- lives in one
// file-scope device array, `scratch`, of a fixed size.  Its device address is
// taken once per device with hipGetSymbolAddress (the array's name in host
// code is only a host-side stand-in) and used for every kernel argument and
// copy after that.  No hipMalloc / hipFree on this path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tm_abi.h"
#include "tm_internal.h"

using namespace tmh;

namespace {

constexpr int kMaxQ = TM_NOALLOC_MAXQ;
constexpr int kMaxWaves = 256 * 32;  // per-wavefront counter slots: 256 CUs x 32 wavefronts
constexpr int kPoolShift = 15;       // donation ring of 2^15 records
constexpr size_t kAlign = 256;

constexpr size_t up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
// layout of the array (byte offsets)
constexpr size_t kOffSched = 0;
constexpr size_t kOffQueries = kOffSched + up(sizeof(tm_sched));
constexpr size_t kOffQtab = kOffQueries + up(sizeof(tm_query_dev) * kMaxQ);
constexpr size_t kOffPrefix = kOffQtab + up(sizeof(int) * 4 * kMaxQ);
constexpr size_t kOffCounts = kOffPrefix + up(sizeof(long long) * (kMaxQ + 1));
constexpr size_t kOffStops = kOffCounts + up(sizeof(unsigned long long) * kMaxQ);
constexpr size_t kOffWstats = kOffStops + up(sizeof(int) * kMaxQ);
constexpr size_t kOffPool = kOffWstats + up(sizeof(unsigned long long) * TM_NSTATS * (size_t)kMaxWaves);
constexpr size_t kOffPoolDone = kOffPool + up(sizeof(unsigned long long) * TM_POOL_FIELDS * ((size_t)1 << kPoolShift));
constexpr size_t kBytes = kOffPoolDone + up(sizeof(unsigned) * ((size_t)1 << kPoolShift));
static_assert(kBytes == TM_NOALLOC_SCRATCH_BYTES, "TM_NOALLOC_SCRATCH_BYTES out of date");

// The scratch array: one copy per GPU the library runs on, held until the
// process ends.
static __device__ __align__(256) unsigned char scratch[kBytes];

__global__ void zero_kernel(unsigned long long *p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

struct DeviceScratch {
  bool ready = false;
  unsigned char *ptr = nullptr;  // device address of `scratch` on this device
  int gen = 0;                   // launches since the ring was cleared (as Slot::gen)
  std::mutex mtx;                // calls on one device take turns
};

DeviceScratch g_dev[64];
std::mutex g_setup;

// Once per device, with it current: the array's device address, and the
// array cleared (the ring's tags, pool_done and the generation start at 0).
tm_status setup(int device, hipStream_t stream, DeviceScratch **out) {
  if (device < 0 || device >= 64) return set_error(TM_ERR_ARG, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_setup);
  DeviceScratch &d = g_dev[device];
  if (!d.ready) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(TM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.ptr = (unsigned char *)p;
    const size_t n = kBytes / sizeof(unsigned long long);
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       (unsigned long long *)d.ptr, n);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(stream)) != hipSuccess)
      return set_error(TM_ERR_HIP, std::string("scratch clear: ") + hipGetErrorString(e));
    d.ready = true;
  }
  *out = &d;
  return TM_OK;
}

}  // namespace

extern "C" {

tm_status tm_batch_count_noalloc(tm_graph *g, int32_t nq, const int32_t *num_edges, const int32_t *msrc,
                                 const int32_t *mdst, const int32_t *deltas, const tm_tuning *tune,
                                 uint64_t *counts) {
  if (!g || nq < 1 || !num_edges || !msrc || !mdst || !deltas || !counts) return set_error(TM_ERR_ARG, "bad argument");
  if (nq > kMaxQ)
    return set_error(TM_ERR_ARG, "tm_batch_count_noalloc takes at most " + std::to_string(kMaxQ) +
                                     " queries per call (its scratch array is sized for that)");
  std::lock_guard<std::mutex> lk(g->mtx);
  hipError_t he = hipSetDevice(g->device);
  if (he != hipSuccess) return set_error(TM_ERR_HIP, hipGetErrorString(he));
  Slot &real = g->slot[0];  // its stream and timing events; its buffers stay untouched
  DeviceScratch *ds = nullptr;
  tm_status st = setup(g->device, real.stream, &ds);
  if (st != TM_OK) return st;
  std::lock_guard<std::mutex> dlk(ds->mtx);
  // a slot whose device buffers are the scratch array's pieces
  Slot view;
  view.index = real.index;
  view.stream = real.stream;
  view.ev0 = real.ev0;
  view.ev1 = real.ev1;
  view.gen = ds->gen;
  view.own_ring = true;  // launches counted on this view: the ring is the scratch array's
  Scratch &s = view.s;
  s.cap_q = kMaxQ;
  s.d_sched = (tm_sched *)(ds->ptr + kOffSched);
  s.d_queries = (tm_query_dev *)(ds->ptr + kOffQueries);
  s.d_qtab = (int *)(ds->ptr + kOffQtab);
  s.d_prefix = (long long *)(ds->ptr + kOffPrefix);
  s.d_counts = (unsigned long long *)(ds->ptr + kOffCounts);
  s.d_stops = (int *)(ds->ptr + kOffStops);
  s.d_wstats = (unsigned long long *)(ds->ptr + kOffWstats);
  s.wstats_cap = kMaxWaves;
  s.d_pool = (unsigned long long *)(ds->ptr + kOffPool);
  s.d_pool_done = (unsigned *)(ds->ptr + kOffPoolDone);
  s.pool_shift = kPoolShift;
  s.pool_cap = 1 << kPoolShift;
  s.h_queries.resize(kMaxQ);
  s.h_qtab.resize(4 * (size_t)kMaxQ);
  s.h_prefix.resize(kMaxQ + 1);
  // the queries by shape class, one launch per class, as tm_batch_count
  std::vector<int> key(nq), order(nq);
  for (int q = 0; q < nq; q++) {
    if (num_edges[q] < 2 || num_edges[q] > TM_MAX_MOTIF_EDGES) return set_error(TM_ERR_MOTIF, "motif needs 2..8 edges");
    key[q] = shape_class_of(num_edges[q], msrc + (size_t)q * TM_MAX_MOTIF_EDGES, mdst + (size_t)q * TM_MAX_MOTIF_EDGES);
    order[q] = q;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<uint64_t> out(nq);
  for (int i = 0; i < nq && st == TM_OK;) {
    int j = i;
    while (j < nq && key[order[j]] == key[order[i]]) j++;
    st = stage_queries(g, view, 0, j - i, order.data() + i, num_edges, msrc, mdst, TM_MAX_MOTIF_EDGES, deltas,
                       nullptr, nullptr);
    const tm_tuning *t = tuning_for_class(g, key[order[i]], tune);
    if (st == TM_OK) st = launch(g, view, 0, j - i, false, t, 0, true);
    ds->gen = view.gen;  // the launch advanced the generation (and cleared the ring on a wrap)
    if (st == TM_OK) st = finish(g, view, 0, j - i, out.data() + i);
    i = j;
  }
  if (st != TM_OK) return st;
  for (int i = 0; i < nq; i++) counts[order[i]] = out[i];
  return TM_OK;
}

}  // extern "C"
