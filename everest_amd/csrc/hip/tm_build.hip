// tm_build.hip -- the graph arrays of the query path built on the device:
// the stable time order of an edge list, and for a time-sorted edge list the
// first-appearance vertex renaming, Eg and both CSR indexes.
//
// Reference: EdgeListLoader::createEdgeListCache sorts by time with
// ranges::stable_sort (src/core/cpu/data.cc:246-248); createGraphData
// renames vertices by first appearance in the sorted list, a source before a
// destination of the same edge (src/core/gpu/data.cu:143-236), and builds the
// CSR positions grouped by renamed source / destination in ascending position
// order with their offsets (data.cu:238-271).  Here every step is a stable
// radix sort or a streaming pass:
//   time order   sort (time, position) pairs by time: stable, so ties keep
//                file order
//   renaming     sort (raw id, slot) over the 2n interleaved endpoints
//                (slot 2i = source of edge i, 2i + 1 = its destination): the
//                head of each id's run holds its first slot; sorting the
//                distinct ids by first slot gives each its new id
//   CSR          sort (new id, position) pairs of the sources (destinations):
//                positions ascending within an id; offsets from the run heads
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>

#include "tm_abi.h"
#include "tm_internal.h"

namespace {

constexpr unsigned kFlip = 0x80000000u;  // signed order as unsigned keys

__global__ void iota_kernel(int *__restrict__ a, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (int)i;
}

__global__ void flip_kernel(const int *__restrict__ x, unsigned *__restrict__ k, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) k[i] = (unsigned)x[i] ^ kFlip;
}

// slot 2i: source of edge i, slot 2i + 1: its destination
__global__ void interleave_kernel(const int *__restrict__ U, const int *__restrict__ V, long long n,
                                  unsigned *__restrict__ key, int *__restrict__ slot) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    key[2 * i] = (unsigned)U[i] ^ kFlip;
    key[2 * i + 1] = (unsigned)V[i] ^ kFlip;
    slot[2 * i] = (int)(2 * i);
    slot[2 * i + 1] = (int)(2 * i + 1);
  }
}

__global__ void head_kernel(const unsigned *__restrict__ key, long long m, int *__restrict__ head) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// per distinct id (group g = run index in id order): its first slot and raw id
__global__ void first_kernel(const unsigned *__restrict__ key, const int *__restrict__ slot,
                             const int *__restrict__ head, const int *__restrict__ grp, long long m,
                             unsigned *__restrict__ first, int *__restrict__ rawk) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && head[i]) {
    const int g = grp[i] - 1;
    first[g] = (unsigned)slot[i];
    rawk[g] = (int)(key[i] ^ kFlip);
  }
}

// order[r] = the group with the r-th smallest first slot: new id r
__global__ void rank_kernel(const int *__restrict__ order, const int *__restrict__ rawk, int nv,
                            int *__restrict__ rank, int *__restrict__ raw) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nv) {
    rank[order[r]] = r;
    raw[r] = rawk[order[r]];
  }
}

__global__ void assign_kernel(const int *__restrict__ slot, const int *__restrict__ grp, const int *__restrict__ rank,
                              long long m, int *__restrict__ Un, int *__restrict__ Vn) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) {
    const int s = slot[i], id = rank[grp[i] - 1];
    if (s & 1) Vn[s >> 1] = id; else Un[s >> 1] = id;
  }
}

__global__ void eg_kernel(const int *__restrict__ Un, const int *__restrict__ Vn, const int *__restrict__ T,
                          long long n, int *__restrict__ Eg) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    Eg[3 * i] = Un[i];
    Eg[3 * i + 1] = Vn[i];
    Eg[3 * i + 2] = T[i];
  }
}

// R[k] = first index i of the id-sorted keys with key >= k, k in [0, nv]
__global__ void offsets_kernel(const unsigned *__restrict__ key, long long n, int nv, int *__restrict__ R) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const long long lo = i == 0 ? 0 : (long long)key[i - 1] + 1;
  const long long hi = i == n ? nv : (long long)key[i];
  for (long long k = lo; k <= hi; k++) R[k] = (int)i;
}

unsigned nblocks(long long n) { return (unsigned)((n + 255) / 256); }

// device buffers of one call, freed on every path
struct Bufs {
  std::vector<void *> p;
  ~Bufs() {
    for (void *x : p) (void)hipFree(x);
  }
  template <class T>
  tm_status get(T **out, size_t count) {
    void *x = nullptr;
    hipError_t e = hipMalloc(&x, sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess) return tmh::set_error(TM_ERR_NOMEM, std::string("graph build: ") + hipGetErrorString(e));
    p.push_back(x);
    *out = (T *)x;
    return TM_OK;
  }
};

#define BCHK(x)                                                                          \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return tmh::set_error(TM_ERR_HIP, std::string("graph build: ") + hipGetErrorString(e_)); \
  } while (0)
#define BST(x)                      \
  do {                              \
    tm_status s_ = (x);             \
    if (s_ != TM_OK) return s_;     \
  } while (0)

// stable sort of (key, value) pairs, n items, into kout / vout
tm_status sort_pairs(Bufs &b, const unsigned *kin, unsigned *kout, const int *vin, int *vout, long long n,
                     hipStream_t s) {
  size_t tmp_bytes = 0;
  BCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, vin, vout, (int)n, 0, 32, s));
  void *tmp = nullptr;
  BST(b.get((char **)&tmp, tmp_bytes));
  BCHK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int)n, 0, 32, s));
  return TM_OK;
}

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

tm_status start(int32_t device, Stream &st) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return tmh::set_error(TM_ERR_NODEVICE, "no GPU visible");
  if (device < 0 || device >= ndev) return tmh::set_error(TM_ERR_ARG, "device out of range");
  BCHK(hipSetDevice(device));
  BCHK(hipStreamCreate(&st.s));
  return TM_OK;
}

}  // namespace

extern "C" {

tm_status tm_time_order(int32_t device, int64_t n, const int32_t *T, int32_t *order) {
  if (n < 0 || (n > 0 && (!T || !order))) return tmh::set_error(TM_ERR_ARG, "bad argument");
  if (n > 0x7fffffffll) return tmh::set_error(TM_ERR_ARG, "edge list too long for int32 positions");
  Stream st;
  BST(start(device, st));
  if (n == 0) return TM_OK;
  Bufs b;
  int *dT, *v0, *v1;
  unsigned *k0, *k1;
  BST(b.get(&dT, n));
  BST(b.get(&v0, n));
  BST(b.get(&v1, n));
  BST(b.get(&k0, n));
  BST(b.get(&k1, n));
  BCHK(hipMemcpyAsync(dT, T, sizeof(int) * n, hipMemcpyHostToDevice, st.s));
  hipLaunchKernelGGL(flip_kernel, dim3(nblocks(n)), dim3(256), 0, st.s, dT, k0, (long long)n);
  hipLaunchKernelGGL(iota_kernel, dim3(nblocks(n)), dim3(256), 0, st.s, v0, (long long)n);
  BCHK(hipGetLastError());
  BST(sort_pairs(b, k0, k1, v0, v1, n, st.s));
  BCHK(hipMemcpyAsync(order, v1, sizeof(int) * n, hipMemcpyDeviceToHost, st.s));
  BCHK(hipStreamSynchronize(st.s));
  return TM_OK;
}

tm_status tm_csr_build(int32_t device, int64_t n, const int32_t *U, const int32_t *V, const int32_t *T, int32_t *Eg,
                       int32_t *outEdgesV, int32_t *inEdgesV, int32_t *outEdgesR, int32_t *inEdgesR,
                       int32_t *vertexIds, int32_t *numVertices) {
  if (n < 0 || !numVertices || !outEdgesR || !inEdgesR ||
      (n > 0 && (!U || !V || !T || !Eg || !outEdgesV || !inEdgesV || !vertexIds)))
    return tmh::set_error(TM_ERR_ARG, "bad argument");
  if (n > 0x3fffffffll) return tmh::set_error(TM_ERR_ARG, "edge list too long for int32 endpoint slots");
  Stream st;
  BST(start(device, st));
  if (n == 0) {
    *numVertices = 0;
    outEdgesR[0] = inEdgesR[0] = 0;
    return TM_OK;
  }
  const long long m = 2 * n;
  Bufs b;
  int *dU, *dV, *dT, *slot0, *slot1, *head, *grp, *rawk, *order, *gidx, *rank, *raw, *Un, *Vn, *dEg, *pos, *P, *R;
  unsigned *key0, *key1, *first, *first1;
  BST(b.get(&dU, n));
  BST(b.get(&dV, n));
  BST(b.get(&dT, n));
  BST(b.get(&key0, m));
  BST(b.get(&key1, m));
  BST(b.get(&slot0, m));
  BST(b.get(&slot1, m));
  BST(b.get(&head, m));
  BST(b.get(&grp, m));
  BCHK(hipMemcpyAsync(dU, U, sizeof(int) * n, hipMemcpyHostToDevice, st.s));
  BCHK(hipMemcpyAsync(dV, V, sizeof(int) * n, hipMemcpyHostToDevice, st.s));
  BCHK(hipMemcpyAsync(dT, T, sizeof(int) * n, hipMemcpyHostToDevice, st.s));
  // ---- renaming
  hipLaunchKernelGGL(interleave_kernel, dim3(nblocks(n)), dim3(256), 0, st.s, dU, dV, (long long)n, key0, slot0);
  BCHK(hipGetLastError());
  BST(sort_pairs(b, key0, key1, slot0, slot1, m, st.s));
  hipLaunchKernelGGL(head_kernel, dim3(nblocks(m)), dim3(256), 0, st.s, key1, m, head);
  BCHK(hipGetLastError());
  {
    size_t tmp_bytes = 0;
    BCHK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, head, grp, (int)m, st.s));
    void *tmp = nullptr;
    BST(b.get((char **)&tmp, tmp_bytes));
    BCHK(hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, head, grp, (int)m, st.s));
  }
  int nv = 0;
  BCHK(hipMemcpyAsync(&nv, grp + (m - 1), sizeof(int), hipMemcpyDeviceToHost, st.s));
  BCHK(hipStreamSynchronize(st.s));
  BST(b.get(&first, nv));
  BST(b.get(&first1, nv));
  BST(b.get(&rawk, nv));
  BST(b.get(&gidx, nv));
  BST(b.get(&order, nv));
  BST(b.get(&rank, nv));
  BST(b.get(&raw, nv));
  hipLaunchKernelGGL(first_kernel, dim3(nblocks(m)), dim3(256), 0, st.s, key1, slot1, head, grp, m, first, rawk);
  hipLaunchKernelGGL(iota_kernel, dim3(nblocks(nv)), dim3(256), 0, st.s, gidx, (long long)nv);
  BCHK(hipGetLastError());
  BST(sort_pairs(b, first, first1, gidx, order, nv, st.s));
  hipLaunchKernelGGL(rank_kernel, dim3(nblocks(nv)), dim3(256), 0, st.s, order, rawk, nv, rank, raw);
  BCHK(hipGetLastError());
  BST(b.get(&Un, n));
  BST(b.get(&Vn, n));
  BST(b.get(&dEg, 3 * n));
  hipLaunchKernelGGL(assign_kernel, dim3(nblocks(m)), dim3(256), 0, st.s, slot1, grp, rank, m, Un, Vn);
  hipLaunchKernelGGL(eg_kernel, dim3(nblocks(n)), dim3(256), 0, st.s, Un, Vn, dT, (long long)n, dEg);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(Eg, dEg, sizeof(int) * 3 * n, hipMemcpyDeviceToHost, st.s));
  BCHK(hipMemcpyAsync(vertexIds, raw, sizeof(int) * nv, hipMemcpyDeviceToHost, st.s));
  // ---- CSR: positions by new source (destination), ascending within each
  BST(b.get(&pos, n));
  BST(b.get(&P, n));
  BST(b.get(&R, (size_t)nv + 1));
  for (int side = 0; side < 2; side++) {
    const int *ids = side == 0 ? Un : Vn;
    hipLaunchKernelGGL(flip_kernel, dim3(nblocks(n)), dim3(256), 0, st.s, ids, key0, (long long)n);  // ids >= 0
    hipLaunchKernelGGL(iota_kernel, dim3(nblocks(n)), dim3(256), 0, st.s, pos, (long long)n);
    BCHK(hipGetLastError());
    BST(sort_pairs(b, key0, key1, pos, P, n, st.s));
    // back to the ids themselves for the offsets (the flip maps id k to k ^ kFlip)
    hipLaunchKernelGGL(flip_kernel, dim3(nblocks(n)), dim3(256), 0, st.s, (const int *)key1, key0, (long long)n);
    hipLaunchKernelGGL(offsets_kernel, dim3(nblocks(n + 1)), dim3(256), 0, st.s, key0, (long long)n, nv, R);
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(side == 0 ? outEdgesV : inEdgesV, P, sizeof(int) * n, hipMemcpyDeviceToHost, st.s));
    BCHK(hipMemcpyAsync(side == 0 ? outEdgesR : inEdgesR, R, sizeof(int) * ((size_t)nv + 1), hipMemcpyDeviceToHost,
                        st.s));
    BCHK(hipStreamSynchronize(st.s));  // key0 / key1 / pos / P / R are reused by the other side
  }
  *numVertices = nv;
  return TM_OK;
}

}  // extern "C"
