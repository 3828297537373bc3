// worker_csr.hip -- a hand-written worker plugin (libCSR.so) in the style of
// the reference's legacy workers: it binds only to the reference's
// DeviceJobData fields (Eg_d, minfo()[i].arrR / arrV, Em_d) and counts with a
// plain one-thread-per-root depth-first search over the CSR lists.
//
// It is the check that code written against the reference's include/data.h
// runs on this core unchanged in shape: getWorker entry point
// (system/templates/GPUWorker/interface.snp:1-7), cumulative count
// (GPUWorker/ctor.snp:3-4: the device counter is zeroed only at
// construction), update_job setting beg = end (GPUWorker/update.snp:1-3).
// The search follows the MotifEdgeInfoV1 pre-decode
// (src/core/cpu/data.cc:422-447): edge i (i >= 1) scans the list of the graph
// node mapped to baseNode (out-list when io == 1, in-list otherwise) after
// the previous edge's position, stops at the first edge past t0 + delta,
// maps the new endpoint to node -constraintNode if it is not mapped yet, or
// for io == -1 requires the source to be the node mapped to constraintNode.
// It is not the library's fast path (that is tm_count behind libTM.so).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "everest/data.h"

using namespace corelib;

namespace {

constexpr int kMaxE = TM_MAX_MOTIF_EDGES;

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("libCSR ") + what + ": " + hipGetErrorString(e));
}

// first slot s in [lo, hi) of the position list v with v[s] > p
__device__ int first_after(const int *__restrict__ v, int lo, int hi, int p) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v[mid] <= p) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256) csr_count_kernel(const TemporalEdge *__restrict__ Eg,
                                                        const MotifEdgeInfoV1 *__restrict__ minfo, int M, int beg,
                                                        int end, int delta, unsigned long long *count) {
  __shared__ unsigned long long part[256];
  const int r = beg + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  unsigned long long c = 0;
  if (r < end) {
    const TemporalEdge e0 = Eg[r];
    if (e0.u != e0.v) {
      int m[kMaxE + 1];  // motif node -> graph node
      int cur[kMaxE], hi[kMaxE], pos[kMaxE];
      m[0] = e0.u;
      m[1] = e0.v;
      pos[0] = r;
      const long long tl = (long long)e0.t + delta;
      int lvl = 1;
      {
        const MotifEdgeInfoV1 mi = minfo[0];
        const int x = m[mi.baseNode];
        hi[1] = mi.arrR[x + 1];
        cur[1] = first_after(mi.arrV, mi.arrR[x], hi[1], r);
      }
      while (lvl >= 1) {
        if (cur[lvl] >= hi[lvl]) { lvl--; continue; }
        const MotifEdgeInfoV1 mi = minfo[lvl - 1];
        const int p = mi.arrV[cur[lvl]++];
        const TemporalEdge e = Eg[p];
        if ((long long)e.t > tl) { lvl--; continue; }  // the list is in time order: the rest is outside
        bool ok = true;
        if (mi.newNode) {
          const int node = mi.io == 1 ? e.v : e.u;
          for (int k = 0; k < mi.mappedNodes; k++) ok = ok && m[k] != node;
          if (ok) m[-mi.constraintNode] = node;
        } else {
          ok = e.u == m[mi.constraintNode];  // in-list of m[baseNode]: the source must match
        }
        if (!ok) continue;
        if (lvl == M - 1) { c++; continue; }
        pos[lvl] = p;
        lvl++;
        const MotifEdgeInfoV1 nx = minfo[lvl - 1];
        const int x = m[nx.baseNode];
        hi[lvl] = nx.arrR[x + 1];
        cur[lvl] = first_after(nx.arrV, nx.arrR[x], hi[lvl], p);
      }
    }
  }
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && part[0]) atomicAdd(count, part[0]);
}

struct GPUWorkerCSR : public GPUWorker {
  unsigned long long *count_d = nullptr;

  explicit GPUWorkerCSR(int gpu) : GPUWorker("CSR", gpu, 256) {
    check(hipSetDevice(gpu), "hipSetDevice");
    check(hipMalloc((void **)&count_d, sizeof(unsigned long long)), "hipMalloc");
    check(hipMemset(count_d, 0, sizeof(unsigned long long)), "hipMemset");
  }
  ~GPUWorkerCSR() override {
    if (count_d) (void)hipFree(count_d);
  }

  unsigned long long run() override {
    if (!job_ || !job_->data) throw std::runtime_error("libCSR: run without a job");
    const data::DeviceJobData *d = job_->data;
    if (!d->Eg_d || !d->minfo()) throw std::runtime_error("libCSR: the job's graph is not on the device");
    const int M = (int)d->motifNumEdges;
    if (M < 2 || M > kMaxE) throw std::runtime_error("libCSR: motif needs 2..8 edges");
    check(hipSetDevice(gpu_), "hipSetDevice");
    const int n = job_->end - job_->beg;
    if (n > 0) {
      hipLaunchKernelGGL(csr_count_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d->Eg_d, d->minfo(), M,
                         job_->beg, job_->end, job_->delta, count_d);
      check(hipGetLastError(), "launch");
    }
    check(hipMemcpy(&count_, count_d, sizeof(count_), hipMemcpyDeviceToHost), "count copy");
    return count_;
  }

  void update_job() override { job_->beg = job_->end; }
};

}  // namespace

extern "C" GPUWorker *getWorker(int gpu) { return new GPUWorkerCSR(gpu); }
