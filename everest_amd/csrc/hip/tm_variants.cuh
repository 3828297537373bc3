// tm_variants.cuh -- definition of one mining-kernel variant function
// (included by the tm_kv_*.hip units only).
#pragma once
#include <string>

#include "tm_internal.h"
#include "tm_kernel.cuh"
#include "tm_variants.h"

#define TMV_DEF(N, E, I, W, M)                                                                              \
  tm_status tmv::TMV_NAME(N, E, I, W, M)(bool occ, hipStream_t stream, int lds, int *blocks,                \
                                         const tm_kernel_args *a, int grid) {                               \
    auto k = tmk::mine_kernel<TM_MAXE, N, (E) != 0, (I), tmk::DynSpec, W, (M) != 0>;                   \
    hipError_t e;                                                                                           \
    if (occ) {                                                                                              \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, 64, lds);                                 \
    } else {                                                                                                \
      hipLaunchKernelGGL(k, dim3(grid), dim3(64), lds, stream, *a);                                         \
      e = hipGetLastError();                                                                                \
    }                                                                                                       \
    if (e != hipSuccess)                                                                                    \
      return tmh::set_error(TM_ERR_HIP, std::string(occ ? "occupancy query: " : "kernel launch: ") +        \
                                            hipGetErrorString(e));                                          \
    return TM_OK;                                                                                           \
  }
