// tm_variants.h -- the mining kernel's instantiations, each behind a plain
// function (defined in the tm_kv_*.hip units, compiled in parallel).
//
// A variant is named by node-map width (4 or 8), enumerate, instrumented (1:
// reference work, 2: metered bytes),
// register cap (1: compiler's choice, 4: 128 VGPRs, 5: 96 VGPRs, 6: 80 VGPRs) and
// single/multi-query form.  occ = true: write the resident workgroups per CU
// for `lds` bytes of dynamic LDS to *blocks; otherwise launch `grid`
// workgroups of 64 threads on `stream`.
#pragma once
#include <hip/hip_runtime.h>

#include "tm_abi.h"
#include "tm_types.h"

namespace tmv {

typedef tm_status (*variant_fn)(bool occ, hipStream_t stream, int lds, int *blocks, const tm_kernel_args *a,
                                int grid);

#define TMV_NAME(N, E, I, W, M) v_##N##_##E##_##I##_##W##_##M
#define TMV_DECL(N, E, I, W, M) \
  tm_status TMV_NAME(N, E, I, W, M)(bool, hipStream_t, int, int *, const tm_kernel_args *, int);

TMV_DECL(4, 0, 0, 1, 0)
TMV_DECL(4, 0, 0, 4, 0)
TMV_DECL(4, 0, 0, 5, 0)
TMV_DECL(4, 0, 0, 6, 0)
TMV_DECL(4, 0, 0, 1, 1)
TMV_DECL(4, 0, 0, 4, 1)
TMV_DECL(4, 0, 0, 5, 1)
TMV_DECL(4, 0, 0, 6, 1)
TMV_DECL(4, 0, 1, 1, 0)
TMV_DECL(4, 1, 0, 1, 0)
TMV_DECL(4, 0, 2, 1, 0)
TMV_DECL(8, 0, 0, 1, 0)
TMV_DECL(8, 0, 0, 4, 0)
TMV_DECL(8, 0, 0, 5, 0)
TMV_DECL(8, 0, 0, 6, 0)
TMV_DECL(8, 0, 0, 1, 1)
TMV_DECL(8, 0, 0, 4, 1)
TMV_DECL(8, 0, 0, 5, 1)
TMV_DECL(8, 0, 0, 6, 1)
TMV_DECL(8, 0, 1, 1, 0)
TMV_DECL(8, 1, 0, 1, 0)
TMV_DECL(8, 0, 2, 1, 0)

// The variant for a launch shape (nmap 4 or 8; instr 0 counting, 1 reference
// work, 2 metered; wpe 1, 4, 5 or 6).
inline variant_fn pick(int nmap, bool enumerate, int instr, int wpe, bool multi) {
  if (nmap == 4) {
    if (instr == 2) return TMV_NAME(4, 0, 2, 1, 0);
    if (instr) return TMV_NAME(4, 0, 1, 1, 0);
    if (enumerate) return TMV_NAME(4, 1, 0, 1, 0);
    if (multi) return wpe == 6 ? TMV_NAME(4, 0, 0, 6, 1) : wpe == 5 ? TMV_NAME(4, 0, 0, 5, 1) : wpe == 4 ? TMV_NAME(4, 0, 0, 4, 1) : TMV_NAME(4, 0, 0, 1, 1);
    return wpe == 6 ? TMV_NAME(4, 0, 0, 6, 0) : wpe == 5 ? TMV_NAME(4, 0, 0, 5, 0) : wpe == 4 ? TMV_NAME(4, 0, 0, 4, 0) : TMV_NAME(4, 0, 0, 1, 0);
  }
  if (instr == 2) return TMV_NAME(8, 0, 2, 1, 0);
  if (instr) return TMV_NAME(8, 0, 1, 1, 0);
  if (enumerate) return TMV_NAME(8, 1, 0, 1, 0);
  if (multi) return wpe == 6 ? TMV_NAME(8, 0, 0, 6, 1) : wpe == 5 ? TMV_NAME(8, 0, 0, 5, 1) : wpe == 4 ? TMV_NAME(8, 0, 0, 4, 1) : TMV_NAME(8, 0, 0, 1, 1);
  return wpe == 6 ? TMV_NAME(8, 0, 0, 6, 0) : wpe == 5 ? TMV_NAME(8, 0, 0, 5, 0) : wpe == 4 ? TMV_NAME(8, 0, 0, 4, 0) : TMV_NAME(8, 0, 0, 1, 0);
}

}  // namespace tmv
