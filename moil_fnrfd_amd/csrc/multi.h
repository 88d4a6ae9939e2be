// Single- and multi-system forms of a kernel from one body.
//
// DEFINE_KERNEL(NAME, ARGS, THREADS, BODY) defines
//   NAME(ARGS)                       one system: BODY(args, blockIdx.x, gridDim.x)
//   NAME_multi(tab, bstart, nsys)    many systems in one grid: workgroup g
//                                    belongs to the system s with
//                                    bstart[s] <= g < bstart[s+1] and runs
//                                    BODY(tab[s], g - bstart[s], bstart[s+1] - bstart[s])
// and the host launchers declared in kargs.h.  BODY(const ARGS &, int bid,
// int nblk) must use bid / nblk in place of blockIdx.x / gridDim.x.
#pragma once

#include <hip/hip_runtime.h>

#include "kargs.h"

// largest s with bstart[s] <= g (systems with no workgroups are skipped
// because their offset equals the next system's)
__device__ __forceinline__ int multi_find(const int32_t *__restrict__ bstart, int nsys, int g) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bstart[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define DEFINE_KERNEL(NAME, ARGS, THREADS, BODY)                                               \
  __global__ __launch_bounds__(THREADS) void NAME(ARGS A) { BODY(A, blockIdx.x, gridDim.x); }  \
  __global__ __launch_bounds__(THREADS) void NAME##_multi(                                     \
      const ARGS *__restrict__ tab, const int32_t *__restrict__ bstart, int nsys) {            \
    const int s_ = multi_find(bstart, nsys, blockIdx.x);                                       \
    const int b0_ = bstart[s_];                                                                \
    BODY(tab[s_], (int)blockIdx.x - b0_, bstart[s_ + 1] - b0_);                                \
  }                                                                                            \
  void launch_##NAME(const ARGS &a, int nblk, hipStream_t st) {                                \
    if (nblk > 0) NAME<<<nblk, THREADS, 0, st>>>(a);                                            \
  }                                                                                            \
  void launch_##NAME##_multi(const ARGS *tab, const int32_t *bstart, int nsys, int total,      \
                             hipStream_t st) {                                                 \
    if (total > 0) NAME##_multi<<<total, THREADS, 0, st>>>(tab, bstart, nsys);                  \
  }
