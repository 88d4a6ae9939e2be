// FP32 pair arithmetic shared by the real-space kernels and their adjoint.
//
// erfc(x) = exp(-x^2) t P(t), t = 1/(1+p x): the functional form of
// Abramowitz & Stegun 7.1.26 with a degree-8 polynomial fitted for this
// library (relative error 3e-8 on [0, 6]).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

#define ERFC_P 0.418f
__device__ __forceinline__ float erfc_poly(float t) {
  float p = 7.663338518e-02f;
  p = fmaf(p, t, -3.531459250e-01f);
  p = fmaf(p, t, 5.087678340e-01f);
  p = fmaf(p, t, -1.952929589e-01f);
  p = fmaf(p, t, 3.047273108e-01f);
  p = fmaf(p, t, 1.823138229e-01f);
  p = fmaf(p, t, 2.404413357e-01f);
  p = fmaf(p, t, 2.355552142e-01f);
  return p * t;
}

// round(f * 2^FIX_SHIFT_MIXED) as a 64-bit integer.  For |f| < 2^31 the
// double sum f + 1.5 2^32 has its last mantissa bit at 2^-20, so its bit
// pattern minus that of 1.5 2^32 is the rounded fixed-point value (one
// conversion, one add, one 64-bit subtract); larger terms take the full
// double-to-integer conversion.  A term beyond the accumulators' per-term
// range (FIX_TERM_UNITS, common.h) or a NaN sets `bad`, which the thread
// reports once with fix_report after its conversions (an atomic on the
// overflow word in the middle of a pair loop costs the loop registers).
static_assert(FIX_SHIFT_MIXED == 20, "fixf's bias constant is that of 2^-20 units");
#define FIX_BIAS20 0x41F8000000000000ull  // the bit pattern of 1.5 2^32
__device__ __forceinline__ long long fixd(double v, double s, bool &bad);
__device__ __forceinline__ long long fixf(float f, bool &bad) {
  if (fabsf(f) < 2147483648.0f)
    return (long long)((unsigned long long)__double_as_longlong((double)f + 6442450944.0) -
                       FIX_BIAS20);
  return fixd((double)f, 1048576.0, bad);
}

// round(v s), s = 2^FIX_SHIFT, for the FP64 conversions, with the same
// range check
__device__ __forceinline__ long long fixd(double v, double s, bool &bad) {
  const double x = v * s;
  bad |= !(fabs(x) < FIX_TERM_UNITS);
  return __double2ll_rn(x);
}

// Sums of four per-thread FP64 values over a workgroup of four wavefronts
// in a fixed order: each wavefront folds its 64 lanes by cross-lane swaps
// (xor 32, 16, ..., 1), lane 0 of each wavefront stores its four sums in
// ws[16], and after one barrier thread k < 4 adds the wavefronts' sums of
// value k as (w0 + w1) + (w2 + w3) and returns it (other threads get 0).
// ws may be reused after the caller's next barrier.
__device__ __forceinline__ double wg_sum4(double v0, double v1, double v2, double v3,
                                          double *ws) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
    v2 += __shfl_xor(v2, off, 64);
    v3 += __shfl_xor(v3, off, 64);
  }
  if (lane == 0) {
    ws[wv] = v0;
    ws[4 + wv] = v1;
    ws[8 + wv] = v2;
    ws[12 + wv] = v3;
  }
  __syncthreads();
  return t < 4 ? (ws[4 * t] + ws[4 * t + 1]) + (ws[4 * t + 2] + ws[4 * t + 3]) : 0.0;
}

// sets the context's overflow word if this thread converted a term out of
// range
__device__ __forceinline__ void fix_report(bool bad, int *ovf) {
  if (bad) atomicOr(ovf, 1);
}

// Each lane walks its item's block mask starting at a member given by the
// item's index (bit b of the rotated mask is block member (b + rot) mod
// 64): the consecutive items of a wavefront start at different members,
// so lanes with similar masks do not all add into the same LDS
// accumulators at once, and an item's sums run in the same order whichever
// lane or chunk takes it.
__device__ __forceinline__ uint64_t rotr64(uint64_t x, int r) {
  return r ? ((x >> r) | (x << (64 - r))) : x;
}

// MOIL's cubic-spline erfc table (pot/cdie_ewald.f:152-160): erfc and its
// derivative at x = beta r from row int(erftbdns x).  MOIL reads that row
// without a bound; its lists keep x inside the table (1.5 relx long) while
// molecules are whole and atoms moved less than the list skin.  Beyond the
// last of the ntab rows (stale lists, a molecule split across the box, a NaN)
// the last row is read at its right end, so no read leaves the table.
__device__ __forceinline__ void erfc_table(const double *__restrict__ erf_arr, double erftbdns,
                                           int ntab, double x, double &erfcc, double &derfc) {
  int idx;
  double dx;
  if (erftbdns * x < (double)ntab) {
    idx = (int)(erftbdns * x);
    dx = x - idx * (1.0 / erftbdns);
  } else {
    idx = ntab - 1;
    dx = 1.0 / erftbdns;
  }
  const double *tb = erf_arr + 4 * idx;
  derfc = -tb[1] - dx * (tb[2] + 0.5 * dx * tb[3]);
  erfcc = dx * (tb[1] + 0.5 * dx * (tb[2] + dx * tb[3] * (1.0 / 3.0))) + tb[0];
}
