// Shared device/host definitions of the MI355X MOIL energy library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "moil_gpu.h"

#define KOFDIE 332.0716  // COMMON/CONVERT.BLOCK:13

// Water-block geometry: 64 waters (one wavefront's worth of mask bits).
#define WBLOCK 64
// A (j, s) entry of a block is split into items of at most WSPLIT pairs
// (consecutive mask bits), so one lane's loop is at most WSPLIT pairs
// long; a block's items are grouped by exact popcount, descending
// (bucket k holds the items with popcount WSPLIT - k).  32 measured best
// on the 98k-atom water box (16: 6 % slower real-space kernel, no split:
// 11 % slower).
#ifndef WSPLIT
#define WSPLIT 32
#endif
#define NBUCKET WSPLIT
// water lists: the listwt1 and the listwt2 items in buckets of their own
// (lists.hip emit_kind)
#define WNBUCKET (2 * WSPLIT)

// Fixed-point force accumulators: int64 with FIX_SHIFT fractional bits.
// Every term converted to fixed point must stay below 2^59 units (a
// sixteenth of the int64 range, so sums of a few such terms still fit):
// |term| < 2^23 kcal/mol/A in FP64 mode (resolution 1.5e-11), < 2^39 in
// mixed mode (resolution 9.5e-7, as fine as the FP32 pair arithmetic for
// forces above ~10).  A term beyond it sets the context's overflow word,
// and the evaluation's energies come back as NaN.
#define FIX_SHIFT_FP64 36
#define FIX_SHIFT_MIXED 20
#define FIX_TERM_UNITS 576460752303423488.0  // 2^59

// Energy accumulator slots (per-workgroup partial sums).
enum {
  EN_BOND = 0, EN_THETA, EN_TORS, EN_IMP, EN_VDW, EN_EL, EN_VSYM, EN_LSYM,
  EN_EL14, EN_VDW14, EN_CORR, EN_RECIP, EN_COUNT
};

// One water-pair work item: water `j` (slot index, the untranslated
// member), MOIL symmetry operation `s` (0 = real pair, 1..13 = the i-side
// water is translated by symop s), and masks over the 64 i-waters of the
// owning block: m1 = listwt1 pairs, m2 = listwt2 pairs.
struct __attribute__((aligned(32))) WItem {
  int32_t j;
  int32_t s;
  uint64_t m1;
  uint64_t m2;
  uint64_t pad;
};

// Candidate record of the water-list builder: partner slot j, symmetry
// operation s and the unsplit listwt1/listwt2 masks over the block.
struct __attribute__((aligned(8))) WRaw {
  int32_t j;
  int32_t s;
  uint64_t m1;
  uint64_t m2;
};

// Atom-pair work item over a 64-atom solute block: partner atom `j`
// (0-based), direction `s` (0 real, 1..13 the block atom is the image
// under symop s, 14..26 the partner is the image under symop s-13), and
// the block atoms' memberships in MOIL's list1/list2/list3.
struct __attribute__((aligned(32))) AItem {
  int32_t j;
  int32_t s;
  uint64_t m1;
  uint64_t m2;
  uint64_t m3;
};

#define HIP_CHECK(x)                                                         \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                __FILE__ + ":" + std::to_string(__LINE__));                  \
      return MOIL_GPU_EHIP;                                                  \
    }                                                                        \
  } while (0)

void set_error(const std::string &msg);

// an atom-list item's direction d: 0 real, 1..13 the block atom is the
// image under symmetry operation s = d, 14..26 the partner is the image
// under s = d - 13 (rev)
__host__ __device__ inline void dir_of(int d, int &s, bool &rev) {
  rev = d > 13;
  s = rev ? d - 13 : d;
}

// MOIL's 13 half-shell symmetry operations in its order
// (symm/syminit.f:80-164); s = 0 is the untranslated (real) cell.
__host__ __device__ inline void symop_of(int s, int *ia, int *jb, int *kc) {
  *ia = (s == 0 || s == 4 || (s >= 8 && s <= 10)) ? 0 : 1;
  *jb = (s <= 4) ? 0 : (s <= 10 ? 1 : -1);
  int k0 = (s >= 11) ? 11 : (s >= 8) ? 8 : (s >= 5) ? 5 : 1;
  *kc = (s == 0) ? 0 : (s == 4) ? 1 : 1 - (s - k0);
}

// LJ coefficients A (of r^-12) and B (of r^-6) of a pair from the i atom's
// factors (ai = epsgm12_i pick, bi) and the partner's (aj, bj).  Geometric
// rule: A = ai aj, B = bi bj.  Arith rule (epsgm6 holds sigma):
// a = ai aj, sigma = (bi + bj)/2, B = a sigma^6, A = a sigma^12 with
// MOIL's products (pot/cdie_ewald.f:377-383,461-470); A and B then enter
// A s6 s6 and B s6 as in the geometric case.
__host__ __device__ inline void lj_coeffs(int arith, double ai, double bi, double aj, double bj,
                                          double &A, double &B) {
  if (!arith) {
    A = ai * aj;
    B = bi * bj;
    return;
  }
  const double a = ai * aj, b = 0.5 * (bi + bj);
  double b1 = b * b * b;
  b1 = b1 * b1;
  B = a * b1;
  b1 = b1 * b1;
  A = a * b1;
}
