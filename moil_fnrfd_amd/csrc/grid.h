// Cell grid and MOIL image-filter helpers shared by the water and atom
// pair-list builders (lists.hip, atoms.hip).
#pragma once

#include "ctx.h"

struct Grid {
  int gx, gy, gz, nbx, nby, nbz;
  double g0x, g0y, g0z, csx, csy, csz;
};

__host__ __device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline int cell_coord(double p, double g0, double cs, int n) {
  double f = floor((p - g0) / cs);
  if (f < 0.0) return 0;
  if (f > (double)(n - 1)) return n - 1;
  return (int)f;
}

__host__ __device__ inline int cell_key(const Grid &g, int cx, int cy, int cz) {
  int bx = cx >> 1, by = cy >> 1, bz = cz >> 1;
  int loc = (cx & 1) + 2 * (cy & 1) + 4 * (cz & 1);
  return ((bz * g.nby + by) * g.nbx + bx) * 8 + loc;
}

// Box geometry and cutoffs used by the image filters.
struct BoxPar {
  double a, b, c, ha, hb, hc;
  double cutmono2, cutvbig2, cutebig2;
};

// rch2 filter, symm/nbmsym.f:178-189
__device__ inline bool image_i_ok(double cix, double ciy, double ciz, double da, double db,
                                  double dc, double large_cutmono, const BoxPar &P) {
#pragma clang fp contract(off)
  double xch = fabs(cix + da) - P.ha, ych = fabs(ciy + db) - P.hb, zch = fabs(ciz + dc) - P.hc;
  if (xch < 0.0) xch = 0.0;
  if (ych < 0.0) ych = 0.0;
  if (zch < 0.0) zch = 0.0;
  double rch2 = xch * xch + ych * ych + zch * zch;
  return !(rch2 > large_cutmono);
}

// need_only face filter, symm/nbmsym.f:97-115
__device__ inline bool need_only_ok(double cjx, double cjy, double cjz, double da, double db,
                                    double dc, const BoxPar &P) {
#pragma clang fp contract(off)
  double barx = 0.5 * da, bary = 0.5 * db, barz = 0.5 * dc;
  if (fabs(barx) > 0) {
    double q = cjx - barx;
    if (q * q > P.cutmono2) return false;
  }
  if (fabs(bary) > 0) {
    double q = cjy - bary;
    if (q * q > P.cutmono2) return false;
  }
  if (fabs(barz) > 0) {
    double q = cjz - barz;
    if (q * q > P.cutmono2) return false;
  }
  return true;
}


// sort of each range [start[k], start[k+1]) of arr (distinct values), ascending
void launch_cell_sort(const int32_t *start, int ncell, int32_t *arr, hipStream_t st);
