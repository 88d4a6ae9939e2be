// Buffered pair-list builder for MOIL's water-group lists on gfx950.
//
// Reproduces the pair set of nbondm (pot/nbondm.f:79-199, real water
// pairs) and nbmsym (symm/nbmsym.f:59-233, the 13 half-shell image
// translations) bit for bit: every predicate is evaluated in FP64 in
// MOIL's operation order with FMA contraction disabled, including the
// partial-distance early exits against cutmono2, the monomer-centre
// "need_only" face filter, the rch2 box-distance filter with its
// large_cutmono switch, and `<=` (real) versus `<` (image) comparisons.
//
// Layout: waters are binned by their centre (the O atom) into a cell grid
// over the box; cells are numbered brick-major (2x2x2-cell bricks) so that
// 64 consecutive waters in cell order (a "block") are spatially compact.
// For every block, one workgroup enumerates candidate partners cell by
// cell and emits one item per partner water j and symmetry operation s
// with a 64-bit mask over the block's waters (the translated member).
// Items are compacted with wavefront ballots/prefix sums in a fixed order
// (bit-reproducible) and grouped by descending popcount so the real-space
// kernel's lanes see similar loop trip counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "grid.h"

#define LB_THREADS 256
#define LB_MAXCELLS 1024

// ---------------------------------------------------------------------------
// monomer centres, pot/nbondm.f:79-124 (water: the atom two before the
// monomer's last atom; otherwise the unweighted mean times 1/n)
__global__ void k_mono_cent(const double *__restrict__ coor, const int32_t *__restrict__ poipt,
                            const int32_t *__restrict__ monoflag, int totmon, double ha,
                            double hb, double hc, double *__restrict__ cent,
                            int32_t *__restrict__ flags) {
#pragma clang fp contract(off)
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= totmon) return;
  double cx, cy, cz;
  if (monoflag[m] & MOIL_MONO_WATER) {
    int j = poipt[m + 1] - 2 - 1;
    cx = coor[3 * j];
    cy = coor[3 * j + 1];
    cz = coor[3 * j + 2];
  } else {
    int ls = poipt[m], le = poipt[m + 1];
    cx = cy = cz = 0.0;
    for (int j = ls; j < le; j++) {
      cx = cx + coor[3 * j];
      cy = cy + coor[3 * j + 1];
      cz = cz + coor[3 * j + 2];
    }
    double r2 = 1.0 / (double)(le - ls);
    cx = cx * r2;
    cy = cy * r2;
    cz = cz * r2;
  }
  cent[3 * m] = cx;
  cent[3 * m + 1] = cy;
  cent[3 * m + 2] = cz;
  // symm/nbmsym.f:87-95: any centre outside the box quadruples cutmono2
  if (fabs(cx) > ha || fabs(cy) > hb || fabs(cz) > hc) atomicOr(&flags[0], 1);
}

__global__ void k_wat_cells(const double *__restrict__ cent, const int32_t *__restrict__ wat_mono,
                            int nwat, Grid g, int32_t *__restrict__ cell_of,
                            int32_t *__restrict__ cnt) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwat) return;
  const double *c = cent + 3 * (wat_mono[w] - 1);
  int cx = cell_coord(c[0], g.g0x, g.csx, g.gx);
  int cy = cell_coord(c[1], g.g0y, g.csy, g.gy);
  int cz = cell_coord(c[2], g.g0z, g.csz, g.gz);
  int k = cell_key(g, cx, cy, cz);
  cell_of[w] = k;
  atomicAdd(&cnt[k], 1);
}

// Exclusive scan of n int32 values into out[0..n] (out[n] = total), any n:
// 2048-element tiles per workgroup (coalesced loads through LDS, eight
// values per thread, wavefront shuffles for the 256 thread sums), the tile
// totals scanned the same way, then the tile offsets added.  Integer
// arithmetic, so the result does not depend on the decomposition.
#define SCAN_TB 256
#define SCAN_PER 8
#define SCAN_TILE (SCAN_TB * SCAN_PER)

__global__ __launch_bounds__(SCAN_TB) void k_scan_tile(const int32_t *__restrict__ in,
                                                       int32_t *__restrict__ out, int n,
                                                       int32_t *__restrict__ tot) {
  __shared__ int32_t v[SCAN_TILE];
  __shared__ int32_t ws[SCAN_TB / 64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
#pragma unroll
  for (int k = 0; k < SCAN_PER; k++) {
    const size_t i = base + k * SCAN_TB + t;
    v[k * SCAN_TB + t] = i < (size_t)n ? in[i] : 0;
  }
  __syncthreads();
  int pre[SCAN_PER], sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; k++) {
    pre[k] = sum;
    sum += v[t * SCAN_PER + k];
  }
  int incl = sum;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) ws[wv] = incl;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wv; w++) wbase += ws[w];
  const int ex = wbase + incl - sum;
#pragma unroll
  for (int k = 0; k < SCAN_PER; k++) v[t * SCAN_PER + k] = ex + pre[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_PER; k++) {
    const size_t i = base + k * SCAN_TB + t;
    if (i < (size_t)n) out[i] = v[k * SCAN_TB + t];
  }
  if (t == SCAN_TB - 1) tot[blockIdx.x] = wbase + incl;
}

__global__ __launch_bounds__(SCAN_TB) void k_scan_add(int32_t *__restrict__ out, int n,
                                                      const int32_t *__restrict__ toff, int ntile) {
  const size_t i = (size_t)blockIdx.x * SCAN_TB + threadIdx.x;
  if (i < (size_t)n) out[i] += toff[i / SCAN_TILE];
  if (i == 0) out[n] = toff[ntile];
}

static size_t scan_ws_need(int n) {
  const size_t nt = ((size_t)n + SCAN_TILE - 1) / SCAN_TILE;
  if (nt <= 1) return 0;
  return nt + (nt + 1) + scan_ws_need((int)nt);
}

static void scan_rec(const int32_t *in, int32_t *out, int n, int32_t *ws, hipStream_t st) {
  const int nt = (int)(((size_t)n + SCAN_TILE - 1) / SCAN_TILE);
  if (nt <= 1) {
    k_scan_tile<<<1, SCAN_TB, 0, st>>>(in, out, n, out + n);
    return;
  }
  int32_t *tot = ws, *toff = ws + nt;
  k_scan_tile<<<nt, SCAN_TB, 0, st>>>(in, out, n, tot);
  scan_rec(tot, toff, nt, ws + nt + (nt + 1), st);
  k_scan_add<<<(unsigned)(((size_t)n + SCAN_TB - 1) / SCAN_TB), SCAN_TB, 0, st>>>(out, n, toff, nt);
}

int32_t scan_excl(moil_ctx *c, const int32_t *in, int32_t *out, int n, hipStream_t st) {
  const size_t need = scan_ws_need(n);
  if (need > c->scan_ws_cap) {
    HIP_CHECK(hipStreamSynchronize(st));
    hipFree(c->d_scan_ws);
    c->d_scan_ws = nullptr;
    c->scan_ws_cap = 0;
    HIP_CHECK(hipMalloc(&c->d_scan_ws, sizeof(int32_t) * (need + need / 2 + 64)));
    c->scan_ws_cap = need + need / 2 + 64;
  }
  scan_rec(in, out, n, c->d_scan_ws, st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

__global__ void k_wat_scatter(const int32_t *__restrict__ cell_of, int nwat,
                              const int32_t *__restrict__ start, int32_t *__restrict__ fill,
                              int32_t *__restrict__ slot2wat) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwat) return;
  int k = cell_of[w];
  int p = start[k] + atomicAdd(&fill[k], 1);
  slot2wat[p] = w;
}

// sort each cell's entries (distinct indices) ascending so the slot order
// is deterministic: one wavefront per cell, the entries staged in LDS and
// each written at its rank (the number of smaller entries); a cell of more
// than CS_MAX entries is sorted by insertion in one lane
#define CS_MAX 512
__global__ __launch_bounds__(64) void k_cell_sort(const int32_t *__restrict__ start, int ncell,
                                                  int32_t *__restrict__ s2w) {
  __shared__ int32_t v[CS_MAX];
  const int k = blockIdx.x, t = threadIdx.x;
  if (k >= ncell) return;
  const int b = start[k], n = start[k + 1] - b;
  if (n <= 1) return;
  if (n > CS_MAX) {
    if (t == 0)
      for (int i = b + 1; i < b + n; i++) {
        const int x = s2w[i];
        int j = i - 1;
        while (j >= b && s2w[j] > x) {
          s2w[j + 1] = s2w[j];
          j--;
        }
        s2w[j + 1] = x;
      }
    return;
  }
  for (int i = t; i < n; i += 64) v[i] = s2w[b + i];
  __syncthreads();
  for (int i = t; i < n; i += 64) {
    const int x = v[i];
    int r = 0;
    for (int j = 0; j < n; j++) r += v[j] < x ? 1 : 0;
    s2w[b + r] = x;
  }
}

void launch_cell_sort(const int32_t *start, int ncell, int32_t *arr, hipStream_t st) {
  if (ncell > 0) k_cell_sort<<<ncell, 64, 0, st>>>(start, ncell, arr);
}

__global__ void k_slot_fill(const int32_t *__restrict__ s2w, const int32_t *__restrict__ wat_mono,
                            const int32_t *__restrict__ wat_o, const double *__restrict__ cent,
                            int nwat, double *__restrict__ scent, int32_t *__restrict__ smono,
                            int32_t *__restrict__ slot_o) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nwat) return;
  int w = s2w[s];
  int m = wat_mono[w];
  smono[s] = m;
  slot_o[s] = wat_o[w];
  scent[3 * s] = cent[3 * (m - 1)];
  scent[3 * s + 1] = cent[3 * (m - 1) + 1];
  scent[3 * s + 2] = cent[3 * (m - 1) + 2];
}

struct WPar : BoxPar {
  double R;
  int nwat;
};

// MOIL's water-pair decision for (i translated by (da,db,dc), j).
// Returns 0 (no pair), 1 (listwt1) or 2 (listwt2).
// Real pairs: pot/nbondm.f:176-199; images: symm/nbmsym.f:197-226.
__device__ inline int wpair_kind(double cix, double ciy, double ciz, double da, double db,
                                 double dc, double cjx, double cjy, double cjz, bool image,
                                 const WPar &P) {
#pragma clang fp contract(off)
  double rx, r2;
  rx = (cix + da) - cjx;
  r2 = rx * rx;
  if (r2 > P.cutmono2) return 0;
  rx = (ciy + db) - cjy;
  r2 = r2 + rx * rx;
  if (r2 > P.cutmono2) return 0;
  rx = (ciz + dc) - cjz;
  r2 = r2 + rx * rx;
  if (!image) {
    if (r2 <= P.cutvbig2) return 1;
    if (r2 <= P.cutebig2) return 2;
  } else {
    if (r2 < P.cutvbig2) return 1;
    if (r2 < P.cutebig2) return 2;
  }
  return 0;
}

// removes the n lowest set bits of m and returns them
__device__ inline uint64_t take_low_bits(uint64_t &m, int n) {
  uint64_t out = 0;
  for (int k = 0; k < n; k++) {
    const uint64_t low = m & (~m + 1ull);
    out |= low;
    m ^= low;
  }
  return out;
}

// Deterministic emission of one batch of candidates (one per thread of
// the workgroup, masks 0 = none).  The two lists stay apart: the listwt1
// members (m1) and the listwt2 members (m2) of a candidate become items of
// their own kind, so the real-space kernel's lanes in one wavefront walk
// masks of one kind and equal length (listwt2 members take part only
// within relx, a cheap test, while every listwt1 member costs a full
// entry).  Each kind's mask is split into nfull items of WSPLIT pairs
// (bucket base + 0) and at most one remainder item (bucket base + WSPLIT -
// popcount remainder), base = 0 for listwt1 and WSPLIT for listwt2;
// per-bucket positions come from wavefront scans and ballots in thread
// order, so the item order is fixed.  Every thread of the workgroup calls
// it.
struct EmitLDS {
  int32_t bk_run[WNBUCKET];
  int32_t wcnt[LB_THREADS / 64][WSPLIT];
};

__device__ inline void emit_kind(EmitLDS &L, int kind, int js, int s, uint64_t mk,
                                 const int32_t *__restrict__ boff_b, WItem *__restrict__ items,
                                 size_t cap) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int base = kind * WSPLIT;
  const int pc = __popcll(mk);
  const int nfull = pc / WSPLIT, rem = pc % WSPLIT;
  const int bk = rem ? WSPLIT - rem : -1;
  int incl = nfull;
  for (int off = 1; off < 64; off <<= 1) {
    int v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  const int wtot0 = __shfl(incl, 63, 64);
  if (lane == 0) L.wcnt[wv][0] = wtot0;
  for (int k = 1; k < WSPLIT; k++) {
    uint64_t bal = __ballot(bk == k);
    if (lane == 0) L.wcnt[wv][k] = __popcll(bal);
  }
  __syncthreads();
  {
    uint64_t rest = mk;  // bits left for the remainder item
    if (nfull > 0) {
      int before0 = 0;
      for (int w2 = 0; w2 < wv; w2++) before0 += L.wcnt[w2][0];
      const int pos0 = boff_b[base] + L.bk_run[base] + before0 + (incl - nfull);
      for (int q = 0; q < nfull; q++) {
        const uint64_t piece = take_low_bits(rest, WSPLIT);
        WItem it;
        it.j = js;
        it.s = s;
        it.m1 = kind ? 0ull : piece;
        it.m2 = kind ? piece : 0ull;
        it.pad = 0;
        if ((size_t)(pos0 + q) < cap) items[pos0 + q] = it;
      }
    }
    for (int k = 1; k < WSPLIT; k++) {
      uint64_t bal = __ballot(bk == k);
      if (bk == k) {
        int before = 0;
        for (int w2 = 0; w2 < wv; w2++) before += L.wcnt[w2][k];
        uint64_t lower = (lane == 0) ? 0ull : (bal & ((1ull << lane) - 1ull));
        const int pos = boff_b[base + k] + L.bk_run[base + k] + before + __popcll(lower);
        WItem it;
        it.j = js;
        it.s = s;
        it.m1 = kind ? 0ull : rest;
        it.m2 = kind ? rest : 0ull;
        it.pad = 0;
        if ((size_t)pos < cap) items[pos] = it;
      }
    }
  }
  __syncthreads();
  if (t < WSPLIT) {
    int tot = 0;
    for (int w2 = 0; w2 < LB_THREADS / 64; w2++) tot += L.wcnt[w2][t];
    L.bk_run[base + t] += tot;
  }
  __syncthreads();
}

__device__ inline void emit_batch(EmitLDS &L, int js, int s, uint64_t m1, uint64_t m2,
                                  const int32_t *__restrict__ boff_b, WItem *__restrict__ items,
                                  size_t cap) {
  // a kind no thread of the batch holds is skipped (uniform over the
  // workgroup)
  if (__syncthreads_or(m1 != 0)) emit_kind(L, 0, js, s, m1, boff_b, items, cap);
  if (__syncthreads_or(m2 != 0)) emit_kind(L, 1, js, s, m2, boff_b, items, cap);
}

// items of one kind's mask per bucket (counts only: order independent)
__device__ inline void count_kind(EmitLDS &L, int kind, uint64_t mk) {
  const int pc = __popcll(mk);
  if (pc >= WSPLIT) atomicAdd(&L.bk_run[kind * WSPLIT], pc / WSPLIT);
  if (pc % WSPLIT) atomicAdd(&L.bk_run[kind * WSPLIT + WSPLIT - pc % WSPLIT], 1);
}

// One workgroup per water block: evaluates MOIL's predicates for every
// candidate partner once, keeps the non-empty (j, s, masks) records in
// candidate order in the block's slice of `raw` (capacity rcap; the count
// is exact even when the slice overflows) and counts the items per bucket.
__global__ __launch_bounds__(LB_THREADS) void k_wblock_items(
    const double *__restrict__ scent, const int32_t *__restrict__ smono,
    const int32_t *__restrict__ cstart, Grid g, WPar P, const int32_t *__restrict__ flags,
    int32_t *__restrict__ bcount, WRaw *__restrict__ raw, int rcap, int32_t *__restrict__ rcount,
    int32_t *__restrict__ rneed, int npart) {
  __shared__ double ic[WBLOCK][3];
  __shared__ int32_t im[WBLOCK];
  __shared__ double bb[6];
  __shared__ double sbb[WBLOCK / 8][6];  // bounding boxes of the 8-water sub-blocks
  __shared__ uint64_t imask;
  __shared__ int32_t cellk[LB_MAXCELLS];
  __shared__ int32_t cellpre[LB_MAXCELLS + 1];
  __shared__ int32_t scanbuf[LB_THREADS];
  __shared__ EmitLDS L;
  __shared__ int32_t rw[LB_THREADS / 64];
  __shared__ int32_t r_run;

  // workgroup = (block, part): the part takes every npart-th batch of
  // candidates of its block (batches numbered over translations and cells)
  const int b = blockIdx.x / npart, part = blockIdx.x - b * npart;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int s0 = b * WBLOCK;
  const int nb = min(WBLOCK, P.nwat - s0);
  int batch = 0;
  if (t < WBLOCK) {
    if (t < nb) {
      ic[t][0] = scent[3 * (s0 + t)];
      ic[t][1] = scent[3 * (s0 + t) + 1];
      ic[t][2] = scent[3 * (s0 + t) + 2];
      im[t] = smono[s0 + t];
    } else {
      im[t] = 0x7fffffff;
    }
  }
  if (t < WNBUCKET) L.bk_run[t] = 0;
  if (t == 0) r_run = 0;
  __syncthreads();
  if (t == 0) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < nb; i++)
      for (int k = 0; k < 3; k++) {
        lo[k] = fmin(lo[k], ic[i][k]);
        hi[k] = fmax(hi[k], ic[i][k]);
      }
    for (int k = 0; k < 3; k++) {
      bb[k] = lo[k];
      bb[3 + k] = hi[k];
    }
  } else if (t >= 64 && t < 64 + WBLOCK / 8) {
    // sub-block u: block members 8u..8u+7 (consecutive slots in cell order,
    // about one cell); an empty one gets a box no candidate is near
    const int u = t - 64;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 8 * u; i < min(8 * u + 8, nb); i++)
      for (int k = 0; k < 3; k++) {
        lo[k] = fmin(lo[k], ic[i][k]);
        hi[k] = fmax(hi[k], ic[i][k]);
      }
    for (int k = 0; k < 3; k++) {
      sbb[u][k] = lo[k];
      sbb[u][3 + k] = hi[k];
    }
  }
  __syncthreads();
  const double large_cutmono = flags[0] ? P.cutmono2 * 4.0 : P.cutmono2;
  const double R2 = P.R * P.R;
  WRaw *rblk = raw + (size_t)blockIdx.x * rcap;

  for (int s = 0; s < 14; s++) {
    int ia, jb, kc;
    symop_of(s, &ia, &jb, &kc);
    const double da = ia * P.a, db = jb * P.b, dc = kc * P.c;
    const bool image = s > 0;
    // i-side filter for this translation
    if (t < 64) {
      bool ok = false;
      if (t < nb) ok = !image || image_i_ok(ic[t][0], ic[t][1], ic[t][2], da, db, dc, large_cutmono, P);
      uint64_t bal = __ballot(ok);
      if (t == 0) imask = bal;
    }
    __syncthreads();
    const uint64_t ivalid = imask;
    if (ivalid == 0) {
      __syncthreads();
      continue;
    }
    // query box: translated block bbox grown by R
    double q0[3], q1[3];
    q0[0] = bb[0] + da - P.R; q1[0] = bb[3] + da + P.R;
    q0[1] = bb[1] + db - P.R; q1[1] = bb[4] + db + P.R;
    q0[2] = bb[2] + dc - P.R; q1[2] = bb[5] + dc + P.R;
    int cx0 = clampi((int)floor((q0[0] - g.g0x) / g.csx), 0, g.gx - 1);
    int cx1 = clampi((int)floor((q1[0] - g.g0x) / g.csx), 0, g.gx - 1);
    int cy0 = clampi((int)floor((q0[1] - g.g0y) / g.csy), 0, g.gy - 1);
    int cy1 = clampi((int)floor((q1[1] - g.g0y) / g.csy), 0, g.gy - 1);
    int cz0 = clampi((int)floor((q0[2] - g.g0z) / g.csz), 0, g.gz - 1);
    int cz1 = clampi((int)floor((q1[2] - g.g0z) / g.csz), 0, g.gz - 1);
    const int nx = cx1 - cx0 + 1, ny = cy1 - cy0 + 1, nz = cz1 - cz0 + 1;
    const int ncq = nx * ny * nz;
    for (int cbeg = 0; cbeg < ncq; cbeg += LB_MAXCELLS) {
      const int ncc = min(LB_MAXCELLS, ncq - cbeg);
      for (int u = t; u < LB_MAXCELLS; u += LB_THREADS) {
        int sz = 0, k = 0;
        if (u < ncc) {
          int q = cbeg + u;
          int cx = cx0 + q % nx, cy = cy0 + (q / nx) % ny, cz = cz0 + q / (nx * ny);
          k = cell_key(g, cx, cy, cz);
          sz = cstart[k + 1] - cstart[k];
        }
        cellk[u] = k;
        cellpre[u] = sz;
      }
      __syncthreads();
      // exclusive scan of the 1024 sizes: 4 per thread, a wavefront scan by
      // shuffles, the wavefront totals through LDS (one barrier)
      {
        static_assert(LB_MAXCELLS == 4 * LB_THREADS, "four cells per thread");
        const int lane = t & 63, wv = t >> 6;
        const int v0 = cellpre[4 * t], v1 = cellpre[4 * t + 1], v2 = cellpre[4 * t + 2],
                  v3 = cellpre[4 * t + 3];
        const int tot = v0 + v1 + v2 + v3;
        int incl = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int v = __shfl_up(incl, off, 64);
          if (lane >= off) incl += v;
        }
        if (lane == 63) scanbuf[wv] = incl;
        __syncthreads();
        int before = 0;
        for (int w2 = 0; w2 < wv; w2++) before += scanbuf[w2];
        incl += before;
        const int base = incl - tot;
        cellpre[4 * t] = base;
        cellpre[4 * t + 1] = base + v0;
        cellpre[4 * t + 2] = base + v0 + v1;
        cellpre[4 * t + 3] = base + v0 + v1 + v2;
        if (t == LB_THREADS - 1) cellpre[LB_MAXCELLS] = incl;
      }
      __syncthreads();
      const int ncand = cellpre[LB_MAXCELLS];
      for (int fb = 0; fb < ncand; fb += LB_THREADS) {
        if (batch++ % npart != part) continue;  // uniform over the workgroup
        const int f = fb + t;
        uint64_t m1 = 0, m2 = 0;
        int js = -1;
        if (f < ncand) {
          // binary search the cell holding flattened candidate f
          int lo = 0, hi = ncc - 1;
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (cellpre[mid] <= f) lo = mid; else hi = mid - 1;
          }
          js = cstart[cellk[lo]] + (f - cellpre[lo]);
          const double cjx = scent[3 * js], cjy = scent[3 * js + 1], cjz = scent[3 * js + 2];
          const int jm = smono[js];
          bool jok = !image || need_only_ok(cjx, cjy, cjz, da, db, dc, P);
          // cheap reject: j farther than the reach R from the (translated)
          // block bounding box is farther than R from every block water
          if (jok) {
            const double ex = fmax(fmax(q0[0] + P.R - cjx, cjx - (q1[0] - P.R)), 0.0);
            const double ey = fmax(fmax(q0[1] + P.R - cjy, cjy - (q1[1] - P.R)), 0.0);
            const double ez = fmax(fmax(q0[2] + P.R - cjz, cjz - (q1[2] - P.R)), 0.0);
            if (ex * ex + ey * ey + ez * ez > R2) jok = false;
          }
          if (jok) {
            // the same test per sub-block: only the members of sub-blocks
            // within R of j are tested below (a member beyond R is no
            // pair: every listed pair has r2 <= R^2, R carrying a margin
            // over the rounding of these distances)
            uint64_t near = 0;
#pragma unroll
            for (int u = 0; u < WBLOCK / 8; u++) {
              const double ex = fmax(fmax((sbb[u][0] + da) - cjx, cjx - (sbb[u][3] + da)), 0.0);
              const double ey = fmax(fmax((sbb[u][1] + db) - cjy, cjy - (sbb[u][4] + db)), 0.0);
              const double ez = fmax(fmax((sbb[u][2] + dc) - cjz, cjz - (sbb[u][5] + dc)), 0.0);
              if (ex * ex + ey * ey + ez * ez <= R2) near |= 0xffull << (8 * u);
            }
            const uint64_t iv = ivalid & near;
            // eight members at a time, their centres and monomers loaded
            // together (the tests of one sub-block are independent), so a
            // lane waits on one round of LDS reads per sub-block instead of
            // one per member
            for (int u = 0; u < WBLOCK / 8; u++) {
              if (((iv >> (8 * u)) & 0xffull) == 0) continue;
#pragma unroll
              for (int q = 0; q < 8; q++) {
                const int i = 8 * u + q;
                const double cx = ic[i][0], cy = ic[i][1], cz = ic[i][2];
                const int mi = im[i];
                if (!((iv >> i) & 1ull) || (!image && !(mi < jm))) continue;
                const int kd = wpair_kind(cx, cy, cz, da, db, dc, cjx, cjy, cjz, image, P);
                if (kd == 1) m1 |= (1ull << i);
                else if (kd == 2) m2 |= (1ull << i);
              }
            }
          }
        }
        // keep the record in candidate order
        const bool has = (m1 | m2) != 0;
        const uint64_t bal = __ballot(has);
        if (lane == 0) rw[wv] = __popcll(bal);
        __syncthreads();
        if (has) {
          int before = 0;
          for (int w2 = 0; w2 < wv; w2++) before += rw[w2];
          const uint64_t lower = (lane == 0) ? 0ull : (bal & ((1ull << lane) - 1ull));
          const int pos = r_run + before + __popcll(lower);
          if (pos < rcap) {
            WRaw r;
            r.j = js;
            r.s = s;
            r.m1 = m1;
            r.m2 = m2;
            rblk[pos] = r;
          }
        }
        __syncthreads();
        if (t == 0) {
          int tot = 0;
          for (int w2 = 0; w2 < LB_THREADS / 64; w2++) tot += rw[w2];
          r_run += tot;
        }
        // items per kind and popcount bucket
        count_kind(L, 0, m1);
        count_kind(L, 1, m2);
      }
      __syncthreads();
    }
    __syncthreads();
  }
  __syncthreads();
  if (t < WNBUCKET) bcount[(size_t)blockIdx.x * WNBUCKET + t] = L.bk_run[t];
  if (t == 0) {
    rcount[blockIdx.x] = r_run;
    if (r_run > rcap) atomicMax(rneed, r_run);
  }
}

// One workgroup per part of a water block: reads the part's records back
// in order and writes its items at boff + the earlier parts' bucket counts
// (bpart, from the predicate pass) + the same deterministic ranks, so the
// item order is that of the block's records part after part.
__global__ __launch_bounds__(LB_THREADS) void k_wblock_write(const WRaw *__restrict__ raw, int rcap,
                                                             const int32_t *__restrict__ rcount,
                                                             int npart,
                                                             const int32_t *__restrict__ bpart,
                                                             const int32_t *__restrict__ boff,
                                                             WItem *__restrict__ items, size_t cap) {
  __shared__ EmitLDS L;
  const int b = blockIdx.x / npart, part = blockIdx.x - b * npart, t = threadIdx.x;
  if (t < WNBUCKET) {
    int v = 0;
    for (int p = 0; p < part; p++) v += bpart[((size_t)b * npart + p) * WNBUCKET + t];
    L.bk_run[t] = v;
  }
  __syncthreads();
  {
    const int slice = blockIdx.x;
    const int n = min(rcount[slice], rcap);
    const WRaw *rblk = raw + (size_t)slice * rcap;
    for (int fb = 0; fb < n; fb += LB_THREADS) {
      const int f = fb + t;
      int js = -1, s = 0;
      uint64_t m1 = 0, m2 = 0;
      if (f < n) {
        const WRaw r = rblk[f];
        js = r.j;
        s = r.s;
        m1 = r.m1;
        m2 = r.m2;
      }
      emit_batch(L, js, s, m1, m2, boff + b * WNBUCKET, items, cap);
    }
  }
}

// item counts per block and bucket: the sum over the block's parts
__global__ void k_bucket_sum(const int32_t *__restrict__ bpart, int nblk, int npart,
                             int32_t *__restrict__ bsum) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nblk * WNBUCKET) return;
  const int b = q / WNBUCKET, k = q - b * WNBUCKET;
  int v = 0;
  for (int p = 0; p < npart; p++) v += bpart[((size_t)b * npart + p) * WNBUCKET + k];
  bsum[q] = v;
}

// items per partner slot j: the real-space kernel leaves each item's
// j-gradient in its own row of jpart, and the per-water reduction reads
// the rows of water j in item order through this CSR
// (nothing when the predicate pass dropped records: the items are then
// incomplete and the build is redone)
__global__ void k_jcount(const WItem *__restrict__ items, const int32_t *__restrict__ nitem,
                         size_t cap, const int32_t *__restrict__ dropped, int32_t *__restrict__ cnt) {
  const size_t n = *dropped ? 0 : min((size_t)*nitem, cap);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[items[q].j], 1);
}
// jpos[item] = the item's row in the partial buffer, inside its partner's
// CSR range (the order within a range does not matter: the rows are
// summed as integers)
// (thread 0 also copies the build's item and chunk totals into the status
// words the host reads at the end)
__global__ void k_jfill(const WItem *__restrict__ items, const int32_t *__restrict__ nitem,
                        size_t cap, const int32_t *__restrict__ dropped,
                        const int32_t *__restrict__ start, int32_t *__restrict__ fill,
                        int32_t *__restrict__ jpos, const int32_t *__restrict__ nchunk,
                        int32_t *__restrict__ st_items, int32_t *__restrict__ st_chunks) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *st_items = *nitem;
    *st_chunks = *nchunk;
  }
  const size_t n = *dropped ? 0 : min((size_t)*nitem, cap);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x) {
    const int j = items[q].j;
    jpos[q] = start[j] + atomicAdd(&fill[j], 1);
  }
}

// ---------------------------------------------------------------------------
// Chunk tables on the device.  A chunk (one workgroup of the pair kernel)
// is a run of one block's items (popcount-descending buckets, boff) holding
// about `pairs` pairs, at least `min_items` items (one per lane) and at
// most `max_items`.  One wavefront per block walks its items twice: to
// count its chunks, then (at the scanned offset) to write them.  Lane k
// holds the block's bucket offset k and the walk reads them with
// readlane, so every lane runs the same (uniform) walk and lane 0 writes.
static_assert(WNBUCKET <= 64, "a wavefront holds one bucket offset per lane");
__device__ inline int chunk_walk(const int32_t *__restrict__ boff, int b, ChunkPar P, int base,
                                 int32_t *__restrict__ chunks, int32_t *__restrict__ cpairs) {
  // bucket k < WSPLIT holds items of popcount WSPLIT - k, and (water
  // lists) bucket WSPLIT + k the listwt2 items of popcount WSPLIT - k,
  // counted as popcount / m2_div pairs (a listwt2 member costs a distance
  // test unless it is within relx)
  const int lane = threadIdx.x & 63;
  const int32_t *op = boff + (size_t)b * P.nbucket;
  const int ol = lane <= P.nbucket ? op[lane] : 0;
  const int end = op[P.nbucket];
  // offset k (0 past P.nbucket)
  auto O = [&](int k) -> int {
    return k < 64 ? __builtin_amdgcn_readlane(ol, k) : (k <= P.nbucket ? end : 0);
  };
  int q = O(0), e = q, n = 0;
  long pairs = 0;
  auto emit = [&]() {
    if (chunks && lane == 0 && base + n < P.cap) {
      chunks[3 * (size_t)(base + n)] = b;
      chunks[3 * (size_t)(base + n) + 1] = q;
      chunks[3 * (size_t)(base + n) + 2] = e;
      if (cpairs) cpairs[base + n] = (int32_t)min(pairs, (long)INT32_MAX);
    }
    n++;
    q = e;
    pairs = 0;
  };
  // a chunk grows while it holds fewer than P.pairs pairs or fewer than
  // P.min_items items, and never past P.max_items items
  for (int k = 0; k < WNBUCKET; k++) {
    // (buckets past P.nbucket have offset 0: no items)
    const int pc = k < WSPLIT ? WSPLIT - k : max(1, (2 * WSPLIT - k) / P.m2_div);
    const int hi = min(O(k + 1), end);
    while (e < hi) {
      // pairs still wanted (none once the budget is met), in 32-bit
      // unsigned arithmetic: the division is the walk's costliest step
      const unsigned need = pairs < P.pairs ? (unsigned)(P.pairs - pairs) : 0u;
      const int want = max(max(1, P.min_items - (e - q)),
                           (int)((need + (unsigned)pc - 1u) / (unsigned)pc));
      int take = min(hi - e, want);
      take = min(take, P.max_items - (e - q));
      pairs += (long)take * pc;
      e += take;
      if (!((pairs < P.pairs || e - q < P.min_items) && e - q < P.max_items)) emit();
    }
  }
  if (e > q) emit();
  return n;
}

__global__ __launch_bounds__(256) void k_chunk_count(const int32_t *__restrict__ boff, ChunkPar P,
                                                     int32_t *__restrict__ ccount) {
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= P.nblk) return;  // uniform over the wavefront
  const int n = chunk_walk(boff, b, P, 0, nullptr, nullptr);
  if ((threadIdx.x & 63) == 0) ccount[b] = n;
}

__global__ __launch_bounds__(256) void k_chunk_write(const int32_t *__restrict__ boff,
                                                     const int32_t *__restrict__ cfirst,
                                                     ChunkPar P, int32_t *__restrict__ chunks,
                                                     int32_t *__restrict__ cpairs) {
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= P.nblk) return;
  chunk_walk(boff, b, P, cfirst[b], chunks, cpairs);
}

// dispatch order, largest chunks first: counting sort on the pair count in
// steps of 32 (one workgroup; ties in any order -- the order changes only
// the timing, each chunk writes its own partial rows)
#define ORDER_KEYS 1024
__global__ __launch_bounds__(1024) void k_chunk_order(const int32_t *__restrict__ nchunk_p,
                                                      const int32_t *__restrict__ cpairs, int cap,
                                                      int32_t *__restrict__ order) {
  __shared__ int32_t cnt[ORDER_KEYS];
  const int t = threadIdx.x;
  const int n = min(*nchunk_p, cap);
  cnt[t] = 0;
  __syncthreads();
  for (int q = t; q < n; q += 1024)
    atomicAdd(&cnt[ORDER_KEYS - 1 - min(cpairs[q] >> 5, ORDER_KEYS - 1)], 1);
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int k = 0; k < ORDER_KEYS; k++) {
      const int v = cnt[k];
      cnt[k] = run;
      run += v;
    }
  }
  __syncthreads();
  for (int q = t; q < n; q += 1024)
    order[atomicAdd(&cnt[ORDER_KEYS - 1 - min(cpairs[q] >> 5, ORDER_KEYS - 1)], 1)] = q;
}

int32_t chunk_table(moil_ctx *c, const int32_t *d_boff, ChunkPar P, int32_t *d_ccount,
                    int32_t *d_cfirst, int32_t *d_chunks, int32_t *d_cpairs, int32_t *d_order,
                    hipStream_t st) {
  const int TB = 256;
  if (P.nblk <= 0) return MOIL_GPU_OK;
  // one wavefront per block, four per workgroup
  k_chunk_count<<<(P.nblk + 3) / 4, TB, 0, st>>>(d_boff, P, d_ccount);
  CHECK_SCAN(scan_excl(c, d_ccount, d_cfirst, P.nblk, st));
  k_chunk_write<<<(P.nblk + 3) / 4, TB, 0, st>>>(d_boff, d_cfirst, P, d_chunks, d_cpairs);
  if (d_order) k_chunk_order<<<1, 1024, 0, st>>>(d_cfirst + P.nblk, d_cpairs, P.cap, d_order);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}


#define CHECK_RC_L(x)                    \
  do {                                   \
    const int32_t rc_ = (x);             \
    if (rc_ != MOIL_GPU_OK) return rc_;  \
  } while (0)

// ---------------------------------------------------------------------------
static void free_lists_buffers(moil_ctx *c) {
  hipFree(c->d_cell_of); c->d_cell_of = nullptr;
  hipFree(c->d_cell_cnt); c->d_cell_cnt = nullptr;
  hipFree(c->d_cell_start); c->d_cell_start = nullptr;
  hipFree(c->d_cell_fill); c->d_cell_fill = nullptr;
}

int32_t lists_init(moil_ctx *c) {
  const moil_params &P = c->par;
  HIP_CHECK(hipMalloc(&c->d_mono_cent, sizeof(double) * 3 * std::max(c->totmon, 1)));
  HIP_CHECK(hipMalloc(&c->d_flags, sizeof(int32_t) * 4));
  HIP_CHECK(hipMalloc(&c->d_status, sizeof(int32_t) * ST_COUNT));
  if (c->nwat > 0) {
    // cell grid over the box with ~8 waters per cell, even counts per
    // dimension are not required (bricks may be partial)
    double L[3] = {P.a, P.b, P.c};
    double vol = P.a * P.b * P.c;
    double edge = cbrt(vol * 8.0 / (double)c->nwat);
    int g[3];
    for (int k = 0; k < 3; k++) {
      g[k] = std::max(1, (int)floor(L[k] / edge));
      c->cs[k] = L[k] / g[k];
      c->gmin[k] = -0.5 * L[k];
    }
    c->gx = g[0]; c->gy = g[1]; c->gz = g[2];
    int nbx = (g[0] + 1) / 2, nby = (g[1] + 1) / 2, nbz = (g[2] + 1) / 2;
    c->ncell = nbx * nby * nbz * 8;
    HIP_CHECK(hipMalloc(&c->d_cell_of, sizeof(int32_t) * c->nwat));
    HIP_CHECK(hipMalloc(&c->d_cell_cnt, sizeof(int32_t) * (c->ncell + 1)));
    HIP_CHECK(hipMalloc(&c->d_cell_start, sizeof(int32_t) * (c->ncell + 1)));
    HIP_CHECK(hipMalloc(&c->d_cell_fill, sizeof(int32_t) * (c->ncell + 1)));
    HIP_CHECK(hipMalloc(&c->d_slot2wat, sizeof(int32_t) * c->nwat));
    HIP_CHECK(hipMalloc(&c->d_slot_o, sizeof(int32_t) * c->nwat));
    HIP_CHECK(hipMalloc(&c->d_scent, sizeof(double) * 3 * c->nwat));
    HIP_CHECK(hipMalloc(&c->d_smono, sizeof(int32_t) * c->nwat));
    c->nwblk = (c->nwat + WBLOCK - 1) / WBLOCK;
    // parts per block for the predicate pass: at least ~2048 workgroups
    c->wnpart = std::min(8, std::max(1, (2048 + c->nwblk - 1) / c->nwblk));
    HIP_CHECK(hipMalloc(&c->d_bpart, sizeof(int32_t) * ((size_t)c->nwblk * c->wnpart * WNBUCKET + 1)));
    HIP_CHECK(hipMalloc(&c->d_bcount, sizeof(int32_t) * (c->nwblk * WNBUCKET + 1)));
    HIP_CHECK(hipMalloc(&c->d_boff, sizeof(int32_t) * (c->nwblk * WNBUCKET + 1)));
  }
  return MOIL_GPU_OK;
}

void lists_free(moil_ctx *c) {
  free_lists_buffers(c);
  hipFree(c->d_mono_cent); c->d_mono_cent = nullptr;
  hipFree(c->d_flags); c->d_flags = nullptr;
  hipFree(c->d_slot2wat); c->d_slot2wat = nullptr;
  hipFree(c->d_slot_o); c->d_slot_o = nullptr;
  hipFree(c->d_scent); c->d_scent = nullptr;
  hipFree(c->d_smono); c->d_smono = nullptr;
  hipFree(c->d_bcount); c->d_bcount = nullptr;
  hipFree(c->d_boff); c->d_boff = nullptr;
  hipFree(c->d_witems); c->d_witems = nullptr;
  hipFree(c->d_wchunk); c->d_wchunk = nullptr;
  hipFree(c->d_wraw); c->d_wraw = nullptr;
  hipFree(c->d_rcount); c->d_rcount = nullptr;
  hipFree(c->d_jcnt); c->d_jcnt = nullptr;
  hipFree(c->d_jfill); c->d_jfill = nullptr;
  hipFree(c->d_jstart); c->d_jstart = nullptr;
  hipFree(c->d_jpos); c->d_jpos = nullptr;
  hipFree(c->d_jpart); c->d_jpart = nullptr;
  hipFree(c->d_ipart); c->d_ipart = nullptr;
  hipFree(c->d_cfirst); c->d_cfirst = nullptr;
  hipFree(c->d_wccount); c->d_wccount = nullptr;
  hipFree(c->d_wcpairs); c->d_wcpairs = nullptr;
  hipFree(c->d_worder); c->d_worder = nullptr;
  hipFree(c->d_status); c->d_status = nullptr;
  hipFree(c->d_bpart); c->d_bpart = nullptr;
}

__global__ void k_clear(ClearSet S) {
  for (int k = 0; k < S.count; k++)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < S.n[k]; i += gridDim.x * blockDim.x)
      S.p[k][i] = S.v[k];
}
void launch_clear(const ClearSet &cs, hipStream_t st) {
  int most = 0;
  for (int k = 0; k < cs.count; k++) most = std::max(most, cs.n[k]);
  if (most > 0) k_clear<<<std::min((most + 255) / 256, 512), 256, 0, st>>>(cs);
}

int32_t lists_build(moil_ctx *c, const double *d_coor, hipStream_t st) {
  const moil_params &P = c->par;
  const int TB = 256;
  // the build's counters and status words, cleared in one launch
  if (c->nwat > 0 && !c->d_jcnt) {
    HIP_CHECK(hipMalloc(&c->d_jcnt, sizeof(int32_t) * (c->nwat + 1)));
    HIP_CHECK(hipMalloc(&c->d_jfill, sizeof(int32_t) * (c->nwat + 1)));
    HIP_CHECK(hipMalloc(&c->d_jstart, sizeof(int32_t) * (c->nwat + 1)));
  }
  {
    ClearSet cs;
    cs.add(c->d_status, ST_COUNT, 0);
    cs.add(c->d_flags, 4, 0);
    if (c->nwat > 0) {
      cs.add(c->d_cell_cnt, c->ncell + 1, 0);
      cs.add(c->d_cell_fill, c->ncell + 1, 0);
      cs.add(c->d_jcnt, c->nwat + 1, 0);
      cs.add(c->d_jfill, c->nwat + 1, 0);
    }
    launch_clear(cs, st);
  }
  if (c->totmon > 0)
    k_mono_cent<<<(c->totmon + TB - 1) / TB, TB, 0, st>>>(d_coor, c->d_poipt, c->d_monoflag,
                                                           c->totmon, 0.5 * P.a, 0.5 * P.b,
                                                           0.5 * P.c, c->d_mono_cent, c->d_flags);
  if (c->nwat > 0) {
    Grid g;
    g.gx = c->gx; g.gy = c->gy; g.gz = c->gz;
    g.nbx = (c->gx + 1) / 2; g.nby = (c->gy + 1) / 2; g.nbz = (c->gz + 1) / 2;
    g.g0x = c->gmin[0]; g.g0y = c->gmin[1]; g.g0z = c->gmin[2];
    g.csx = c->cs[0]; g.csy = c->cs[1]; g.csz = c->cs[2];
    int nblk = (c->nwat + TB - 1) / TB;
    k_wat_cells<<<nblk, TB, 0, st>>>(c->d_mono_cent, c->d_wat_mono, c->nwat, g, c->d_cell_of,
                                     c->d_cell_cnt);
    CHECK_SCAN(scan_excl(c, c->d_cell_cnt, c->d_cell_start, c->ncell, st));
    k_wat_scatter<<<nblk, TB, 0, st>>>(c->d_cell_of, c->nwat, c->d_cell_start, c->d_cell_fill,
                                       c->d_slot2wat);
    launch_cell_sort(c->d_cell_start, c->ncell, c->d_slot2wat, st);
    double *d_scent = c->d_scent;
    int32_t *d_smono = c->d_smono;
    k_slot_fill<<<nblk, TB, 0, st>>>(c->d_slot2wat, c->d_wat_mono, c->d_wat_o, c->d_mono_cent,
                                     c->nwat, d_scent, d_smono, c->d_slot_o);
    WPar wp;
    wp.a = P.a; wp.b = P.b; wp.c = P.c;
    wp.ha = 0.5 * P.a; wp.hb = 0.5 * P.b; wp.hc = 0.5 * P.c;
    wp.cutmono2 = P.cutmono2; wp.cutvbig2 = P.cutvbig2; wp.cutebig2 = P.cutebig2;
    // candidate reach: every listed pair has r2 <= max(cutvbig2, cutebig2)
    double R = sqrt(std::max(std::max(P.cutvbig2, P.cutebig2), 0.0));
    wp.R = R * (1.0 + 1e-12) + 1e-9;
    wp.nwat = c->nwat;
    // Everything below is enqueued without waiting: the predicate pass
    // keeps its records per block (capacity from earlier builds), the items
    // and the chunk table are written on the device, and every capacity
    // overflow is reported in the status words that moil_gpu_build_lists_dev
    // reads once at the end (the build is then redone with enough room).
    const int np = c->wnpart;
    if (!c->d_rcount) HIP_CHECK(hipMalloc(&c->d_rcount, sizeof(int32_t) * (c->nwblk * np + 1)));
    if (!c->d_wccount) HIP_CHECK(hipMalloc(&c->d_wccount, sizeof(int32_t) * (c->nwblk + 1)));
    if (!c->d_cfirst) HIP_CHECK(hipMalloc(&c->d_cfirst, sizeof(int32_t) * (c->nwblk + 1)));
    if (!c->d_wraw) {
      c->wraw_cap = std::max(c->wraw_cap, 1024);
      HIP_CHECK(hipMalloc(&c->d_wraw, sizeof(WRaw) * (size_t)c->wraw_cap * c->nwblk * np));
    }
    if (!c->d_witems) {
      c->witem_cap = std::max<size_t>(c->witem_cap, 4096);
      HIP_CHECK(hipMalloc(&c->d_witems, sizeof(WItem) * c->witem_cap));
      HIP_CHECK(hipMalloc(&c->d_jpos, sizeof(int32_t) * c->witem_cap));
      HIP_CHECK(hipMalloc(&c->d_jpart, sizeof(long long) * 9 * c->witem_cap));
    }
    if (!c->d_wchunk) {
      c->wchunk_cap = std::max(c->wchunk_cap, 1024);
      HIP_CHECK(hipMalloc(&c->d_wchunk, sizeof(int32_t) * 3 * (size_t)c->wchunk_cap));
      HIP_CHECK(hipMalloc(&c->d_wcpairs, sizeof(int32_t) * (size_t)c->wchunk_cap));
      HIP_CHECK(hipMalloc(&c->d_worder, sizeof(int32_t) * (size_t)c->wchunk_cap));
    }
    k_wblock_items<<<c->nwblk * np, LB_THREADS, 0, st>>>(d_scent, d_smono, c->d_cell_start, g, wp,
                                                         c->d_flags, c->d_bpart, c->d_wraw,
                                                         c->wraw_cap, c->d_rcount,
                                                         c->d_status + ST_WRAW, np);
    k_bucket_sum<<<(c->nwblk * WNBUCKET + TB - 1) / TB, TB, 0, st>>>(c->d_bpart, c->nwblk, np,
                                                                    c->d_bcount);
    CHECK_SCAN(scan_excl(c, c->d_bcount, c->d_boff, c->nwblk * WNBUCKET, st));
    k_wblock_write<<<c->nwblk * np, LB_THREADS, 0, st>>>(c->d_wraw, c->wraw_cap, c->d_rcount,
                                                         np, c->d_bpart, c->d_boff, c->d_witems,
                                                         c->witem_cap);
    ChunkPar cp;
    cp.nblk = c->nwblk;
    cp.pairs = c->wchunk_pairs;
    cp.max_items = c->wchunk_max_items;
    cp.min_items = LB_THREADS;
    cp.cap = c->wchunk_cap;
    cp.nbucket = WNBUCKET;
    cp.m2_div = 8;
    CHECK_RC_L(chunk_table(c, c->d_boff, cp, c->d_wccount, c->d_cfirst, c->d_wchunk, c->d_wcpairs,
                           c->d_worder, st));
    // row ranges of the partial buffer by partner slot j
    const int32_t *d_nitem = c->d_boff + (size_t)c->nwblk * WNBUCKET;
    k_jcount<<<1024, TB, 0, st>>>(c->d_witems, d_nitem, c->witem_cap, c->d_status + ST_WRAW,
                                  c->d_jcnt);
    CHECK_SCAN(scan_excl(c, c->d_jcnt, c->d_jstart, c->nwat, st));
    k_jfill<<<1024, TB, 0, st>>>(c->d_witems, d_nitem, c->witem_cap, c->d_status + ST_WRAW,
                                 c->d_jstart, c->d_jfill, c->d_jpos, c->d_cfirst + c->nwblk,
                                 c->d_status + ST_WITEMS, c->d_status + ST_WCHUNK);
    HIP_CHECK(hipGetLastError());
  }
  return MOIL_GPU_OK;
}

// capacities after a build whose status reported an overflow (false: the
// build fitted); grown buffers are empty, the build must be redone
bool lists_grow(moil_ctx *c, const int32_t *status) {
  bool again = false;
  if (c->nwat <= 0) return false;
  if (status[ST_WRAW] > c->wraw_cap) {
    hipFree(c->d_wraw);
    c->d_wraw = nullptr;
    c->wraw_cap = status[ST_WRAW] + status[ST_WRAW] / 4 + 256;
    again = true;
  }
  if ((size_t)status[ST_WITEMS] > c->witem_cap) {
    hipFree(c->d_witems);
    hipFree(c->d_jpos);
    hipFree(c->d_jpart);
    c->d_witems = nullptr;
    c->d_jpos = nullptr;
    c->d_jpart = nullptr;
    c->witem_cap = (size_t)status[ST_WITEMS] + status[ST_WITEMS] / 4 + 1024;
    again = true;
  }
  if (status[ST_WCHUNK] > c->wchunk_cap) {
    hipFree(c->d_wchunk);
    hipFree(c->d_wcpairs);
    hipFree(c->d_worder);
    c->d_wchunk = c->d_wcpairs = c->d_worder = nullptr;
    c->wchunk_cap = status[ST_WCHUNK] + status[ST_WCHUNK] / 4 + 64;
    again = true;
  }
  return again;
}

// sizes of a build that fitted; per-chunk partial rows sized to match
int32_t lists_finish(moil_ctx *c, const int32_t *status) {
  if (c->nwat <= 0) {
    c->nwitems = 0;
    c->nwchunk = 0;
    return MOIL_GPU_OK;
  }
  c->nwitems = (size_t)status[ST_WITEMS];
  c->nwchunk = status[ST_WCHUNK];
  if (c->nwchunk > c->ipart_cap || c->d_ipart == nullptr) {
    hipFree(c->d_ipart);
    c->ipart_cap = std::max(c->nwchunk + c->nwchunk / 4 + 64, 256);
    HIP_CHECK(hipMalloc(&c->d_ipart, sizeof(long long) * WBLOCK * 9 * (size_t)c->ipart_cap));
  }
  return MOIL_GPU_OK;
}

