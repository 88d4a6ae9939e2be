// MOIL's atom-pair lists and their real-space energy on gfx950: every pair
// with at least one atom outside a water monomer.
//
// Lists (pot/nbond.f:34-149 behind pot/nbondm.f:172-211 for real pairs;
// symm/nbsym.f:58-142 behind symm/nbmsym.f:178-233 for the 13 image
// translations), reproduced bit for bit with FP64 predicates in MOIL's
// operation order and FMA contraction off:
//  * real pairs: the two monomers are one non-water monomer (atoms l < m)
//    or two monomers, not both water, whose centres satisfy
//    r2 <= cutmono2; excluded pairs are dropped (pot/nbond.f:95-103);
//  * image pairs (l + T, m): the image monomer passes the rch2 box filter,
//    the real monomer the need_only face filter, centre r2 < cutmono2,
//    not both water, no exclusion test (symm/nbsym.f:79-142);
//  * classification on the atom distance: r2 < cutvbig2 -> list1 (both
//    charged) or list2, else r2 < cutebig2 and both charged -> list3; the
//    hydrogen rule (epsgm12 < 1e-3 within cutebig2: list3 if both
//    charged, else dropped) applies to real pairs when hvdw0 is set
//    (pot/nbond.f:115-126) and always to image pairs (symm/nbsym.f:105-115).
//
// Layout: the atoms of non-water monomers ("solute") are sorted by cell
// into 64-atom blocks; partners j come from a cell grid over all atoms.
// A block is searched in 27 directions: real, forward images (the block
// atom is the translated member l, partner any atom) and reverse images
// (a water atom is the translated member l, the block atom is m).  Items
// are (partner atom, direction, list1/list2/list3 masks over the block),
// split, bucketed and chunked exactly like the water items.
//
// Energy (pot/cdie_ewald.f:309-673, symm/symcdie_ewald.f:100-549): list1
// Coulomb always plus LJ when r2 <= rvmx^2; list2 LJ when r2 <= rvmx^2;
// list3 Coulomb when r2 <= relx^2; LJ A/r^12 - B/r^6 with geometric
// combination A = epsgm12_l epsgm12_m, B = epsgm6_l epsgm6_m; Coulomb
// q_l q_m (kofdie/eps) erfc(beta r)/r.  Gradients go to per-item and
// per-chunk partial rows (no global atomics) that one reduction pass adds
// to the fixed-point accumulators.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "multi.h"
#include "pairmath.h"
#include "grid.h"

#define ABT 256
#define AB_MAXCELLS 1024
#define ADIRS 27

// per-atom flag bits
#define AF_CHG 1    // flagchr
#define AF_SMALL 2  // epsgm12 < 1e-3 (hydrogen rule)
#define AF_WAT 4    // atom of a water monomer

struct APar : BoxPar {
  double R;
  int hvdw0;
  int nsol;
};


// exclusion test for atoms lo < hi (0-based): partners up to lo+64 from
// the bit mask, farther ones from MOIL's exclusion list (exc1/exc2)
__device__ inline bool excluded(int lo, int hi, const uint64_t *__restrict__ xmask,
                                const int32_t *__restrict__ xfar, const int32_t *__restrict__ exc1,
                                const int32_t *__restrict__ exc2) {
  const int d = hi - lo;
  if (d <= 64) return (xmask[lo] >> (d - 1)) & 1ull;
  if (!xfar[lo]) return false;
  for (int k = exc1[lo]; k < exc1[lo + 1]; k++)
    if (exc2[k] - 1 == hi) return true;
  return false;
}

// MOIL's list choice for an admitted pair (0 = none, 1/2/3 = list1/2/3)
__device__ inline int apair_class(double r2, int fl, int fm, bool hrule, const APar &P) {
  const bool chg = (fl & AF_CHG) && (fm & AF_CHG);
  if (hrule && ((fl | fm) & AF_SMALL) && r2 < P.cutebig2) return chg ? 3 : 0;
  if (r2 < P.cutvbig2) return chg ? 1 : 2;
  if (r2 < P.cutebig2 && chg) return 3;
  return 0;
}

__device__ inline double dist2(double ax, double ay, double az, double bx, double by, double bz) {
#pragma clang fp contract(off)
  double rx = ax - bx, ry = ay - by, rz = az - bz;
  return rx * rx + ry * ry + rz * rz;
}

// ---------------------------------------------------------------------------
// binning of atom sets into the cell grid
__global__ void k_atom_cells(const double *__restrict__ coor, const int32_t *__restrict__ ids, int n,
                             Grid g, int32_t *__restrict__ cell_of, int32_t *__restrict__ cnt) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int a = ids ? ids[q] : q;
  int cx = cell_coord(coor[3 * a], g.g0x, g.csx, g.gx);
  int cy = cell_coord(coor[3 * a + 1], g.g0y, g.csy, g.gy);
  int cz = cell_coord(coor[3 * a + 2], g.g0z, g.csz, g.gz);
  int k = cell_key(g, cx, cy, cz);
  cell_of[q] = k;
  atomicAdd(&cnt[k], 1);
}

__global__ void k_atom_scatter(const int32_t *__restrict__ cell_of, const int32_t *__restrict__ ids,
                               int n, const int32_t *__restrict__ start, int32_t *__restrict__ fill,
                               int32_t *__restrict__ out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int k = cell_of[q];
  out[start[k] + atomicAdd(&fill[k], 1)] = ids ? ids[q] : q;
}

__global__ void k_slot_of(const int32_t *__restrict__ sslot2atom, int nslot,
                          int32_t *__restrict__ slot_of) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nslot && sslot2atom[q] >= 0) slot_of[sslot2atom[q]] = q;
}

// Solute slots padded per brick row: the cell keys of one row of bricks
// (fixed bz, by) are contiguous, and each row's atoms start a new block,
// so a block never spans the end of a row (whose next atoms lie a whole
// box length away in x).  Padding slots hold -1.
__global__ void k_row_pad(const int32_t *__restrict__ sstart, int nrow, int rowcells,
                          int32_t *__restrict__ rowcnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  const int n = sstart[(r + 1) * rowcells] - sstart[r * rowcells];
  rowcnt[r] = (n + WBLOCK - 1) / WBLOCK * WBLOCK;
}

__global__ void k_pad_slots(const int32_t *__restrict__ sslot_u, int nsol,
                            const int32_t *__restrict__ sstart, const int32_t *__restrict__ rowoff,
                            int nrow, int rowcells, int32_t *__restrict__ sslot2atom) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nsol) return;
  int lo = 0, hi = nrow - 1;  // last row starting at or before u
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sstart[mid * rowcells] <= u) lo = mid; else hi = mid - 1;
  }
  sslot2atom[rowoff[lo] + (u - sstart[lo * rowcells])] = sslot_u[u];
}

// ---------------------------------------------------------------------------
// One workgroup per (solute block, part): evaluates MOIL's predicates once
// for every candidate partner and direction, the part taking every
// npart-th batch of its block's candidates (batches numbered over
// directions and cells; several parts keep the GPU busy when the system
// has few solute blocks), keeps the non-empty records (partner,
// direction, list1/2/3 masks over the block) in candidate order in its
// slice of `raw` (capacity rcap; the count is exact even when the slice
// overflows), and counts the items per popcount bucket.
// For each (block atom, partner) the distance test comes first: a pair
// beyond max(rvbg, rebg) is in no list whatever the monomer prefilter and
// the exclusions say, so those (with their memory lookups) are only
// evaluated for pairs inside the reach.  All tests are conjunctions with
// MOIL's operations and order, so the pair set is unchanged.
__global__ __launch_bounds__(ABT) void k_ablock_raw(
    const double *__restrict__ coor, const int32_t *__restrict__ sslot2atom,
    const int32_t *__restrict__ aslot2atom, const int32_t *__restrict__ acstart, Grid g,
    const int32_t *__restrict__ mono_of, const int32_t *__restrict__ aflag,
    const double *__restrict__ mcent, const int32_t *__restrict__ flags,
    const uint64_t *__restrict__ xmask, const int32_t *__restrict__ xfar,
    const int32_t *__restrict__ exc1, const int32_t *__restrict__ exc2, APar P,
    int32_t *__restrict__ bcount, AItem *__restrict__ raw, int rcap,
    int32_t *__restrict__ rcount, int32_t *__restrict__ rneed, int npart) {
#pragma clang fp contract(off)
  __shared__ double ix[WBLOCK][3], ic[WBLOCK][3];
  __shared__ int32_t iatom[WBLOCK], imono[WBLOCK], iflag[WBLOCK];
  __shared__ double bb[6];
  __shared__ uint64_t imask;
  __shared__ int32_t cellk[AB_MAXCELLS];
  __shared__ int32_t cellpre[AB_MAXCELLS + 1];
  __shared__ int32_t scanbuf[ABT];
  __shared__ int32_t bk_cnt[NBUCKET];
  __shared__ int32_t rw[ABT / 64];
  __shared__ int32_t r_run;

  const int b = blockIdx.x / npart, part = blockIdx.x - b * npart;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int s0 = b * WBLOCK;
  int batch = 0;
  if (t < WBLOCK) {
    const int a = (s0 + t < P.nsol) ? sslot2atom[s0 + t] : -1;
    iatom[t] = a;
    if (a >= 0) {
      const int m = mono_of[a];
      imono[t] = m;
      iflag[t] = aflag[a];
      for (int k = 0; k < 3; k++) {
        ix[t][k] = coor[3 * a + k];
        ic[t][k] = mcent[3 * m + k];
      }
    }
  }
  if (t < NBUCKET) bk_cnt[t] = 0;
  if (t == 0) r_run = 0;
  __syncthreads();
  if (t == 0) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < WBLOCK; i++) {
      if (iatom[i] < 0) continue;
      for (int k = 0; k < 3; k++) {
        lo[k] = fmin(lo[k], ix[i][k]);
        hi[k] = fmax(hi[k], ix[i][k]);
      }
    }
    for (int k = 0; k < 3; k++) {
      bb[k] = lo[k];
      bb[3 + k] = hi[k];
    }
  }
  __syncthreads();
  const double large_cutmono = flags[0] ? P.cutmono2 * 4.0 : P.cutmono2;
  const double reach2 = fmax(P.cutvbig2, P.cutebig2);
  AItem *rblk = raw + (size_t)blockIdx.x * rcap;

  for (int d = 0; d < ADIRS; d++) {
    int s;
    bool rev;
    dir_of(d, s, rev);
    int ia, jb, kc;
    symop_of(s, &ia, &jb, &kc);
    const double da = ia * P.a, db = jb * P.b, dc = kc * P.c;
    const bool image = s > 0;
    // block-atom filter for this direction
    if (t < 64) {
      bool ok = false;
      if (iatom[t] >= 0) {
        if (!image) ok = true;
        else if (!rev) ok = image_i_ok(ic[t][0], ic[t][1], ic[t][2], da, db, dc, large_cutmono, P);
        else ok = need_only_ok(ic[t][0], ic[t][1], ic[t][2], da, db, dc, P);
      }
      uint64_t bal = __ballot(ok);
      if (t == 0) imask = bal;
    }
    __syncthreads();
    const uint64_t ivalid = imask;
    if (ivalid == 0) {
      __syncthreads();
      continue;
    }
    // query box: block bbox moved to the partner's frame, grown by R
    const double sx = rev ? -da : da, sy = rev ? -db : db, sz = rev ? -dc : dc;
    double q0[3], q1[3];
    q0[0] = bb[0] + sx - P.R; q1[0] = bb[3] + sx + P.R;
    q0[1] = bb[1] + sy - P.R; q1[1] = bb[4] + sy + P.R;
    q0[2] = bb[2] + sz - P.R; q1[2] = bb[5] + sz + P.R;
    const int cx0 = clampi((int)floor((q0[0] - g.g0x) / g.csx), 0, g.gx - 1);
    const int cx1 = clampi((int)floor((q1[0] - g.g0x) / g.csx), 0, g.gx - 1);
    const int cy0 = clampi((int)floor((q0[1] - g.g0y) / g.csy), 0, g.gy - 1);
    const int cy1 = clampi((int)floor((q1[1] - g.g0y) / g.csy), 0, g.gy - 1);
    const int cz0 = clampi((int)floor((q0[2] - g.g0z) / g.csz), 0, g.gz - 1);
    const int cz1 = clampi((int)floor((q1[2] - g.g0z) / g.csz), 0, g.gz - 1);
    const int nx = cx1 - cx0 + 1, ny = cy1 - cy0 + 1, nz = cz1 - cz0 + 1;
    const int ncq = nx * ny * nz;
    for (int cbeg = 0; cbeg < ncq; cbeg += AB_MAXCELLS) {
      const int ncc = min(AB_MAXCELLS, ncq - cbeg);
      for (int u = t; u < AB_MAXCELLS; u += ABT) {
        int sz2 = 0, k = 0;
        if (u < ncc) {
          int q = cbeg + u;
          int cx = cx0 + q % nx, cy = cy0 + (q / nx) % ny, cz = cz0 + q / (nx * ny);
          k = cell_key(g, cx, cy, cz);
          sz2 = acstart[k + 1] - acstart[k];
        }
        cellk[u] = k;
        cellpre[u] = sz2;
      }
      __syncthreads();
      {
        // exclusive scan of the cell sizes: four cells per thread, a
        // wavefront scan by shuffles, the wavefront totals through LDS
        // (one barrier instead of a barrier pair per step)
        static_assert(AB_MAXCELLS == 4 * ABT, "four cells per thread");
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
        if (t == ABT - 1) cellpre[AB_MAXCELLS] = incl;
      }
      __syncthreads();
      const int ncand = cellpre[AB_MAXCELLS];
      for (int fb = 0; fb < ncand; fb += ABT) {
        if (batch++ % npart != part) continue;  // uniform over the workgroup
        const int f = fb + t;
        uint64_t m1 = 0, m2 = 0, m3 = 0;
        int ja = -1;
        if (f < ncand) {
          int lo = 0, hi = ncc - 1;
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (cellpre[mid] <= f) lo = mid; else hi = mid - 1;
          }
          ja = aslot2atom[acstart[cellk[lo]] + (f - cellpre[lo])];
          const double xj = coor[3 * ja], yj = coor[3 * ja + 1], zj = coor[3 * ja + 2];
          // j farther than the reach R from the block's bounding box (in
          // j's frame) is farther than R from every block atom
          const double ex = fmax(fmax(q0[0] + P.R - xj, xj - (q1[0] - P.R)), 0.0);
          const double ey = fmax(fmax(q0[1] + P.R - yj, yj - (q1[1] - P.R)), 0.0);
          const double ez = fmax(fmax(q0[2] + P.R - zj, zj - (q1[2] - P.R)), 0.0);
          bool jok = ex * ex + ey * ey + ez * ez <= P.R * P.R;
          if (jok) {
            const int jm = mono_of[ja], jf = aflag[ja];
            const double cjx = mcent[3 * jm], cjy = mcent[3 * jm + 1], cjz = mcent[3 * jm + 2];
            if (image && !rev) jok = need_only_ok(cjx, cjy, cjz, da, db, dc, P);
            if (rev) jok = (jf & AF_WAT) && image_i_ok(cjx, cjy, cjz, da, db, dc, large_cutmono, P);
            uint64_t iv = jok ? ivalid : 0ull;
            while (iv) {
              const int i = __ffsll((long long)iv) - 1;
              iv &= iv - 1;
              const int ai = iatom[i];
              int cls = 0;
              if (!image) {
                // real pair (pot/nbondm.f:176-199, pot/nbond.f:67-149)
                if (!(jf & AF_WAT) && !(ai < ja)) continue;
                const double r2 = dist2(ix[i][0], ix[i][1], ix[i][2], xj, yj, zj);
                if (!(r2 < reach2)) continue;
                if (imono[i] != jm) {
                  if (!(dist2(ic[i][0], ic[i][1], ic[i][2], cjx, cjy, cjz) <= P.cutmono2)) continue;
                }
                const int lo2 = min(ai, ja), hi2 = max(ai, ja);
                if (excluded(lo2, hi2, xmask, xfar, exc1, exc2)) continue;
                cls = apair_class(r2, iflag[i], jf, P.hvdw0 != 0, P);
              } else if (!rev) {
                // block atom translated (symm/nbmsym.f:197-233, symm/nbsym.f:79-142)
                const double r2 = dist2(ix[i][0] + da, ix[i][1] + db, ix[i][2] + dc, xj, yj, zj);
                if (!(r2 < reach2)) continue;
                if (!(dist2(ic[i][0] + da, ic[i][1] + db, ic[i][2] + dc, cjx, cjy, cjz) < P.cutmono2))
                  continue;
                cls = apair_class(r2, iflag[i], jf, true, P);
              } else {
                // water atom translated, block atom real
                const double r2 = dist2(xj + da, yj + db, zj + dc, ix[i][0], ix[i][1], ix[i][2]);
                if (!(r2 < reach2)) continue;
                if (!(dist2(cjx + da, cjy + db, cjz + dc, ic[i][0], ic[i][1], ic[i][2]) < P.cutmono2))
                  continue;
                cls = apair_class(r2, jf, iflag[i], true, P);
              }
              if (cls == 1) m1 |= 1ull << i;
              else if (cls == 2) m2 |= 1ull << i;
              else if (cls == 3) m3 |= 1ull << i;
            }
          }
        }
        // bucket counts (order independent) and the record in candidate order
        const uint64_t mm = m1 | m2 | m3;
        const int pc = __popcll(mm);
        if (pc >= WSPLIT) atomicAdd(&bk_cnt[0], pc / WSPLIT);
        if (pc % WSPLIT) atomicAdd(&bk_cnt[WSPLIT - pc % WSPLIT], 1);
        const bool has = mm != 0;
        const uint64_t bal = __ballot(has);
        if (lane == 0) rw[wv] = __popcll(bal);
        __syncthreads();
        if (has) {
          int before = 0;
          for (int w2 = 0; w2 < wv; w2++) before += rw[w2];
          const uint64_t lower = (lane == 0) ? 0ull : (bal & ((1ull << lane) - 1ull));
          const int pos = r_run + before + __popcll(lower);
          if (pos < rcap) {
            AItem r;
            r.j = ja;
            r.s = d;
            r.m1 = m1;
            r.m2 = m2;
            r.m3 = m3;
            rblk[pos] = r;
          }
        }
        __syncthreads();
        if (t == 0) {
          int tot = 0;
          for (int w2 = 0; w2 < ABT / 64; w2++) tot += rw[w2];
          r_run += tot;
        }
      }
      __syncthreads();
    }
    __syncthreads();
  }
  __syncthreads();
  if (t < NBUCKET) bcount[(size_t)blockIdx.x * NBUCKET + t] = bk_cnt[t];
  if (t == 0) {
    rcount[blockIdx.x] = r_run;
    if (r_run > rcap) atomicMax(rneed, r_run);
  }
}

// item counts per block and bucket: the sum over the block's parts
__global__ void k_abucket_sum(const int32_t *__restrict__ bpart, int nblk, int npart,
                              int32_t *__restrict__ bsum) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nblk * NBUCKET) return;
  const int b = q / NBUCKET, k = q - b * NBUCKET;
  int v = 0;
  for (int p = 0; p < npart; p++) v += bpart[((size_t)b * npart + p) * NBUCKET + k];
  bsum[q] = v;
}

// One workgroup per part of a solute block: reads the part's records back
// in order and writes its items -- each record split into items of <=
// WSPLIT pairs (consecutive mask bits), bucketed by popcount -- at boff +
// the earlier parts' bucket counts (bcount, from the predicate pass) +
// ranks taken in thread order (wavefront scans and ballots), so the item
// order is that of the block's records part after part.
__global__ __launch_bounds__(ABT) void k_ablock_write(const AItem *__restrict__ raw, int rcap,
                                                      const int32_t *__restrict__ rcount,
                                                      int npart,
                                                      const int32_t *__restrict__ bcount,
                                                      const int32_t *__restrict__ boff,
                                                      AItem *__restrict__ items, size_t cap) {
  __shared__ int32_t bk_run[NBUCKET];
  __shared__ int32_t wcnt[ABT / 64][NBUCKET];
  const int b = blockIdx.x / npart, part = blockIdx.x - b * npart;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t < NBUCKET) {
    int v = 0;
    for (int p = 0; p < part; p++) v += bcount[((size_t)b * npart + p) * NBUCKET + t];
    bk_run[t] = v;
  }
  __syncthreads();
  {
  const int slice = blockIdx.x;
  const int n = min(rcount[slice], rcap);
  const AItem *rblk = raw + (size_t)slice * rcap;
  for (int fb = 0; fb < n; fb += ABT) {
    const int f = fb + t;
    int ja = -1, d = 0;
    uint64_t m1 = 0, m2 = 0, m3 = 0;
    if (f < n) {
      const AItem r = rblk[f];
      ja = r.j;
      d = r.s;
      m1 = r.m1;
      m2 = r.m2;
      m3 = r.m3;
    }
    const uint64_t mm = m1 | m2 | m3;
    const int pc = __popcll(mm);
    const int nfull = pc / WSPLIT, rem = pc % WSPLIT;
    const int bk = rem ? WSPLIT - rem : -1;
    int incl = nfull;
    for (int off = 1; off < 64; off <<= 1) {
      int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    const int wtot0 = __shfl(incl, 63, 64);
    if (lane == 0) wcnt[wv][0] = wtot0;
    for (int k = 1; k < NBUCKET; k++) {
      uint64_t bal = __ballot(bk == k);
      if (lane == 0) wcnt[wv][k] = __popcll(bal);
    }
    __syncthreads();
    uint64_t rest = mm;
    if (nfull > 0) {
      int before0 = 0;
      for (int w2 = 0; w2 < wv; w2++) before0 += wcnt[w2][0];
      const int pos0 = boff[b * NBUCKET] + bk_run[0] + before0 + (incl - nfull);
      for (int q = 0; q < nfull; q++) {
        uint64_t piece = 0;
        for (int k = 0; k < WSPLIT; k++) {
          const uint64_t low = rest & (~rest + 1ull);
          piece |= low;
          rest ^= low;
        }
        AItem it;
        it.j = ja;
        it.s = d;
        it.m1 = m1 & piece;
        it.m2 = m2 & piece;
        it.m3 = m3 & piece;
        if ((size_t)(pos0 + q) < cap) items[pos0 + q] = it;
      }
    }
    for (int k = 1; k < NBUCKET; k++) {
      uint64_t bal = __ballot(bk == k);
      if (bk == k) {
        int before = 0;
        for (int w2 = 0; w2 < wv; w2++) before += wcnt[w2][k];
        uint64_t lower = (lane == 0) ? 0ull : (bal & ((1ull << lane) - 1ull));
        const int pos = boff[b * NBUCKET + k] + bk_run[k] + before + __popcll(lower);
        AItem it;
        it.j = ja;
        it.s = d;
        it.m1 = m1 & rest;
        it.m2 = m2 & rest;
        it.m3 = m3 & rest;
        if ((size_t)pos < cap) items[pos] = it;
      }
    }
    __syncthreads();
    if (t < NBUCKET) {
      int tot = 0;
      for (int w2 = 0; w2 < ABT / 64; w2++) tot += wcnt[w2][t];
      bk_run[t] += tot;
    }
    __syncthreads();
  }
  }
}

// (nothing when the predicate pass dropped records: the items are then
// incomplete and the build is redone)
__global__ void k_ajcount(const AItem *__restrict__ items, const int32_t *__restrict__ nitem,
                          size_t cap, const int32_t *__restrict__ dropped,
                          int32_t *__restrict__ cnt) {
  const size_t n = *dropped ? 0 : min((size_t)*nitem, cap);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[items[q].j], 1);
}
// (thread 0 also copies the build's item and chunk totals into the status
// words the host reads at the end)
__global__ void k_ajfill(const AItem *__restrict__ items, const int32_t *__restrict__ nitem,
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
// real-space energy over the atom items

template <bool F64>
__device__ __forceinline__ void atoms_impl(
    const int32_t *__restrict__ chunks, const AItem *__restrict__ items,
    const double *__restrict__ coor, const int32_t *__restrict__ sslot2atom, int nsol,
    const double *__restrict__ q, const double *__restrict__ e12, const double *__restrict__ e6,
    const double *__restrict__ erf_arr, const ARS &P, const int32_t *__restrict__ jpos,
    long long *__restrict__ jpart,
    long long *__restrict__ ipart, double *__restrict__ epart, int max_parts, int part0, int bid) {
  __shared__ double ix[WBLOCK][3];
  __shared__ double iq[WBLOCK], ia12[WBLOCK], ib6[WBLOCK];
  __shared__ unsigned long long iacc[WBLOCK * 3];
  __shared__ double red[16];
  const int ch = bid, t = threadIdx.x;
  const int blk = chunks[3 * ch], beg = chunks[3 * ch + 1], end = chunks[3 * ch + 2];
  const int s0 = blk * WBLOCK;
  if (t < WBLOCK) {
    const int a = (s0 + t < nsol) ? sslot2atom[s0 + t] : -1;
    if (a >= 0) {
      ix[t][0] = coor[3 * a];
      ix[t][1] = coor[3 * a + 1];
      ix[t][2] = coor[3 * a + 2];
      iq[t] = q[a] * P.qfac;
      ia12[t] = e12[a] * P.pick;
      ib6[t] = P.arith ? e6[a] : e6[a] * P.pick;  // sigma under arith
    } else {
      ix[t][0] = ix[t][1] = ix[t][2] = 0.0;
      iq[t] = ia12[t] = ib6[t] = 0.0;
    }
  }
  for (int u = t; u < WBLOCK * 3; u += ABT) iacc[u] = 0;
  __syncthreads();
  bool bad = false;  // a term beyond the fixed-point range (or a NaN)
  const float betaf = (float)P.beta, c2b = (float)(2.0 / sqrt(M_PI) * P.beta);
  const float cutvdw2f = (float)P.cutvdw2, cutele2f = (float)P.cutele2;
  const float log2e = 1.4426950408889634f;
  double ev_r = 0, ee_r = 0, ev_i = 0, ee_i = 0;
  for (int it = beg + t; it < end; it += ABT) {
    // the mask walk starts at a member given by the item, so the order of
    // an item's sums does not depend on the lane or chunk that runs it
    const int rot = it & 63;
    const AItem item = items[it];
    const int ja = item.j;
    int s;
    bool rev;
    dir_of(item.s, s, rev);
    int sa, sb, sc;
    symop_of(s, &sa, &sb, &sc);
    const double T[3] = {sa * P.a, sb * P.b, sc * P.c};
    double xj[3] = {coor[3 * ja], coor[3 * ja + 1], coor[3 * ja + 2]};
    // reverse images: the partner is the translated member, x_j + T
    if (rev)
      for (int k = 0; k < 3; k++) xj[k] = xj[k] + T[k];
    const double qj = q[ja], aj = e12[ja], bj = e6[ja];
    double fj[3] = {0, 0, 0};
    double ev = 0, ee = 0;
    const uint64_t m2 = item.m2, m3 = item.m3;
    uint64_t mask = rotr64(item.m1 | m2 | m3, rot);
    while (mask) {
      const int i = (__ffsll((long long)mask) - 1 + rot) & 63;
      mask &= mask - 1;
      const int kind = ((m2 >> i) & 1ull) ? 2 : (((m3 >> i) & 1ull) ? 3 : 1);
      // d = x_l - x_m with the translation on the image member
      double d[3];
      if (s > 0 && !rev) {
#pragma clang fp contract(off)
        for (int k = 0; k < 3; k++) d[k] = (ix[i][k] + T[k]) - xj[k];
      } else {
        for (int k = 0; k < 3; k++) d[k] = ix[i][k] - xj[k];
      }
      const double qq = iq[i] * qj;
      double A, B;
      lj_coeffs(P.arith, ia12[i], ib6[i], aj, bj, A, B);
      double df = 0.0, e_v = 0.0, e_e = 0.0;
      if (F64) {
#pragma clang fp contract(off)
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double s2 = 1.0 / r2;
        const bool lj = (kind != 3) && r2 <= P.cutvdw2;
        const bool el = (kind == 1) || (kind == 3 && r2 <= P.cutele2);
        if (kind == 2 && lj) {
          const double s6 = s2 * s2 * s2, A12 = A * s6 * s6, B6 = B * s6;
          e_v = A12 - B6;
          df = -6.0 * s2 * (A12 + e_v);
        } else if (el) {
          const double rij = sqrt(r2), sv = 1.0 / rij;
          double erfcc, derfc;
          erfc_table(erf_arr, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
          e_e = qq * sv * erfcc;
          df = -qq * sv * (s2 * erfcc + sv * P.beta * derfc);
          if (kind == 1 && lj) {
            const double s6 = s2 * s2 * s2, A12 = A * s6 * s6, B6 = B * s6;
            e_v = A12 - B6;
            df += -6.0 * s2 * (A12 + e_v);
          }
        }
      } else {
        const float dx = (float)d[0], dy = (float)d[1], dz = (float)d[2];
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float sv = __builtin_amdgcn_rsqf(r2);
        const float s2 = sv * sv;
        const bool lj = (kind != 3) && r2 <= cutvdw2f;
        const bool el = (kind == 1) || (kind == 3 && r2 <= cutele2f);
        float dff = 0.f;
        if (el) {
          const float qqf = (float)qq;
          const float x = betaf * (r2 * sv);
          const float tt = __builtin_amdgcn_rcpf(fmaf(ERFC_P, x, 1.0f));
          const float ex = __builtin_amdgcn_exp2f(-(x * x) * log2e);
          const float e = qqf * (ex * erfc_poly(tt)) * sv;
          e_e = (double)e;
          dff = -s2 * (e + qqf * c2b * ex);
        }
        if (lj) {
          const float s6 = s2 * s2 * s2;
          const float A12 = (float)A * s6 * s6, B6 = (float)B * s6;
          const float e1 = A12 - B6;
          e_v = (double)e1;
          dff += -6.0f * s2 * (A12 + e1);
        }
        df = (double)dff;
      }
      ev += e_v;
      ee += e_e;
      // gradient: +df d on the block atom, -df d on the partner
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double g = df * d[k];
        fj[k] -= g;
        const long long v = F64 ? fixd(g, P.fix_scale, bad) : fixf((float)g, bad);
        if (v) atomicAdd(&iacc[i * 3 + k], (unsigned long long)v);
      }
    }
    long long *jrow = jpart + (size_t)jpos[it] * 3;
#pragma unroll
    for (int k = 0; k < 3; k++)
      jrow[k] = F64 ? fixd(fj[k], P.fix_scale, bad) : fixf((float)fj[k], bad);
    if (s == 0) {
      ev_r += ev;
      ee_r += ee;
    } else {
      ev_i += ev;
      ee_i += ee;
    }
  }
  fix_report(bad, P.ovf);
  __syncthreads();
  for (int u = t; u < WBLOCK * 3; u += ABT) ipart[(size_t)ch * (WBLOCK * 3) + u] = (long long)iacc[u];
  const double v = wg_sum4(ev_r, ee_r, ev_i, ee_i, red);
  if (t < 4) epart[(t == 0 ? EN_VDW : t == 1 ? EN_EL : t == 2 ? EN_VSYM : EN_LSYM) * max_parts +
                   part0 + ch] = v;
}

template <bool F64>
__device__ __forceinline__ void atoms_body(const AtomArgs &A, int bid, int nblk) {
  atoms_impl<F64>(A.chunks, A.items, A.coor, A.sslot2atom, A.nsol, A.q, A.e12, A.e6, A.erf, A.P,
                  A.jpos, A.jpart, A.ipart, A.epart, A.max_parts, A.part0, bid);
}
DEFINE_KERNEL(k_atoms_f32, AtomArgs, ABT, atoms_body<false>)
DEFINE_KERNEL(k_atoms_f64, AtomArgs, ABT, atoms_body<true>)

// one thread per (atom, component): the atom's partner rows (contiguous)
// and, for a block atom, its entries in its block's chunk rows, all
// 64-bit fixed point, summed as integers and added to its accumulator
__device__ __forceinline__ void atoms_reduce_impl(
    const int32_t *__restrict__ jstart, const long long *__restrict__ jpart,
    const int32_t *__restrict__ slot_of, const int32_t *__restrict__ cfirst,
    const long long *__restrict__ ipart, int npt, long long *__restrict__ facc, int bid) {
  const int g = bid * 256 + threadIdx.x;
  if (g >= 3 * npt) return;
  const int a = g / 3, k = g - 3 * a;
  long long acc = 0;
  const int q1 = jstart[a + 1];
  for (int q = jstart[a]; q < q1; q++) acc += jpart[(size_t)q * 3 + k];
  const int s = slot_of[a];
  if (s >= 0) {
    const int b = s / WBLOCK, i = s % WBLOCK;
    for (int c = cfirst[b]; c < cfirst[b + 1]; c++) acc += ipart[(size_t)c * (WBLOCK * 3) + i * 3 + k];
  }
  if (acc) facc[3 * (size_t)a + k] += acc;
}
__device__ __forceinline__ void atoms_reduce_body(const AReduceArgs &A, int bid, int nblk) {
  atoms_reduce_impl(A.jstart, A.jpart, A.slot_of, A.cfirst, A.ipart, A.npt, A.facc, bid);
}
DEFINE_KERNEL(k_atoms_reduce, AReduceArgs, 256, atoms_reduce_body)

// ---------------------------------------------------------------------------
int32_t atoms_init(moil_ctx *c, const moil_topology *t) {
  const moil_params &P = c->par;
  std::vector<int32_t> mono_of(c->npt), aflag(c->npt, 0);
  std::vector<uint64_t> xmask(c->npt, 0);
  std::vector<int32_t> xfar(c->npt, 0);
  for (int m = 0; m < t->totmon; m++)
    for (int a = t->poipt[m]; a < t->poipt[m + 1]; a++) {
      mono_of[a] = m;
      if (t->monoflag[m] & MOIL_MONO_WATER) aflag[a] |= AF_WAT;
      else c->h_sol.push_back(a);
    }
  for (int a = 0; a < c->npt; a++) {
    if (t->flagchr[a]) aflag[a] |= AF_CHG;
    // the hydrogen rule is off under the arith rule (pot/nbond.f:115-117,
    // symm/nbsym.f:104-106)
    if (t->epsgm12[a] < 1e-3 && !(P.terms & MOIL_TERM_ARITH)) aflag[a] |= AF_SMALL;
    for (int k = t->exc1[a]; k < t->exc1[a + 1]; k++) {
      const int m = t->exc2[k] - 1;
      if (m > a && m - a <= 64) xmask[a] |= 1ull << (m - a - 1);
      else xfar[a] = 1;
    }
  }
  c->nsol = (int)c->h_sol.size();
  HIP_CHECK(hipMalloc(&c->d_mono_of, sizeof(int32_t) * std::max(c->npt, 1)));
  HIP_CHECK(hipMemcpy(c->d_mono_of, mono_of.data(), sizeof(int32_t) * c->npt, hipMemcpyHostToDevice));
  HIP_CHECK(hipMalloc(&c->d_aflag, sizeof(int32_t) * std::max(c->npt, 1)));
  HIP_CHECK(hipMemcpy(c->d_aflag, aflag.data(), sizeof(int32_t) * c->npt, hipMemcpyHostToDevice));
  if (c->nsol == 0) return MOIL_GPU_OK;
  HIP_CHECK(hipMalloc(&c->d_xmask, sizeof(uint64_t) * c->npt));
  HIP_CHECK(hipMemcpy(c->d_xmask, xmask.data(), sizeof(uint64_t) * c->npt, hipMemcpyHostToDevice));
  HIP_CHECK(hipMalloc(&c->d_xfar, sizeof(int32_t) * c->npt));
  HIP_CHECK(hipMemcpy(c->d_xfar, xfar.data(), sizeof(int32_t) * c->npt, hipMemcpyHostToDevice));
  const int nexc = t->exc1[c->npt];
  HIP_CHECK(hipMalloc(&c->d_exc1, sizeof(int32_t) * (c->npt + 1)));
  HIP_CHECK(hipMemcpy(c->d_exc1, t->exc1, sizeof(int32_t) * (c->npt + 1), hipMemcpyHostToDevice));
  HIP_CHECK(hipMalloc(&c->d_exc2, sizeof(int32_t) * std::max(nexc, 1)));
  if (nexc) HIP_CHECK(hipMemcpy(c->d_exc2, t->exc2, sizeof(int32_t) * nexc, hipMemcpyHostToDevice));
  HIP_CHECK(hipMalloc(&c->d_sol, sizeof(int32_t) * c->nsol));
  HIP_CHECK(hipMemcpy(c->d_sol, c->h_sol.data(), sizeof(int32_t) * c->nsol, hipMemcpyHostToDevice));
  // all-atom cell grid, ~8 atoms per cell
  const double L[3] = {P.a, P.b, P.c};
  const double edge = cbrt(P.a * P.b * P.c * 8.0 / (double)c->npt);
  for (int k = 0; k < 3; k++) {
    c->ag[k] = std::max(1, (int)floor(L[k] / edge));
    c->acs[k] = L[k] / c->ag[k];
  }
  c->ancell = ((c->ag[0] + 1) / 2) * ((c->ag[1] + 1) / 2) * ((c->ag[2] + 1) / 2) * 8;
  HIP_CHECK(hipMalloc(&c->d_acell_of, sizeof(int32_t) * c->npt));
  HIP_CHECK(hipMalloc(&c->d_acnt, sizeof(int32_t) * 4 * (size_t)(c->ancell + 1)));
  HIP_CHECK(hipMalloc(&c->d_astart, sizeof(int32_t) * (c->ancell + 1)));
  HIP_CHECK(hipMalloc(&c->d_aslot2atom, sizeof(int32_t) * c->npt));
  HIP_CHECK(hipMalloc(&c->d_sstart, sizeof(int32_t) * (c->ancell + 1)));
  // solute slots: each brick row padded to whole blocks
  c->nrow = ((c->ag[1] + 1) / 2) * ((c->ag[2] + 1) / 2);
  c->rowcells = ((c->ag[0] + 1) / 2) * 8;
  c->nablk = (c->nsol + (WBLOCK - 1) * c->nrow + WBLOCK - 1) / WBLOCK;
  c->nsslot = c->nablk * WBLOCK;
  HIP_CHECK(hipMalloc(&c->d_sslot_u, sizeof(int32_t) * c->nsol));
  HIP_CHECK(hipMalloc(&c->d_sslot2atom, sizeof(int32_t) * c->nsslot));
  HIP_CHECK(hipMalloc(&c->d_rowcnt, sizeof(int32_t) * (c->nrow + 1)));
  HIP_CHECK(hipMalloc(&c->d_rowoff, sizeof(int32_t) * (c->nrow + 1)));
  HIP_CHECK(hipMalloc(&c->d_sslot_of, sizeof(int32_t) * c->npt));
  // parts per block for the predicate pass: at least ~2048 workgroups
  c->anpart = std::min(8, std::max(1, (2048 + c->nablk - 1) / std::max(c->nablk, 1)));
  HIP_CHECK(hipMalloc(&c->d_abpart, sizeof(int32_t) * ((size_t)c->nablk * c->anpart * NBUCKET + 1)));
  HIP_CHECK(hipMalloc(&c->d_abcount, sizeof(int32_t) * (c->nablk * NBUCKET + 1)));
  HIP_CHECK(hipMalloc(&c->d_aboff, sizeof(int32_t) * (c->nablk * NBUCKET + 1)));
  HIP_CHECK(hipMalloc(&c->d_ajcnt, sizeof(int32_t) * (c->npt + 1)));
  HIP_CHECK(hipMalloc(&c->d_ajfill, sizeof(int32_t) * (c->npt + 1)));
  HIP_CHECK(hipMalloc(&c->d_ajstart, sizeof(int32_t) * (c->npt + 1)));
  HIP_CHECK(hipMalloc(&c->d_acfirst, sizeof(int32_t) * (c->nablk + 1)));
  HIP_CHECK(hipMalloc(&c->d_arcount, sizeof(int32_t) * ((size_t)c->nablk * c->anpart + 1)));
  return MOIL_GPU_OK;
}

void atoms_free(moil_ctx *c) {
  void *ptrs[] = {c->d_mono_of, c->d_aflag, c->d_xmask, c->d_xfar, c->d_exc1, c->d_exc2,
                  c->d_sol, c->d_acell_of, c->d_acnt, c->d_astart, c->d_aslot2atom,
                  c->d_sstart, c->d_sslot2atom, c->d_sslot_of, c->d_abcount, c->d_aboff,
                  c->d_aitems, c->d_achunk, c->d_ajpart, c->d_aipart, c->d_ajcnt, c->d_ajfill, c->d_ajstart,
                  c->d_ajpos, c->d_acfirst, c->d_araw, c->d_arcount, c->d_sslot_u,
                  c->d_rowcnt, c->d_rowoff, c->d_account, c->d_abpart};
  for (void *p : ptrs)
    if (p) hipFree(p);
}

int32_t atoms_build(moil_ctx *c, const double *d_coor, hipStream_t st) {
  c->naitems = 0;
  c->nachunk = 0;
  if (c->nsol == 0) return MOIL_GPU_OK;
  const moil_params &P = c->par;
  const int TB = 256;
  Grid g;
  g.gx = c->ag[0]; g.gy = c->ag[1]; g.gz = c->ag[2];
  g.nbx = (g.gx + 1) / 2; g.nby = (g.gy + 1) / 2; g.nbz = (g.gz + 1) / 2;
  g.g0x = -0.5 * P.a; g.g0y = -0.5 * P.b; g.g0z = -0.5 * P.c;
  g.csx = c->acs[0]; g.csy = c->acs[1]; g.csz = c->acs[2];
  // the build's counters (four cell-count passes, the partner CSR) and
  // the padding markers, set in one launch
  const size_t nc1 = (size_t)c->ancell + 1;
  int32_t *cnt0 = c->d_acnt, *cnt1 = cnt0 + nc1, *cnt2 = cnt1 + nc1, *cnt3 = cnt2 + nc1;
  {
    ClearSet cs;
    cs.add(c->d_acnt, (int)(4 * nc1), 0);
    cs.add(c->d_ajcnt, c->npt + 1, 0);
    cs.add(c->d_ajfill, c->npt + 1, 0);
    cs.add(c->d_sslot_of, c->npt, -1);
    cs.add(c->d_sslot2atom, c->nsslot, -1);
    launch_clear(cs, st);
  }
  // all atoms by cell (partner side)
  k_atom_cells<<<(c->npt + TB - 1) / TB, TB, 0, st>>>(d_coor, nullptr, c->npt, g, c->d_acell_of,
                                                      cnt0);
  CHECK_SCAN(scan_excl(c, cnt0, c->d_astart, c->ancell, st));
  k_atom_scatter<<<(c->npt + TB - 1) / TB, TB, 0, st>>>(c->d_acell_of, nullptr, c->npt,
                                                        c->d_astart, cnt1, c->d_aslot2atom);
  launch_cell_sort(c->d_astart, c->ancell, c->d_aslot2atom, st);
  // solute atoms by cell (block side)
  k_atom_cells<<<(c->nsol + TB - 1) / TB, TB, 0, st>>>(d_coor, c->d_sol, c->nsol, g, c->d_acell_of,
                                                       cnt2);
  CHECK_SCAN(scan_excl(c, cnt2, c->d_sstart, c->ancell, st));
  k_atom_scatter<<<(c->nsol + TB - 1) / TB, TB, 0, st>>>(c->d_acell_of, c->d_sol, c->nsol,
                                                         c->d_sstart, cnt3, c->d_sslot_u);
  launch_cell_sort(c->d_sstart, c->ancell, c->d_sslot_u, st);
  // blocks end with their brick row (padding slots -1)
  k_row_pad<<<(c->nrow + TB - 1) / TB, TB, 0, st>>>(c->d_sstart, c->nrow, c->rowcells, c->d_rowcnt);
  CHECK_SCAN(scan_excl(c, c->d_rowcnt, c->d_rowoff, c->nrow, st));
  k_pad_slots<<<(c->nsol + TB - 1) / TB, TB, 0, st>>>(c->d_sslot_u, c->nsol, c->d_sstart,
                                                      c->d_rowoff, c->nrow, c->rowcells,
                                                      c->d_sslot2atom);
  k_slot_of<<<(c->nsslot + TB - 1) / TB, TB, 0, st>>>(c->d_sslot2atom, c->nsslot, c->d_sslot_of);

  APar ap;
  ap.a = P.a; ap.b = P.b; ap.c = P.c;
  ap.ha = 0.5 * P.a; ap.hb = 0.5 * P.b; ap.hc = 0.5 * P.c;
  ap.cutmono2 = P.cutmono2; ap.cutvbig2 = P.cutvbig2; ap.cutebig2 = P.cutebig2;
  ap.R = sqrt(std::max(std::max(P.cutvbig2, P.cutebig2), 0.0)) * (1.0 + 1e-12) + 1e-9;
  ap.hvdw0 = (P.terms & MOIL_TERM_HVDW0) ? 1 : 0;
  ap.nsol = c->nsslot;
  // enqueued without waiting, like the water lists: capacities from earlier
  // builds, overflows reported in the status words (ST_A*)
  if (!c->d_araw) {
    c->araw_cap = std::max(c->araw_cap, 1024);
    HIP_CHECK(hipMalloc(&c->d_araw, sizeof(AItem) * (size_t)c->araw_cap * c->nablk * c->anpart));
  }
  if (!c->d_aitems) {
    c->aitem_cap = std::max<size_t>(c->aitem_cap, 4096);
    HIP_CHECK(hipMalloc(&c->d_aitems, sizeof(AItem) * c->aitem_cap));
    HIP_CHECK(hipMalloc(&c->d_ajpos, sizeof(int32_t) * c->aitem_cap));
    HIP_CHECK(hipMalloc(&c->d_ajpart, sizeof(long long) * 3 * c->aitem_cap));
  }
  if (!c->d_achunk) {
    c->achunk_cap = std::max(c->achunk_cap, 1024);
    HIP_CHECK(hipMalloc(&c->d_achunk, sizeof(int32_t) * 3 * (size_t)c->achunk_cap));
  }
  if (!c->d_account) HIP_CHECK(hipMalloc(&c->d_account, sizeof(int32_t) * (c->nablk + 1)));
  const int np = c->anpart;
  k_ablock_raw<<<c->nablk * np, ABT, 0, st>>>(
      d_coor, c->d_sslot2atom, c->d_aslot2atom, c->d_astart, g, c->d_mono_of, c->d_aflag,
      c->d_mono_cent, c->d_flags, c->d_xmask, c->d_xfar, c->d_exc1, c->d_exc2, ap, c->d_abpart,
      c->d_araw, c->araw_cap, c->d_arcount, c->d_status + ST_ARAW, np);
  k_abucket_sum<<<(c->nablk * NBUCKET + TB - 1) / TB, TB, 0, st>>>(c->d_abpart, c->nablk, np,
                                                                  c->d_abcount);
  CHECK_SCAN(scan_excl(c, c->d_abcount, c->d_aboff, c->nablk * NBUCKET, st));
  k_ablock_write<<<c->nablk * np, ABT, 0, st>>>(c->d_araw, c->araw_cap, c->d_arcount, np,
                                                c->d_abpart, c->d_aboff, c->d_aitems,
                                                c->aitem_cap);
  ChunkPar cp;
  cp.nblk = c->nablk;
  cp.pairs = c->achunk_pairs;
  cp.max_items = c->wchunk_max_items;
  cp.min_items = 1;
  cp.cap = c->achunk_cap;
  cp.nbucket = NBUCKET;
  cp.m2_div = 1;
  CHECK_SCAN(chunk_table(c, c->d_aboff, cp, c->d_account, c->d_acfirst, c->d_achunk, nullptr,
                         nullptr, st));
  const int32_t *d_nitem = c->d_aboff + (size_t)c->nablk * NBUCKET;
  k_ajcount<<<1024, TB, 0, st>>>(c->d_aitems, d_nitem, c->aitem_cap, c->d_status + ST_ARAW,
                                 c->d_ajcnt);
  CHECK_SCAN(scan_excl(c, c->d_ajcnt, c->d_ajstart, c->npt, st));
  k_ajfill<<<1024, TB, 0, st>>>(c->d_aitems, d_nitem, c->aitem_cap, c->d_status + ST_ARAW,
                                c->d_ajstart, c->d_ajfill, c->d_ajpos, c->d_acfirst + c->nablk,
                                c->d_status + ST_AITEMS, c->d_status + ST_ACHUNK);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

bool atoms_grow(moil_ctx *c, const int32_t *status) {
  bool again = false;
  if (c->nsol == 0) return false;
  if (status[ST_ARAW] > c->araw_cap) {
    hipFree(c->d_araw);
    c->d_araw = nullptr;
    c->araw_cap = status[ST_ARAW] + status[ST_ARAW] / 4 + 256;
    again = true;
  }
  if ((size_t)status[ST_AITEMS] > c->aitem_cap) {
    hipFree(c->d_aitems);
    hipFree(c->d_ajpos);
    hipFree(c->d_ajpart);
    c->d_aitems = nullptr;
    c->d_ajpos = nullptr;
    c->d_ajpart = nullptr;
    c->aitem_cap = (size_t)status[ST_AITEMS] + status[ST_AITEMS] / 4 + 1024;
    again = true;
  }
  if (status[ST_ACHUNK] > c->achunk_cap) {
    hipFree(c->d_achunk);
    c->d_achunk = nullptr;
    c->achunk_cap = status[ST_ACHUNK] + status[ST_ACHUNK] / 4 + 64;
    again = true;
  }
  return again;
}

int32_t atoms_finish(moil_ctx *c, const int32_t *status) {
  if (c->nsol == 0) {
    c->naitems = 0;
    c->nachunk = 0;
    return MOIL_GPU_OK;
  }
  c->naitems = (size_t)status[ST_AITEMS];
  c->nachunk = status[ST_ACHUNK];
  if (c->nachunk > c->aipart_cap || c->d_aipart == nullptr) {
    hipFree(c->d_aipart);
    c->aipart_cap = std::max(c->nachunk + c->nachunk / 4 + 64, 256);
    HIP_CHECK(hipMalloc(&c->d_aipart, sizeof(long long) * WBLOCK * 3 * (size_t)c->aipart_cap));
  }
  return MOIL_GPU_OK;
}

// the atom real-space launches of one evaluation (none with an empty
// list); the energy partials follow the water kernel's (part0)
void atoms_plan(const moil_ctx *c, const double *d_coor, int part0, EvalPlan *p) {
  p->n_atoms = p->n_ared = 0;
  if (c->nachunk == 0) return;
  ARS P;
  const uint32_t T = c->par.terms;
  P.qfac = (T & MOIL_TERM_EL) ? KOFDIE / c->par.eps : 0.0;
  P.pick = (T & MOIL_TERM_VDW) ? 1.0 : 0.0;
  P.arith = (T & MOIL_TERM_ARITH) ? 1 : 0;
  P.beta = c->beta;
  P.cutvdw2 = c->par.cutvdw2;
  P.cutele2 = c->par.cutele2;
  P.erftbdns = c->erftbdns;
  P.ntab = c->ntab;
  P.a = c->par.a; P.b = c->par.b; P.c = c->par.c;
  P.fix_scale = c->fix_scale;
  P.ovf = c->d_ovf;
  AtomArgs &A = p->atoms;
  A.chunks = c->d_achunk;
  A.items = c->d_aitems;
  A.coor = d_coor;
  A.sslot2atom = c->d_sslot2atom;
  A.nsol = c->nsslot;
  A.q = c->d_ptchg;
  A.e12 = c->d_eps12;
  A.e6 = c->d_eps6;
  A.erf = c->d_erf;
  A.P = P;
  A.jpos = c->d_ajpos;
  A.jpart = c->d_ajpart;
  A.ipart = c->d_aipart;
  A.epart = c->d_epart;
  A.max_parts = c->max_parts;
  A.part0 = part0;
  p->n_atoms = c->nachunk;
  AReduceArgs &R = p->ared;
  R.jstart = c->d_ajstart;
  R.jpart = c->d_ajpart;
  R.slot_of = c->d_sslot_of;
  R.cfirst = c->d_acfirst;
  R.ipart = c->d_aipart;
  R.npt = c->npt;
  R.facc = c->d_facc;
  p->n_ared = (3 * c->npt + 255) / 256;
}

void launch_k_atoms_ev(const AtomArgs &a, bool f64, int nblk, hipStream_t st, hipEvent_t e0,
                       hipEvent_t e1) {
  if (nblk <= 0) return;
  if (f64)
    hipExtLaunchKernelGGL(k_atoms_f64, dim3(nblk), dim3(ABT), 0, st, e0, e1, 0, a);
  else
    hipExtLaunchKernelGGL(k_atoms_f32, dim3(nblk), dim3(ABT), 0, st, e0, e1, 0, a);
}
