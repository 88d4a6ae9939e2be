// Smooth particle-mesh Ewald reciprocal sum for MOIL
// (pot/ewald_receip.f:8-213 do_pmesh_kspace): order-4 B-spline charge
// spreading, forward 3-D FFT, influence-function convolution with the
// energy sum, inverse FFT and force gather.
//
// Conventions kept from MOIL: box centred at the origin with
// fr = n (w - anint(w) + 0.5), w = x/a (pot/ewald_receip.f:250-260);
// stencil points int(fr)-3 .. int(fr) with periodic wrap
// (pot/ewald_receip.f:637-656); charges scaled by sqrt(kofdie/eps);
// B-spline moduli with the <1e-7 neighbour fix (pot/ewald_receip.f:
// 455-481); E = 1/2 sum_{m != 0} eterm(m) |S(m)|^2 with
// eterm = exp(-pi^2 m^2/beta^2) / (pi V |b1 b2 b3|^2 m^2)
// (pot/ewald_receip.f:723-812).
//
// GPU design: at every list build the atoms are counting-sorted by grid
// tile (8x8x8 points; the tile holding the atom's stencil base point
// int(fr)), and the assignment is kept until the next build.  Each
// evaluation runs one workgroup per tile: its atoms' B-splines go to LDS,
// then every atom adds its 64 weighted charges (one lane per stencil
// point) into the tile's region in LDS -- the tile, the 3-point stencil
// reach and a margin of PME_MARGIN base points on each side, so an atom
// that has moved up to that far since the build still lands inside --
// with 64-bit fixed-point atomics, which are order independent
// (bit-reproducible).  An atom that moved further adds its stencil to a
// fixed-point overflow grid with device atomics instead (the same
// integers, so the sum is unchanged).  A per-point pass sums the tile
// regions covering each point and the overflow grid and writes the float
// (mixed) or double (FP64) grid for rocFFT's real-to-complex transform --
// the only library call.  The convolution runs on the half spectrum with
// the influence function tabulated once in FP64 at init; the reciprocal
// energy is summed in FP64 with self-conjugate planes weighted once and
// all others twice.  The gather runs one workgroup per tile with the
// tile's region of the convolved grid in LDS, recomputes the B-splines
// and derivatives from the coordinates (an atom outside the region reads
// the grid itself), and writes each atom's reciprocal gradient directly
// (no atomics).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "multi.h"

#define PME_THREADS 256
#define QFIX_SHIFT 40  // grid charge fixed point: 2^-40
// round(v 2^40) for |v| < 2^11 as the bit pattern of v + 1.5 2^12 (whose
// last mantissa bit is 2^-40) minus that of 1.5 2^12: one add and one
// 64-bit subtract in place of a double-to-integer conversion.  Every
// stencil weight q sqrt(kofdie/eps) theta1 theta2 theta3 is below
// max |q| sqrt(kofdie/eps), which pme_init holds below 2^11.
static_assert(QFIX_SHIFT == 40, "qfix's bias is that of 2^-40 units");
#define QFIX_BIAS 6144.0  // 1.5 2^12
#define QFIX_MAXW 2048.0
__device__ __forceinline__ long long qfix(double v) {
  return (long long)((unsigned long long)__double_as_longlong(v + QFIX_BIAS) -
                     (unsigned long long)__double_as_longlong(QFIX_BIAS));
}
// The reciprocal-space kernels run beside the real-space pair kernel (on
// the context's second stream) and form the evaluation's longer chain:
// their waves take the SIMDs' issue slots first.
#define PME_WAVE_PRIORITY() __builtin_amdgcn_s_setprio(2)
// base points a tile's region accepts beyond the tile on each side, and
// the largest region edge (tile <= 8 points + the stencil's 3 + margins)
#define PME_MARGIN 1
#define PME_RMAX (8 + 3 + 2 * PME_MARGIN)


// order-4 cardinal B-spline weights and derivatives at fractional offset
// w (pot/ewald_receip.f:1013-1080: linear init, one recursion to order 3,
// derivative from order 3, final recursion to order 4)
template <typename R>
__device__ __forceinline__ R pick4(const R t[4], int a) {
  return a == 0 ? t[0] : (a == 1 ? t[1] : (a == 2 ? t[2] : t[3]));
}

template <typename R>
__device__ __forceinline__ void bspline4(R w, R th[4], R dth[4]) {
  R c1 = (R)1 - w, c2 = w;
  // order 3
  R d3 = (R)0.5;
  R o3 = d3 * w * c2;
  R o2 = d3 * ((w + (R)1) * c1 + ((R)2 - w) * c2);
  R o1 = d3 * ((R)1 - w) * c1;
  dth[0] = -o1;
  dth[1] = o1 - o2;
  dth[2] = o2 - o3;
  dth[3] = o3;
  // order 4
  R d4 = (R)1 / (R)3;
  th[3] = d4 * w * o3;
  th[2] = d4 * ((w + (R)1) * o2 + ((R)3 - w) * o3);
  th[1] = d4 * ((w + (R)2) * o1 + ((R)2 - w) * o2);
  th[0] = d4 * ((R)1 - w) * o1;
}

__device__ __forceinline__ void frac_coord(double x, double r, int n, int *i0, double *w) {
  double u = x * r;
  double f = (double)n * (u - round(u) + 0.5);
  int fi = (int)f;
  *w = f - (double)fi;
  *i0 = fi - 3;  // first stencil point int(fr)-3
}

// Grid tiles: dimension k is cut into nt_k = ceil(n_k/8) tiles with
// boundaries lo = floor(t n_k / nt_k) (sizes 4..8 for n_k >= 4, so a
// tile's 3-point low-side halo reaches one neighbour only).  An atom
// belongs to the tile holding its stencil base point int(fr); its 64
// stencil points base-3 .. base lie in that tile or its halo.
// (TilePar: r1 r2 r3 = LDS region dims, max tile size + 3 + 2 margins;
// rsz = r1 r2 r3)

__device__ __forceinline__ int tile_of(int g, int n, int nt) { return ((g + 1) * nt - 1) / n; }
__device__ __forceinline__ int tile_lo(int t, int n, int nt) { return (t * n) / nt; }

// Region of tile [lo, hi) in one dimension: grid points lo - M - 3 ..
// hi + M - 1 (periodic), the stencils of base points lo - M .. hi + M - 1.
// local_base gives an atom's base point c relative to the region (the
// region index of its first stencil point c - 3), using whichever
// periodic image of c lies in the accepted range, or -1 if none does.
__device__ __forceinline__ int local_base(int c, int lo, int hi, int n) {
#pragma unroll
  for (int k = -1; k <= 1; k++) {
    const int b = c + k * n;
    if (b >= lo - PME_MARGIN && b < hi + PME_MARGIN) return b - lo + PME_MARGIN;
  }
  return -1;
}

__device__ __forceinline__ void base_point(const double *x, const PmePar &P, int &c1, int &c2,
                                           int &c3, double &w1, double &w2, double &w3) {
  int i0, j0, k0;
  frac_coord(x[0], P.ra, P.n1, &i0, &w1);
  frac_coord(x[1], P.rb, P.n2, &j0, &w2);
  frac_coord(x[2], P.rc, P.n3, &k0, &w3);
  // base point int(fr) = i0 + 3, in [0, n] (n only when fr == n exactly)
  c1 = i0 + 3; c2 = j0 + 3; c3 = k0 + 3;
  if (c1 >= P.n1) c1 -= P.n1;
  if (c2 >= P.n2) c2 -= P.n2;
  if (c3 >= P.n3) c3 -= P.n3;
}

// tile of every atom and its rank inside the tile (counting sort).  Atoms
// are visited in the previous evaluation's tile order, so the lanes of a
// wavefront mostly share a tile, and each wavefront issues one counter
// atomic per distinct tile (ranks inside a tile may come out in any
// order: the spreading sums are integer and the gather is per atom)
__device__ __forceinline__ void tbin_impl(const double *__restrict__ coor, int npt,
                                          const int32_t *__restrict__ order, const PmePar &P,
                                          const TilePar &T, int32_t *__restrict__ tcnt,
                                          int32_t *__restrict__ atile,
                                          int32_t *__restrict__ arank, int bid) {
  const int p = bid * PME_THREADS + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = p < npt;
  const int n = valid ? order[p] : 0;
  int tile = -1;
  if (valid) {
    int c1, c2, c3;
    double w1, w2, w3;
    base_point(coor + 3 * n, P, c1, c2, c3, w1, w2, w3);
    tile = (tile_of(c3, T.n3, T.nt3) * T.nt2 + tile_of(c2, T.n2, T.nt2)) * T.nt1 +
           tile_of(c1, T.n1, T.nt1);
    atile[n] = tile;
  }
  uint64_t todo = __ballot(valid);
  const uint64_t below = (lane == 0) ? 0ull : ((1ull << lane) - 1ull);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lt = __shfl(tile, leader, 64);
    const uint64_t same = __ballot(valid && tile == lt);
    int base = 0;
    if (lane == leader) base = atomicAdd(&tcnt[lt], __popcll(same));
    base = __shfl(base, leader, 64);
    if (valid && tile == lt) arank[n] = base + __popcll(same & below);
    todo &= ~same;
  }
}
// With at most TSCAN_FUSED tiles the last workgroup of the binning pass
// scans the counts itself (saving the dispatch of a one-workgroup kernel
// behind the real-space work): every workgroup takes a ticket after its
// counter atomics have returned; the last one reads the counts with
// agent-scope atomic loads (the counts live at the coherence point, where
// the atomics performed them), writes the exclusive scan and clears the
// counts and the ticket with agent-scope atomic stores -- no fence that
// would write back the L2.
#define TSCAN_FUSED 4096
__device__ __forceinline__ void tscan_last(int32_t *__restrict__ cnt, int n,
                                           int32_t *__restrict__ start) {
  __shared__ int part[PME_THREADS];
  const int t = threadIdx.x;
  const int per = (n + PME_THREADS - 1) / PME_THREADS;  // <= TSCAN_FUSED / PME_THREADS
  const int b = t * per, e = min(n, b + per);
  int v[TSCAN_FUSED / PME_THREADS];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < TSCAN_FUSED / PME_THREADS; k++) {
    v[k] = (b + k < e) ? __hip_atomic_load(&cnt[b + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                       : 0;
    sum += v[k];
  }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < PME_THREADS; off <<= 1) {
    const int x = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  int run = part[t] - sum;
#pragma unroll
  for (int k = 0; k < TSCAN_FUSED / PME_THREADS; k++) {
    if (b + k < e) {
      start[b + k] = run;
      run += v[k];
      __hip_atomic_store(&cnt[b + k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (t == PME_THREADS - 1) {
    start[n] = part[PME_THREADS - 1];
    __hip_atomic_store(&cnt[n], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
__device__ __forceinline__ void tbin_body(const TbinArgs &A, int bid, int nblk) {
  tbin_impl(A.coor, A.npt, A.order, A.P, A.T, A.tcnt, A.atile, A.arank, bid);
  if (!A.fuse_scan) return;  // a launch argument: uniform
  __shared__ int last;
  const int ntile = A.T.nt1 * A.T.nt2 * A.T.nt3;
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(&A.tcnt[ntile], 1) == nblk - 1;
  __syncthreads();
  if (__builtin_amdgcn_readfirstlane(last)) tscan_last(A.tcnt, ntile, A.tstart);
}
DEFINE_KERNEL(k_pme_tbin, TbinArgs, PME_THREADS, tbin_body)

// exclusive scan of the tile counts (one workgroup); clears the counts
// for the next evaluation
__device__ __forceinline__ void tscan_impl(int32_t *__restrict__ cnt, int n,
                                           int32_t *__restrict__ start) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = min(n, b + per);
  int sum = 0;
  for (int i = b; i < e; i++) sum += cnt[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  for (int i = b; i < e; i++) {
    int v = cnt[i];
    start[i] = run;
    run += v;
    cnt[i] = 0;
  }
  if (t == 1023) start[n] = part[1023];
}
__device__ __forceinline__ void tscan_body(const TscanArgs &A, int bid, int nblk) {
  tscan_impl(A.cnt, A.n, A.start);
}
DEFINE_KERNEL(k_pme_tscan, TscanArgs, 1024, tscan_body)

// counting-sort placement: the atoms of every tile, contiguous in tile
// order (sidx)
__device__ __forceinline__ void tplace_body(const TplaceArgs &A, int bid, int nblk) {
  const int n = bid * PME_THREADS + threadIdx.x;
  if (n >= A.npt) return;
  A.sidx[A.tstart[A.atile[n]] + A.arank[n]] = n;
}
DEFINE_KERNEL(k_pme_tplace, TplaceArgs, PME_THREADS, tplace_body)

// one workgroup per tile: up to TS_STAGE of the tile's atoms at a time get
// their B-spline weights theta1(4), theta2(4), q sqrt(kofdie/eps)
// theta3(4) and region base in LDS (one thread per atom), then every atom
// adds its 64 stencil weights (one lane per stencil point) into the
// tile's region in LDS, in 64-bit fixed point (order independent); the
// region is then written to the tile's slot of tbuf.  An atom whose base
// point left the region's accepted range adds its 64 weights to the
// overflow grid with device atomics (the same fixed-point integers).
#define TS_STAGE 256
#define RB_GRID 0x40000000
template <typename R>
__device__ __forceinline__ void tspread_impl(const TspreadArgs &A, int bid) {
  PME_WAVE_PRIORITY();
  __shared__ unsigned long long acc[PME_RMAX * PME_RMAX * PME_RMAX];
  __shared__ R rec[TS_STAGE * 12];
  __shared__ int32_t rb[TS_STAGE];
  const TilePar &T = A.T;
  const PmePar &P = A.P;
  const int tile = bid, t = threadIdx.x;
  const int tt1 = tile % T.nt1, tt2 = (tile / T.nt1) % T.nt2, tt3 = tile / (T.nt1 * T.nt2);
  const int lo1 = tile_lo(tt1, T.n1, T.nt1), lo2 = tile_lo(tt2, T.n2, T.nt2),
            lo3 = tile_lo(tt3, T.n3, T.nt3);
  const int hi1 = tile_lo(tt1 + 1, T.n1, T.nt1), hi2 = tile_lo(tt2 + 1, T.n2, T.nt2),
            hi3 = tile_lo(tt3 + 1, T.n3, T.nt3);
  for (int u = t; u < T.rsz; u += PME_THREADS) acc[u] = 0;
  const int lane = t & 63, wv = t >> 6;
  const int la = lane & 3, lb = (lane >> 2) & 3, lc = lane >> 4;
  const int beg = A.tstart[tile], end = A.tstart[tile + 1];
  for (int b0 = beg; b0 < end; b0 += TS_STAGE) {
    const int n = min(TS_STAGE, end - b0);
    __syncthreads();
    if (t < n) {
      const int at = A.sidx[b0 + t];
      int c1, c2, c3;
      double w1, w2, w3;
      base_point(A.coor + 3 * at, P, c1, c2, c3, w1, w2, w3);
      R t1[4], t2[4], t3[4], dd[4];
      bspline4<R>((R)w1, t1, dd);
      bspline4<R>((R)w2, t2, dd);
      bspline4<R>((R)w3, t3, dd);
      const R qf = (R)(A.q[at] * P.fc);
      R *o = rec + 12 * t;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        o[k] = t1[k];
        o[4 + k] = t2[k];
        o[8 + k] = qf * t3[k];
      }
      const int b1 = local_base(c1, lo1, hi1, T.n1), b2 = local_base(c2, lo2, hi2, T.n2),
                b3 = local_base(c3, lo3, hi3, T.n3);
      // the local base in the region, or (bit 30) the grid base point of
      // an atom outside the region (it moved beyond the margin since the
      // list build; grid dimensions are below 1024)
      rb[t] = (b1 >= 0 && b2 >= 0 && b3 >= 0) ? (b1 | (b2 << 8) | (b3 << 16))
                                               : (RB_GRID | c1 | (c2 << 10) | (c3 << 20));
    }
    __syncthreads();
    for (int a = wv; a < n; a += PME_THREADS / 64) {
      const int bp = rb[a];
      const R *o = rec + 12 * a;
      const R v = o[8 + lc] * o[4 + lb] * o[la];
      const long long f = qfix((double)v);
      if (!(bp & RB_GRID)) {  // uniform over the wavefront
        const int l1 = (bp & 255) + la, l2 = ((bp >> 8) & 255) + lb, l3 = (bp >> 16) + lc;
        if (f) atomicAdd(&acc[(l3 * T.r2 + l2) * T.r1 + l1], (unsigned long long)f);
      } else {
        // the stencil point c-3+l straight into the overflow grid
        int g1 = (bp & 1023) - 3 + la, g2 = ((bp >> 10) & 1023) - 3 + lb,
            g3 = ((bp >> 20) & 1023) - 3 + lc;
        g1 += g1 < 0 ? T.n1 : 0;
        g2 += g2 < 0 ? T.n2 : 0;
        g3 += g3 < 0 ? T.n3 : 0;
        if (f) atomicAdd(&A.qover[((size_t)g3 * T.n2 + g2) * T.n1 + g1], (unsigned long long)f);
      }
    }
  }
  __syncthreads();
  long long *dst = A.tbuf + (size_t)tile * T.rsz;
  for (int u = t; u < T.rsz; u += PME_THREADS) dst[u] = (long long)acc[u];
}
template <typename R>
__device__ __forceinline__ void tspread_body(const TspreadArgs &A, int bid, int nblk) {
  tspread_impl<R>(A, bid);
}
DEFINE_KERNEL(k_pme_tspread_f32, TspreadArgs, PME_THREADS, tspread_body<float>)
DEFINE_KERNEL(k_pme_tspread_f64, TspreadArgs, PME_THREADS, tspread_body<double>)

// per grid point: the entries of every tile region covering it (the
// point's own tile and its neighbours in each dimension, every periodic
// image the region holds) plus the overflow grid's entry, which is
// cleared for the next evaluation; converted to the transform's real type
template <typename R>
__device__ __forceinline__ void tsum_impl(const TsumArgs &A, int bid) {
  PME_WAVE_PRIORITY();
  const TilePar &T = A.T;
  const size_t G = (size_t)T.n1 * T.n2 * T.n3;
  size_t g = (size_t)bid * PME_THREADS + threadIdx.x;
  if (g >= G) return;
  const int gg[3] = {(int)(g % T.n1), (int)((g / T.n1) % T.n2), (int)(g / ((size_t)T.n1 * T.n2))};
  const int nn[3] = {T.n1, T.n2, T.n3}, nt[3] = {T.nt1, T.nt2, T.nt3};
  // per dimension: (tile, region index) of every region cell holding the
  // point (at most 4: see the region sizes)
  int ct[3][4], cl[3][4], nc[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int own = tile_of(gg[d], nn[d], nt[d]);
    const int prev = own == 0 ? nt[d] - 1 : own - 1, next = own + 1 == nt[d] ? 0 : own + 1;
    const int cand[3] = {own, prev, next};
    const int ncand = nt[d] == 1 ? 1 : (nt[d] == 2 ? 2 : 3);
    int m = 0;
    for (int q = 0; q < ncand; q++) {
      const int tk = cand[q];
      const int lo = tile_lo(tk, nn[d], nt[d]), hi = tile_lo(tk + 1, nn[d], nt[d]);
      const int s0 = lo - PME_MARGIN - 3, rlen = hi - lo + 3 + 2 * PME_MARGIN;
      for (int k = -1; k <= 1; k++) {
        const int l = gg[d] + k * nn[d] - s0;
        if (l >= 0 && l < rlen && m < 4) {
          ct[d][m] = tk;
          cl[d][m] = l;
          m++;
        }
      }
    }
    nc[d] = m;
  }
  long long sum = 0;
  for (int c = 0; c < nc[2]; c++)
    for (int b = 0; b < nc[1]; b++)
      for (int a = 0; a < nc[0]; a++) {
        const int tile = (ct[2][c] * T.nt2 + ct[1][b]) * T.nt1 + ct[0][a];
        sum += A.tbuf[(size_t)tile * T.rsz + (cl[2][c] * T.r2 + cl[1][b]) * T.r1 + cl[0][a]];
      }
  const unsigned long long ov = A.qover[g];
  if (ov) {
    sum += (long long)ov;
    A.qover[g] = 0;
  }
  ((R *)A.grid)[g] = (R)((double)sum * A.qinv);
}
template <typename R>
__device__ __forceinline__ void tsum_body(const TsumArgs &A, int bid, int nblk) {
  tsum_impl<R>(A, bid);
}
DEFINE_KERNEL(k_pme_tsum_f32, TsumArgs, PME_THREADS, tsum_body<float>)
DEFINE_KERNEL(k_pme_tsum_f64, TsumArgs, PME_THREADS, tsum_body<double>)

// half-spectrum convolution + energy (weights 1 for the k1 = 0 and
// k1 = n1/2 planes, 2 otherwise)
template <typename R>
__device__ __forceinline__ void conv_impl(R *__restrict__ S, const double *__restrict__ eterm,
                                          int nc1, int n1, size_t H, double *__restrict__ epart,
                                          int bid) {
  PME_WAVE_PRIORITY();
  __shared__ double red[PME_THREADS];
  size_t g = (size_t)bid * PME_THREADS + threadIdx.x;
  double e = 0.0;
  if (g < H) {
    int k1 = (int)(g % nc1);
    double w = (k1 == 0 || (2 * k1 == n1)) ? 1.0 : 2.0;
    double et = eterm[g];
    R re = S[2 * g], im = S[2 * g + 1];
    e = w * et * ((double)re * (double)re + (double)im * (double)im);
    S[2 * g] = (R)(re * (R)et);
    S[2 * g + 1] = (R)(im * (R)et);
  }
  int t = threadIdx.x;
  red[t] = e;
  __syncthreads();
  for (int off = PME_THREADS / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) epart[bid] = 0.5 * red[0];
}
template <typename R>
__device__ __forceinline__ void conv_body(const ConvArgs &A, int bid, int nblk) {
  conv_impl<R>((R *)A.S, A.eterm, A.nc1, A.n1, A.H, A.epart, bid);
}
DEFINE_KERNEL(k_conv_f32, ConvArgs, PME_THREADS, conv_body<float>)
DEFINE_KERNEL(k_conv_f64, ConvArgs, PME_THREADS, conv_body<double>)

// one workgroup per tile: the tile's region of the convolved grid goes to
// LDS, then one thread per atom of the tile (tile order) recomputes its
// B-splines and derivatives from the coordinates and sums its 64 stencil
// points from LDS (or, for an atom outside the region, from the grid);
// each atom's reciprocal gradient is written directly (no atomics)
template <typename R>
__device__ __forceinline__ void load_region(const TilePar &T, int lo1, int lo2, int lo3,
                                            const R *__restrict__ grid, R *reg) {
  for (int u = threadIdx.x; u < T.rsz; u += PME_THREADS) {
    const int l1 = u % T.r1, l2 = (u / T.r1) % T.r2, l3 = u / (T.r1 * T.r2);
    // region index l <-> grid point lo - M - 3 + l (periodic); entries
    // past a smaller tile's own region are never read
    int g1 = lo1 - PME_MARGIN - 3 + l1, g2 = lo2 - PME_MARGIN - 3 + l2,
        g3 = lo3 - PME_MARGIN - 3 + l3;
    g1 = g1 < 0 ? g1 + T.n1 : (g1 >= T.n1 ? g1 - T.n1 : g1);
    g2 = g2 < 0 ? g2 + T.n2 : (g2 >= T.n2 ? g2 - T.n2 : g2);
    g3 = g3 < 0 ? g3 + T.n3 : (g3 >= T.n3 ? g3 - T.n3 : g3);
    reg[u] = grid[((size_t)g3 * T.n2 + g2) * T.n1 + g1];
  }
}

// the 4x4x4 stencil values of an atom: LOCAL, from the LDS region at
// local base (b1, b2, b3); else from the grid around base point
// (b1, b2, b3) = (c1, c2, c3).  The two forms are separate code paths of
// the callers, so the rare grid path costs the common one no registers.
template <typename R, bool LOCAL>
__device__ __forceinline__ R stencil_value(const TilePar &T, const R *__restrict__ src, int b1,
                                           int b2, int b3, int a, int b, int c) {
  if (LOCAL) return src[((b3 + c) * T.r2 + (b2 + b)) * T.r1 + b1 + a];
  int g1 = b1 - 3 + a, g2 = b2 - 3 + b, g3 = b3 - 3 + c;
  g1 += g1 < 0 ? T.n1 : 0;
  g2 += g2 < 0 ? T.n2 : 0;
  g3 += g3 < 0 ? T.n3 : 0;
  return src[((size_t)g3 * T.n2 + g2) * T.n1 + g1];
}

// the gather's three sums over an atom's stencil (derivative weights in
// one dimension each)
template <typename R, bool LOCAL>
__device__ __forceinline__ void gather_sums(const TilePar &T, const R *__restrict__ src, int b1,
                                            int b2, int b3, const R *t1, const R *t2,
                                            const R *t3, const R *d1, const R *d2, const R *d3,
                                            double &f1, double &f2, double &f3) {
#pragma unroll
  for (int c = 0; c < 4; c++) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      R g1 = 0, g2 = 0;
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const R Q = stencil_value<R, LOCAL>(T, src, b1, b2, b3, a, b, c);
        g1 += Q * d1[a];
        g2 += Q * t1[a];
      }
      f1 += (double)(g1 * t2[b] * t3[c]);
      f2 += (double)(g2 * d2[b] * t3[c]);
      f3 += (double)(g2 * t2[b] * d3[c]);
    }
  }
}

// the potential's sum over an atom's stencil
template <typename R, bool LOCAL>
__device__ __forceinline__ double potential_sum(const TilePar &T, const R *__restrict__ src,
                                                int b1, int b2, int b3, const R *t1,
                                                const R *t2, const R *t3) {
  double s = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      R g = 0;
#pragma unroll
      for (int a = 0; a < 4; a++) g += stencil_value<R, LOCAL>(T, src, b1, b2, b3, a, b, c) * t1[a];
      s += (double)(g * t2[b] * t3[c]);
    }
  }
  return s;
}

template <typename R>
__device__ __forceinline__ void gather_impl(const double *__restrict__ coor,
                                            const double *__restrict__ q,
                                            const int32_t *__restrict__ tstart,
                                            const int32_t *__restrict__ sidx, const PmePar &P,
                                            const TilePar &T, const R *__restrict__ grid,
                                            double *__restrict__ fpme, int bid) {
  PME_WAVE_PRIORITY();
  __shared__ R reg[PME_RMAX * PME_RMAX * PME_RMAX];
  const int tile = bid, t = threadIdx.x;
  const int tt1 = tile % T.nt1, tt2 = (tile / T.nt1) % T.nt2, tt3 = tile / (T.nt1 * T.nt2);
  const int lo1 = tile_lo(tt1, T.n1, T.nt1), lo2 = tile_lo(tt2, T.n2, T.nt2),
            lo3 = tile_lo(tt3, T.n3, T.nt3);
  const int hi1 = tile_lo(tt1 + 1, T.n1, T.nt1), hi2 = tile_lo(tt2 + 1, T.n2, T.nt2),
            hi3 = tile_lo(tt3 + 1, T.n3, T.nt3);
  load_region<R>(T, lo1, lo2, lo3, grid, reg);
  __syncthreads();
  const int beg = tstart[tile], end = tstart[tile + 1];
  for (int p = beg + t; p < end; p += PME_THREADS) {
    const int n = sidx[p];
    const double qn = q[n] * P.fc;
    if (qn == 0.0) {
      fpme[3 * n] = fpme[3 * n + 1] = fpme[3 * n + 2] = 0.0;
      continue;
    }
    int c1, c2, c3;
    double w1, w2, w3;
    base_point(coor + 3 * n, P, c1, c2, c3, w1, w2, w3);
    R t1[4], t2[4], t3[4], d1[4], d2[4], d3[4];
    bspline4<R>((R)w1, t1, d1);
    bspline4<R>((R)w2, t2, d2);
    bspline4<R>((R)w3, t3, d3);
    const int b1 = local_base(c1, lo1, hi1, T.n1), b2 = local_base(c2, lo2, hi2, T.n2),
              b3 = local_base(c3, lo3, hi3, T.n3);
    double f1 = 0, f2 = 0, f3 = 0;
    if (b1 >= 0 && b2 >= 0 && b3 >= 0)
      gather_sums<R, true>(T, reg, b1, b2, b3, t1, t2, t3, d1, d2, d3, f1, f2, f3);
    else  // outside the region (moved beyond the margin): the grid itself
      gather_sums<R, false>(T, grid, c1, c2, c3, t1, t2, t3, d1, d2, d3, f1, f2, f3);
    // dE/dfr_k = qn * sum(...) * n_k ; dpot -= recip * f with f = -dE/dfr
    fpme[3 * n] = qn * f1 * (double)P.n1 * P.ra;
    fpme[3 * n + 1] = qn * f2 * (double)P.n2 * P.rb;
    fpme[3 * n + 2] = qn * f3 * (double)P.n3 * P.rc;
  }
}
template <typename R>
__device__ __forceinline__ void gather_body(const GatherArgs &A, int bid, int nblk) {
  gather_impl<R>(A.coor, A.q, A.tstart, A.sidx, A.P, A.T, (const R *)A.grid, A.fpme, bid);
}
DEFINE_KERNEL(k_gather_f32, GatherArgs, PME_THREADS, gather_body<float>)
DEFINE_KERNEL(k_gather_f64, GatherArgs, PME_THREADS, gather_body<double>)

// ---------------------------------------------------------------------------
static void bsp_moduli(std::vector<double> &mod, int nfft) {
  // fill_bspline(0): order-4 values at integer offsets (1/6, 2/3, 1/6, 0)
  double arr[4], darr[4];
  {
    double w = 0.0;
    double c1 = 1 - w, c2 = w;
    double o3 = 0.5 * w * c2, o2 = 0.5 * ((w + 1) * c1 + (2 - w) * c2), o1 = 0.5 * (1 - w) * c1;
    darr[0] = -o1;
    (void)darr;
    arr[3] = (1.0 / 3.0) * w * o3;
    arr[2] = (1.0 / 3.0) * ((w + 1) * o2 + (3 - w) * o3);
    arr[1] = (1.0 / 3.0) * ((w + 2) * o1 + (2 - w) * o2);
    arr[0] = (1.0 / 3.0) * (1 - w) * o1;
  }
  std::vector<double> b(nfft + 8, 0.0);
  for (int k = 2; k <= 5; k++) b[k - 1] = arr[k - 2];
  const double twopi = 2.0 * 3.14159265358979323846;
  mod.assign(nfft, 0.0);
  for (int k = 1; k <= nfft; k++) {
    double s1 = 0, s2 = 0;
    for (int j = 1; j <= nfft; j++) {
      double arg = twopi * (k - 1) * (j - 1) / nfft;
      s1 += b[j - 1] * cos(arg);
      s2 += b[j - 1] * sin(arg);
    }
    mod[k - 1] = s1 * s1 + s2 * s2;
  }
  for (int k = 1; k <= nfft; k++)
    if (mod[k - 1] < 1e-7) mod[k - 1] = 0.5 * (mod[(k - 2 + nfft) % nfft] + mod[k % nfft]);
}

static const char *rf_err(rocfft_status s) {
  return s == rocfft_status_success ? nullptr : "rocFFT failure";
}

#define RF_CHECK(x)                                                   \
  do {                                                                \
    rocfft_status s_ = (x);                                           \
    if (s_ != rocfft_status_success) {                                \
      set_error(std::string("rocFFT error ") + std::to_string((int)s_) + \
                " at " + __FILE__ + ":" + std::to_string(__LINE__));  \
      return MOIL_GPU_EHIP;                                           \
    }                                                                 \
  } while (0)

// Tiles of at most `edge` points a side (4..8): per dimension
// nt = ceil(n / edge) tiles, fewer where a tile would hold fewer than 4
// points (the halo sum reads the regions of a point's own tile and its two
// neighbours only, which covers every region holding the point while
// tiles are at least 4 wide); the tile buffers are sized to match.  The
// atoms' tiles are assigned at the next list build.
int32_t pme_set_tile(moil_ctx *c, int edge) {
  if (edge < 4 || edge > 8) {
    set_error("PME spreading tile edge must lie in 4..8");
    return MOIL_GPU_EINVAL;
  }
  const int n[3] = {c->nf1, c->nf2, c->nf3};
  int nt[3], r[3];
  for (int d = 0; d < 3; d++) {
    nt[d] = (n[d] + edge - 1) / edge;
    while (nt[d] > 1 && n[d] / nt[d] < 4) nt[d]--;
    r[d] = (n[d] + nt[d] - 1) / nt[d] + 3 + 2 * PME_MARGIN;
    if (r[d] > PME_RMAX) {
      set_error("PME grid dimension " + std::to_string(n[d]) + " has no tiling with edge " +
                std::to_string(edge));
      return MOIL_GPU_EINVAL;
    }
  }
  if (c->d_tbuf && nt[0] == c->pt_nt[0] && nt[1] == c->pt_nt[1] && nt[2] == c->pt_nt[2]) {
    c->pt_edge = edge;
    return MOIL_GPU_OK;
  }
  hipFree(c->d_tcnt); c->d_tcnt = nullptr;
  hipFree(c->d_tstart); c->d_tstart = nullptr;
  hipFree(c->d_tbuf); c->d_tbuf = nullptr;
  for (int d = 0; d < 3; d++) {
    c->pt_nt[d] = nt[d];
    c->pt_r[d] = r[d];
  }
  c->pt_ntile = nt[0] * nt[1] * nt[2];
  c->pt_edge = edge;
  const size_t rsz = (size_t)r[0] * r[1] * r[2];
  HIP_CHECK(hipMalloc(&c->d_tcnt, sizeof(int32_t) * (c->pt_ntile + 1)));
  HIP_CHECK(hipMemset(c->d_tcnt, 0, sizeof(int32_t) * (c->pt_ntile + 1)));
  HIP_CHECK(hipMalloc(&c->d_tstart, sizeof(int32_t) * (c->pt_ntile + 1)));
  HIP_CHECK(hipMalloc(&c->d_tbuf, sizeof(long long) * rsz * c->pt_ntile));
  return MOIL_GPU_OK;
}

int32_t pme_init(moil_ctx *c) {
  const moil_params &P = c->par;
  c->nf1 = P.nfft1; c->nf2 = P.nfft2; c->nf3 = P.nfft3;
  if (c->nf1 < 4 || c->nf2 < 4 || c->nf3 < 4 || c->nf1 > 1023 || c->nf2 > 1023 || c->nf3 > 1023) {
    set_error("PME grid dimensions must lie in 4..1023");
    return MOIL_GPU_EINVAL;
  }
  c->nc1 = c->nf1 / 2 + 1;
  bsp_moduli(c->bsp[0], c->nf1);
  bsp_moduli(c->bsp[1], c->nf2);
  bsp_moduli(c->bsp[2], c->nf3);
  const size_t G = (size_t)c->nf1 * c->nf2 * c->nf3;
  const size_t H = (size_t)c->nc1 * c->nf2 * c->nf3;
  // influence function on the half spectrum, pot/ewald_receip.f:770-798
  std::vector<double> et(H);
  {
    const double pi = 3.14159265358979323846;
    const double fac = pi * pi / (c->beta * c->beta);
    const double pivol = pi * (P.a * P.b * P.c);
    int nf[3] = {c->nf1 / 2, c->nf2 / 2, c->nf3 / 2};
    if (2 * nf[0] < c->nf1) nf[0]++;
    if (2 * nf[1] < c->nf2) nf[1]++;
    if (2 * nf[2] < c->nf3) nf[2]++;
    for (int k3 = 0; k3 < c->nf3; k3++)
      for (int k2 = 0; k2 < c->nf2; k2++)
        for (int k1 = 0; k1 < c->nc1; k1++) {
          size_t g = ((size_t)k3 * c->nf2 + k2) * c->nc1 + k1;
          if (k1 == 0 && k2 == 0 && k3 == 0) {
            et[g] = 0.0;
            continue;
          }
          int m1 = (k1 + 1 > nf[0]) ? k1 - c->nf1 : k1;
          int m2 = (k2 + 1 > nf[1]) ? k2 - c->nf2 : k2;
          int m3 = (k3 + 1 > nf[2]) ? k3 - c->nf3 : k3;
          double mh1 = m1 / P.a, mh2 = m2 / P.b, mh3 = m3 / P.c;
          double msq = mh1 * mh1 + mh2 * mh2 + mh3 * mh3;
          double denom = pivol * c->bsp[0][k1] * c->bsp[1][k2] * c->bsp[2][k3] * msq;
          et[g] = exp(-fac * msq) / denom;
        }
  }
  HIP_CHECK(hipMalloc(&c->d_eterm, sizeof(double) * H));
  HIP_CHECK(hipMemcpy(c->d_eterm, et.data(), sizeof(double) * H, hipMemcpyHostToDevice));
  const bool f32 = c->par.precision == MOIL_PREC_MIXED;
  const size_t rs = f32 ? sizeof(float) : sizeof(double);
  const int na = std::max(c->npt, 1);
  // grid tiles for the spreading (8 points a side; the autotuner may
  // choose smaller ones)
  {
    const int32_t rc = pme_set_tile(c, 8);
    if (rc != MOIL_GPU_OK) return rc;
  }
  HIP_CHECK(hipMalloc(&c->d_atile, sizeof(int32_t) * na));
  HIP_CHECK(hipMalloc(&c->d_arank, sizeof(int32_t) * na));
  HIP_CHECK(hipMalloc(&c->d_qover, sizeof(unsigned long long) * G));
  HIP_CHECK(hipMemset(c->d_qover, 0, sizeof(unsigned long long) * G));
  HIP_CHECK(hipMalloc(&c->d_sidx, sizeof(int32_t) * na));
  {
    // the first binning visits the atoms in index order
    std::vector<int32_t> iota(na);
    for (int k = 0; k < na; k++) iota[k] = k;
    HIP_CHECK(hipMemcpy(c->d_sidx, iota.data(), sizeof(int32_t) * na, hipMemcpyHostToDevice));
  }
  HIP_CHECK(hipMalloc(&c->d_qreal, rs * G));
  HIP_CHECK(hipMalloc(&c->d_qcplx, 2 * rs * H));
  HIP_CHECK(hipMalloc(&c->d_fpme, sizeof(double) * 3 * std::max(c->npt, 1)));
  RF_CHECK(rocfft_setup());
  size_t len[3] = {(size_t)c->nf1, (size_t)c->nf2, (size_t)c->nf3};
  rocfft_precision prec = f32 ? rocfft_precision_single : rocfft_precision_double;
  RF_CHECK(rocfft_plan_create(&c->plan_fwd, rocfft_placement_notinplace,
                              rocfft_transform_type_real_forward, prec, 3, len, 1, nullptr));
  RF_CHECK(rocfft_plan_create(&c->plan_bwd, rocfft_placement_notinplace,
                              rocfft_transform_type_real_inverse, prec, 3, len, 1, nullptr));
  size_t w1 = 0, w2 = 0;
  RF_CHECK(rocfft_plan_get_work_buffer_size(c->plan_fwd, &w1));
  RF_CHECK(rocfft_plan_get_work_buffer_size(c->plan_bwd, &w2));
  c->fftwork_bytes = std::max(w1, w2);
  if (c->fftwork_bytes) HIP_CHECK(hipMalloc(&c->d_fftwork, c->fftwork_bytes));
  RF_CHECK(rocfft_execution_info_create(&c->exec_fwd));
  RF_CHECK(rocfft_execution_info_create(&c->exec_bwd));
  if (c->fftwork_bytes) {
    RF_CHECK(rocfft_execution_info_set_work_buffer(c->exec_fwd, c->d_fftwork, c->fftwork_bytes));
    RF_CHECK(rocfft_execution_info_set_work_buffer(c->exec_bwd, c->d_fftwork, c->fftwork_bytes));
  }
  (void)rf_err;
  return MOIL_GPU_OK;
}

void pme_free(moil_ctx *c) {
  if (c->plan_fwd) rocfft_plan_destroy(c->plan_fwd);
  if (c->plan_bwd) rocfft_plan_destroy(c->plan_bwd);
  if (c->exec_fwd) rocfft_execution_info_destroy(c->exec_fwd);
  if (c->exec_bwd) rocfft_execution_info_destroy(c->exec_bwd);
  c->plan_fwd = c->plan_bwd = nullptr;
  c->exec_fwd = c->exec_bwd = nullptr;
  hipFree(c->d_fftwork); c->d_fftwork = nullptr;
  hipFree(c->d_eterm); c->d_eterm = nullptr;
  hipFree(c->d_tcnt); c->d_tcnt = nullptr;
  hipFree(c->d_tstart); c->d_tstart = nullptr;
  hipFree(c->d_tbuf); c->d_tbuf = nullptr;
  hipFree(c->d_atile); c->d_atile = nullptr;
  hipFree(c->d_arank); c->d_arank = nullptr;
  hipFree(c->d_qover); c->d_qover = nullptr;
  hipFree(c->d_sidx); c->d_sidx = nullptr;
  hipFree(c->d_qreal); c->d_qreal = nullptr;
  hipFree(c->d_qcplx); c->d_qcplx = nullptr;
  hipFree(c->d_fpme); c->d_fpme = nullptr;
}

bool pme_charges_fit(const double *q, size_t n, double eps, std::string *why) {
  const double fc = sqrt(KOFDIE / eps);
  for (size_t i = 0; i < n; i++)
    if (!(fabs(q[i]) * fc < QFIX_MAXW)) {
      if (why)
        *why = "charge " + std::to_string(q[i]) + " of atom " + std::to_string(i + 1) +
               " beyond the reciprocal grid's fixed-point range (|q| sqrt(332.0716/eps) < 2048)";
      return false;
    }
  return true;
}

int pme_conv_blocks(const moil_ctx *c) {
  const size_t H = (size_t)c->nc1 * c->nf2 * c->nf3;
  return (int)((H + PME_THREADS - 1) / PME_THREADS);
}

// the reciprocal-space launches of one evaluation (grid buffers of the
// context; a batch substitutes its own transform buffers)
void pme_plan(const moil_ctx *c, const double *d_coor, EvalPlan *p) {
  const size_t G = (size_t)c->nf1 * c->nf2 * c->nf3;
  const size_t H = (size_t)c->nc1 * c->nf2 * c->nf3;
  PmePar P;
  P.n1 = c->nf1; P.n2 = c->nf2; P.n3 = c->nf3; P.nc1 = c->nc1;
  P.ra = 1.0 / c->par.a; P.rb = 1.0 / c->par.b; P.rc = 1.0 / c->par.c;
  P.fc = (c->par.terms & MOIL_TERM_EL) ? sqrt(KOFDIE / c->par.eps) : 0.0;
  P.qscale = ldexp(1.0, QFIX_SHIFT);
  P.qinv = ldexp(1.0, -QFIX_SHIFT);
  TilePar T;
  T.n1 = c->nf1; T.n2 = c->nf2; T.n3 = c->nf3;
  T.nt1 = c->pt_nt[0]; T.nt2 = c->pt_nt[1]; T.nt3 = c->pt_nt[2];
  T.r1 = c->pt_r[0]; T.r2 = c->pt_r[1]; T.r3 = c->pt_r[2];
  T.rsz = T.r1 * T.r2 * T.r3;
  p->pme = true;
  p->n_tile = c->pt_ntile;
  p->tspread = {d_coor, c->d_ptchg, c->d_tstart, c->d_sidx, P, T, c->d_tbuf, c->d_qover};
  p->n_tsum = (int)((G + PME_THREADS - 1) / PME_THREADS);
  p->tsum = {c->d_tbuf, c->d_qover, T, P.qinv, c->d_qreal};
  p->n_conv = (int)((H + PME_THREADS - 1) / PME_THREADS);
  p->conv = {c->d_qcplx, c->d_eterm, c->nc1, c->nf1, H,
             c->d_epart + (size_t)EN_RECIP * c->max_parts};
  p->gather = {d_coor, c->d_ptchg, c->d_tstart, c->d_sidx, P, T, c->d_qreal, c->d_fpme};
}

// spreading (tile regions, grid sum)
void pme_launch_spread(const EvalPlan &p, hipStream_t st) {
  if (p.f64) {
    launch_k_pme_tspread_f64(p.tspread, p.n_tile, st);
    launch_k_pme_tsum_f64(p.tsum, p.n_tsum, st);
  } else {
    launch_k_pme_tspread_f32(p.tspread, p.n_tile, st);
    launch_k_pme_tsum_f32(p.tsum, p.n_tsum, st);
  }
}

// The atoms' spreading tiles from coordinates d_coor (counting sort by the
// tile of the stencil base point), kept until the next list build: the
// atoms are binned in the previous sort's order, so the lanes of a
// wavefront mostly share a tile; with at most TSCAN_FUSED tiles the last
// binning workgroup also scans the counts.
int32_t pme_bin(moil_ctx *c, const double *d_coor, hipStream_t st) {
  PmePar P;
  P.n1 = c->nf1; P.n2 = c->nf2; P.n3 = c->nf3; P.nc1 = c->nc1;
  P.ra = 1.0 / c->par.a; P.rb = 1.0 / c->par.b; P.rc = 1.0 / c->par.c;
  P.fc = 0.0;
  P.qscale = P.qinv = 0.0;
  TilePar T;
  T.n1 = c->nf1; T.n2 = c->nf2; T.n3 = c->nf3;
  T.nt1 = c->pt_nt[0]; T.nt2 = c->pt_nt[1]; T.nt3 = c->pt_nt[2];
  T.r1 = c->pt_r[0]; T.r2 = c->pt_r[1]; T.r3 = c->pt_r[2];
  T.rsz = T.r1 * T.r2 * T.r3;
  const int nblk = (c->npt + PME_THREADS - 1) / PME_THREADS;
  const int fuse = c->pt_ntile <= TSCAN_FUSED ? 1 : 0;
  TbinArgs B = {d_coor, c->npt, c->d_sidx, P, T, c->d_tcnt, c->d_atile, c->d_arank,
                c->d_tstart, fuse};
  launch_k_pme_tbin(B, nblk, st);
  if (!fuse) launch_k_pme_tscan(TscanArgs{c->d_tcnt, c->pt_ntile, c->d_tstart}, 1, st);
  launch_k_pme_tplace(TplaceArgs{c->npt, c->d_atile, c->d_arank, c->d_tstart, c->d_sidx}, nblk,
                      st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

int32_t pme_run(moil_ctx *c, const EvalPlan &p, hipStream_t st) {
  pme_launch_spread(p, st);
  HIP_CHECK(hipGetLastError());
  void *in[1] = {c->d_qreal}, *out[1] = {c->d_qcplx};
  RF_CHECK(rocfft_execution_info_set_stream(c->exec_fwd, st));
  RF_CHECK(rocfft_execute(c->plan_fwd, in, out, c->exec_fwd));
  // the structure factor before the convolution scales it in place
  if (p.vir) launch_pme_virial(c, p, st);
  if (p.f64) launch_k_conv_f64(p.conv, p.n_conv, st);
  else launch_k_conv_f32(p.conv, p.n_conv, st);
  void *in2[1] = {c->d_qcplx}, *out2[1] = {c->d_qreal};
  RF_CHECK(rocfft_execution_info_set_stream(c->exec_bwd, st));
  RF_CHECK(rocfft_execute(c->plan_bwd, in2, out2, c->exec_bwd));
  if (p.f64) launch_k_gather_f64(p.gather, p.n_tile, st);
  else launch_k_gather_f32(p.gather, p.n_tile, st);
  if (p.vir) launch_rel_virial(c, p, st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

// reciprocal potential for the charge gradient: one workgroup per tile
// (as the gather), dE/dq_n = fc sum_p theta1 theta2 theta3 psi(p) with psi
// the convolved grid after the inverse transform (the same grid whose
// derivative-weighted sum the gather turns into forces); added to dq
template <typename R>
__global__ __launch_bounds__(PME_THREADS) void k_recip_pot(
    const double *__restrict__ coor, const int32_t *__restrict__ tstart,
    const int32_t *__restrict__ sidx, PmePar P, TilePar T, const R *__restrict__ grid,
    double *__restrict__ dq) {
  __shared__ R reg[PME_RMAX * PME_RMAX * PME_RMAX];
  const int tile = blockIdx.x, t = threadIdx.x;
  const int tt1 = tile % T.nt1, tt2 = (tile / T.nt1) % T.nt2, tt3 = tile / (T.nt1 * T.nt2);
  const int lo1 = tile_lo(tt1, T.n1, T.nt1), lo2 = tile_lo(tt2, T.n2, T.nt2),
            lo3 = tile_lo(tt3, T.n3, T.nt3);
  const int hi1 = tile_lo(tt1 + 1, T.n1, T.nt1), hi2 = tile_lo(tt2 + 1, T.n2, T.nt2),
            hi3 = tile_lo(tt3 + 1, T.n3, T.nt3);
  load_region<R>(T, lo1, lo2, lo3, grid, reg);
  __syncthreads();
  const int beg = tstart[tile], end = tstart[tile + 1];
  for (int p = beg + t; p < end; p += PME_THREADS) {
    const int n = sidx[p];
    int c1, c2, c3;
    double w1, w2, w3;
    base_point(coor + 3 * n, P, c1, c2, c3, w1, w2, w3);
    R t1[4], t2[4], t3[4], dd[4];
    bspline4<R>((R)w1, t1, dd);
    bspline4<R>((R)w2, t2, dd);
    bspline4<R>((R)w3, t3, dd);
    const int b1 = local_base(c1, lo1, hi1, T.n1), b2 = local_base(c2, lo2, hi2, T.n2),
              b3 = local_base(c3, lo3, hi3, T.n3);
    const double s = (b1 >= 0 && b2 >= 0 && b3 >= 0)
                         ? potential_sum<R, true>(T, reg, b1, b2, b3, t1, t2, t3)
                         : potential_sum<R, false>(T, grid, c1, c2, c3, t1, t2, t3);
    dq[n] += P.fc * s;
  }
}

int32_t pme_potential(moil_ctx *c, const EvalPlan &p, double *d_dq, hipStream_t st) {
  pme_launch_spread(p, st);
  void *in[1] = {c->d_qreal}, *out[1] = {c->d_qcplx};
  RF_CHECK(rocfft_execution_info_set_stream(c->exec_fwd, st));
  RF_CHECK(rocfft_execute(c->plan_fwd, in, out, c->exec_fwd));
  if (p.f64) launch_k_conv_f64(p.conv, p.n_conv, st);
  else launch_k_conv_f32(p.conv, p.n_conv, st);
  void *in2[1] = {c->d_qcplx}, *out2[1] = {c->d_qreal};
  RF_CHECK(rocfft_execution_info_set_stream(c->exec_bwd, st));
  RF_CHECK(rocfft_execute(c->plan_bwd, in2, out2, c->exec_bwd));
  const GatherArgs &g = p.gather;
  if (p.f64)
    k_recip_pot<double><<<p.n_tile, PME_THREADS, 0, st>>>(g.coor, g.tstart, g.sidx, g.P, g.T,
                                                           (const double *)g.grid, d_dq);
  else
    k_recip_pot<float><<<p.n_tile, PME_THREADS, 0, st>>>(g.coor, g.tstart, g.sidx, g.P, g.T,
                                                          (const float *)g.grid, d_dq);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}
