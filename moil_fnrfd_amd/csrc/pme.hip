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
// GPU design: spreading is done without atomics on the grid.  Atoms are
// binned by their stencil base point int(fr) with a counting sort (one
// integer atomic per atom on the bin counts, a device-wide scan, a
// scatter); then one thread per grid point gathers the 4x4x4 bins whose
// stencils cover it and sums the B-spline weighted charges in fixed point
// (order independent, so bit-reproducible), writing the float (mixed) or
// double (FP64) grid for rocFFT's real-to-complex transform -- the only
// library call.  The convolution runs on the half spectrum
// with the influence function tabulated once in FP64 at init; the
// reciprocal energy is summed in FP64 with self-conjugate planes weighted
// once and all others twice.  The gather recomputes the B-splines from
// the coordinates instead of storing them, and writes each atom's
// reciprocal gradient directly (no atomics).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctx.h"

#define PME_THREADS 256
#define QFIX_SHIFT 40

struct PmePar {
  int n1, n2, n3, nc1;
  double ra, rb, rc;  // 1/a, 1/b, 1/c
  double fc;          // sqrt(kofdie/eps)
  double qscale, qinv;
};

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

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// stencil base bin of every atom, counting-sort histogram
template <typename R>
__global__ __launch_bounds__(PME_THREADS) void k_pme_bin(const double *__restrict__ coor,
                                                         const double *__restrict__ q, int npt,
                                                         PmePar P, int32_t *__restrict__ cnt,
                                                         int32_t *__restrict__ abin,
                                                         int32_t *__restrict__ arank,
                                                         R *__restrict__ aw) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= npt) return;
  int i0, j0, k0;
  double w1, w2, w3;
  frac_coord(coor[3 * n], P.ra, P.n1, &i0, &w1);
  frac_coord(coor[3 * n + 1], P.rb, P.n2, &j0, &w2);
  frac_coord(coor[3 * n + 2], P.rc, P.n3, &k0, &w3);
  // base point int(fr) = i0 + 3, in [0, n] (n only when fr == n exactly)
  int c1 = i0 + 3, c2 = j0 + 3, c3 = k0 + 3;
  if (c1 >= P.n1) c1 -= P.n1;
  if (c2 >= P.n2) c2 -= P.n2;
  if (c3 >= P.n3) c3 -= P.n3;
  int bin = (c3 * P.n2 + c2) * P.n1 + c1;
  abin[n] = bin;
  arank[n] = atomicAdd(&cnt[bin], 1);
  aw[4 * n] = (R)w1;
  aw[4 * n + 1] = (R)w2;
  aw[4 * n + 2] = (R)w3;
  aw[4 * n + 3] = (R)(q[n] * P.fc);
}

template <typename R>
__global__ __launch_bounds__(PME_THREADS) void k_pme_scatter(int npt, const int32_t *__restrict__ abin,
                                                             const int32_t *__restrict__ arank,
                                                             const int32_t *__restrict__ start,
                                                             const R *__restrict__ aw,
                                                             R *__restrict__ sw,
                                                             int32_t *__restrict__ sx, int n1) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= npt) return;
  int b = abin[n];
  int p = start[b] + arank[n];
  sw[4 * p] = aw[4 * n];
  sw[4 * p + 1] = aw[4 * n + 1];
  sw[4 * p + 2] = aw[4 * n + 2];
  sw[4 * p + 3] = aw[4 * n + 3];
  sx[p] = b % n1;
}

__device__ __forceinline__ void fix_split_f(float f, int &h, int &l) {
  const float x = f * 256.0f;
  const float hr = rintf(x);
  const float r = x - hr;
  h = (int)hr;
  l = (int)rintf(r * 16777216.0f);
}

// one thread per grid point: Q(g) = sum over atoms with base point b,
// b - 3 <= g <= b (periodic), of q fc theta1 theta2 theta3
template <typename R>
__global__ __launch_bounds__(PME_THREADS) void k_pme_spread(const int32_t *__restrict__ start,
                                                            const R *__restrict__ sw,
                                                            const int32_t *__restrict__ sx,
                                                            PmePar P, R *__restrict__ grid) {
  const size_t G = (size_t)P.n1 * P.n2 * P.n3;
  size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const int g1 = (int)(g % P.n1);
  const int g2 = (int)((g / P.n1) % P.n2);
  const int g3 = (int)(g / ((size_t)P.n1 * P.n2));
  int acc_h = 0, acc_l = 0;  // FP32 path: fixed point 2^-32 split h*2^24+l
  long long acc = 0;         // FP64 path: fixed point 2^-40
#pragma unroll
  for (int d3 = 0; d3 < 4; d3++) {
    int c3 = g3 + d3;
    if (c3 >= P.n3) c3 -= P.n3;
#pragma unroll
    for (int d2 = 0; d2 < 4; d2++) {
      int c2 = g2 + d2;
      if (c2 >= P.n2) c2 -= P.n2;
      const int row = (c3 * P.n2 + c2) * P.n1;
      // bins g1..g1+3 of this row: one or two contiguous runs
      int e1 = g1 + 3;
      int lo[2], hi[2], nr = 1;
      if (e1 < P.n1) {
        lo[0] = start[row + g1];
        hi[0] = start[row + e1 + 1];
      } else {
        lo[0] = start[row + g1];
        hi[0] = start[row + P.n1];
        lo[1] = start[row];
        hi[1] = start[row + (e1 - P.n1) + 1];
        nr = 2;
      }
      for (int rr = 0; rr < nr; rr++) {
        for (int a = lo[rr]; a < hi[rr]; a++) {
          int x1 = sx[a] - g1;
          if (x1 < 0) x1 += P.n1;  // 0..3
          R t1[4], t2[4], t3[4], dd[4];
          bspline4<R>(sw[4 * a], t1, dd);
          bspline4<R>(sw[4 * a + 1], t2, dd);
          bspline4<R>(sw[4 * a + 2], t3, dd);
          const R v = sw[4 * a + 3] * t3[3 - d3] * t2[3 - d2] * pick4<R>(t1, 3 - x1);
          if (sizeof(R) == 4) {
            int h, l;
            fix_split_f((float)v, h, l);
            acc_h += h;
            acc_l += l;
          } else {
            acc += __double2ll_rn((double)v * P.qscale);
          }
        }
      }
    }
  }
  if (sizeof(R) == 4)
    grid[g] = (R)((double)(((long long)acc_h << 24) + (long long)acc_l) * 0x1p-32);
  else
    grid[g] = (R)((double)acc * P.qinv);
}

// device-wide exclusive scan in three passes (per-block sums, scan of the
// block sums, per-block rescan); the last pass also clears the counts for
// the next evaluation
#define SCAN_PER 2048
__global__ __launch_bounds__(256) void k_scan_sums(const int32_t *__restrict__ in, size_t n,
                                                   int32_t *__restrict__ bsum) {
  __shared__ int red[256];
  size_t b0 = (size_t)blockIdx.x * SCAN_PER;
  int s = 0;
  for (int k = threadIdx.x; k < SCAN_PER; k += 256) {
    size_t i = b0 + k;
    if (i < n) s += in[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void k_scan_top(int32_t *__restrict__ bsum, int nb) {
  __shared__ int part[1024];
  int t = threadIdx.x;
  int per = (nb + 1023) / 1024;
  int b = t * per, e = min(nb, b + per);
  int s = 0;
  for (int i = b; i < e; i++) s += bsum[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t == 0) ? 0 : part[t - 1];
  for (int i = b; i < e; i++) {
    int v = bsum[i];
    bsum[i] = run;
    run += v;
  }
  if (t == 1023) bsum[nb] = part[1023];
}

__global__ __launch_bounds__(256) void k_scan_apply(int32_t *__restrict__ cnt, size_t n,
                                                    const int32_t *__restrict__ bsum,
                                                    int32_t *__restrict__ out) {
  __shared__ int part[256];
  size_t b0 = (size_t)blockIdx.x * SCAN_PER;
  const int per = SCAN_PER / 256;  // 8 consecutive elements per thread
  int v[per];
  int s = 0;
  for (int k = 0; k < per; k++) {
    size_t i = b0 + (size_t)threadIdx.x * per + k;
    v[k] = (i < n) ? cnt[i] : 0;
    s += v[k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int a = (threadIdx.x >= off) ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += a;
    __syncthreads();
  }
  int run = bsum[blockIdx.x] + part[threadIdx.x] - s;
  for (int k = 0; k < per; k++) {
    size_t i = b0 + (size_t)threadIdx.x * per + k;
    if (i < n) {
      out[i] = run;
      cnt[i] = 0;
    }
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) out[n] = bsum[gridDim.x];
}

// half-spectrum convolution + energy (weights 1 for the k1 = 0 and
// k1 = n1/2 planes, 2 otherwise)
template <typename R>
__global__ __launch_bounds__(PME_THREADS) void k_conv(R *__restrict__ S, const double *__restrict__ eterm,
                                                      int nc1, int n1, size_t H,
                                                      double *__restrict__ epart) {
  __shared__ double red[PME_THREADS];
  size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
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
  if (t == 0) epart[blockIdx.x] = 0.5 * red[0];
}

template <typename R>
__global__ __launch_bounds__(PME_THREADS) void k_gather(const double *__restrict__ coor,
                                                        const double *__restrict__ q, int npt,
                                                        PmePar P, const R *__restrict__ grid,
                                                        double *__restrict__ fpme) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= npt) return;
  double qn = q[n] * P.fc;
  if (qn == 0.0) {
    fpme[3 * n] = fpme[3 * n + 1] = fpme[3 * n + 2] = 0.0;
    return;
  }
  int i0, j0, k0;
  double w1, w2, w3;
  frac_coord(coor[3 * n], P.ra, P.n1, &i0, &w1);
  frac_coord(coor[3 * n + 1], P.rb, P.n2, &j0, &w2);
  frac_coord(coor[3 * n + 2], P.rc, P.n3, &k0, &w3);
  R t1[4], t2[4], t3[4], d1[4], d2[4], d3[4];
  bspline4<R>((R)w1, t1, d1);
  bspline4<R>((R)w2, t2, d2);
  bspline4<R>((R)w3, t3, d3);
  double f1 = 0, f2 = 0, f3 = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int k = wrapi(k0 + c, P.n3);
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int j = wrapi(j0 + b, P.n2);
      const R *row = grid + ((size_t)k * P.n2 + j) * P.n1;
      R g1 = 0, g2 = 0, g3 = 0;
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const int i = wrapi(i0 + a, P.n1);
        const R Q = row[i];
        g1 += Q * d1[a];
        g2 += Q * t1[a];
      }
      g3 = g2;
      f1 += (double)(g1 * t2[b] * t3[c]);
      f2 += (double)(g2 * d2[b] * t3[c]);
      f3 += (double)(g3 * t2[b] * d3[c]);
    }
  }
  // dE/dfr_k = qn * sum(...) * n_k ; dpot -= recip * f with f = -dE/dfr
  fpme[3 * n] = qn * f1 * (double)P.n1 * P.ra;
  fpme[3 * n + 1] = qn * f2 * (double)P.n2 * P.rb;
  fpme[3 * n + 2] = qn * f3 * (double)P.n3 * P.rc;
}

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

int32_t pme_init(moil_ctx *c) {
  const moil_params &P = c->par;
  c->nf1 = P.nfft1; c->nf2 = P.nfft2; c->nf3 = P.nfft3;
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
  const size_t nscan = (G + SCAN_PER - 1) / SCAN_PER;
  HIP_CHECK(hipMalloc(&c->d_pcnt, sizeof(int32_t) * (G + 1)));
  HIP_CHECK(hipMemset(c->d_pcnt, 0, sizeof(int32_t) * (G + 1)));
  HIP_CHECK(hipMalloc(&c->d_pstart, sizeof(int32_t) * (G + 1)));
  HIP_CHECK(hipMalloc(&c->d_pbsum, sizeof(int32_t) * (nscan + 1)));
  HIP_CHECK(hipMalloc(&c->d_abin, sizeof(int32_t) * na));
  HIP_CHECK(hipMalloc(&c->d_arank, sizeof(int32_t) * na));
  HIP_CHECK(hipMalloc(&c->d_aw, 4 * rs * na));
  HIP_CHECK(hipMalloc(&c->d_sw, 4 * rs * na));
  HIP_CHECK(hipMalloc(&c->d_sx, sizeof(int32_t) * na));
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
  hipFree(c->d_pcnt); c->d_pcnt = nullptr;
  hipFree(c->d_pstart); c->d_pstart = nullptr;
  hipFree(c->d_pbsum); c->d_pbsum = nullptr;
  hipFree(c->d_abin); c->d_abin = nullptr;
  hipFree(c->d_arank); c->d_arank = nullptr;
  hipFree(c->d_aw); c->d_aw = nullptr;
  hipFree(c->d_sw); c->d_sw = nullptr;
  hipFree(c->d_sx); c->d_sx = nullptr;
  hipFree(c->d_qreal); c->d_qreal = nullptr;
  hipFree(c->d_qcplx); c->d_qcplx = nullptr;
  hipFree(c->d_fpme); c->d_fpme = nullptr;
}

int pme_conv_blocks(const moil_ctx *c) {
  const size_t H = (size_t)c->nc1 * c->nf2 * c->nf3;
  return (int)((H + PME_THREADS - 1) / PME_THREADS);
}

template <typename R>
static int32_t pme_run_t(moil_ctx *c, const double *d_coor, hipStream_t st) {
  const size_t G = (size_t)c->nf1 * c->nf2 * c->nf3;
  const size_t H = (size_t)c->nc1 * c->nf2 * c->nf3;
  PmePar P;
  P.n1 = c->nf1; P.n2 = c->nf2; P.n3 = c->nf3; P.nc1 = c->nc1;
  P.ra = 1.0 / c->par.a; P.rb = 1.0 / c->par.b; P.rc = 1.0 / c->par.c;
  P.fc = (c->par.terms & MOIL_TERM_EL) ? sqrt(KOFDIE / c->par.eps) : 0.0;
  P.qscale = ldexp(1.0, QFIX_SHIFT);
  P.qinv = ldexp(1.0, -QFIX_SHIFT);
  const int gb = (int)((G + PME_THREADS - 1) / PME_THREADS);
  const int ab = (c->npt + PME_THREADS - 1) / PME_THREADS;
  const int nscan = (int)((G + SCAN_PER - 1) / SCAN_PER);
  k_pme_bin<R><<<ab, PME_THREADS, 0, st>>>(d_coor, c->d_ptchg, c->npt, P, c->d_pcnt, c->d_abin,
                                           c->d_arank, (R *)c->d_aw);
  k_scan_sums<<<nscan, 256, 0, st>>>(c->d_pcnt, G, c->d_pbsum);
  k_scan_top<<<1, 1024, 0, st>>>(c->d_pbsum, nscan);
  k_scan_apply<<<nscan, 256, 0, st>>>(c->d_pcnt, G, c->d_pbsum, c->d_pstart);
  k_pme_scatter<R><<<ab, PME_THREADS, 0, st>>>(c->npt, c->d_abin, c->d_arank, c->d_pstart,
                                               (const R *)c->d_aw, (R *)c->d_sw, c->d_sx, c->nf1);
  k_pme_spread<R><<<gb, PME_THREADS, 0, st>>>(c->d_pstart, (const R *)c->d_sw, c->d_sx, P,
                                              (R *)c->d_qreal);
  HIP_CHECK(hipGetLastError());
  void *in[1] = {c->d_qreal}, *out[1] = {c->d_qcplx};
  RF_CHECK(rocfft_execution_info_set_stream(c->exec_fwd, st));
  RF_CHECK(rocfft_execute(c->plan_fwd, in, out, c->exec_fwd));
  const int hb = (int)((H + PME_THREADS - 1) / PME_THREADS);
  k_conv<R><<<hb, PME_THREADS, 0, st>>>((R *)c->d_qcplx, c->d_eterm, c->nc1, c->nf1, H,
                                          c->d_epart + (size_t)EN_RECIP * c->max_parts);
  void *in2[1] = {c->d_qcplx}, *out2[1] = {c->d_qreal};
  RF_CHECK(rocfft_execution_info_set_stream(c->exec_bwd, st));
  RF_CHECK(rocfft_execute(c->plan_bwd, in2, out2, c->exec_bwd));
  k_gather<R><<<ab, PME_THREADS, 0, st>>>(d_coor, c->d_ptchg, c->npt, P, (const R *)c->d_qreal,
                                          c->d_fpme);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

int32_t pme_run(moil_ctx *c, const double *d_coor, hipStream_t st) {
  if (c->par.precision == MOIL_PREC_MIXED) return pme_run_t<float>(c, d_coor, st);
  return pme_run_t<double>(c, d_coor, st);
}
