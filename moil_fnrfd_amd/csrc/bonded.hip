// Bonded terms, scaled 1-4 pairs and the Ewald exclusion correction, in
// FP64 for both precision modes (these are O(N) and carry the large
// cancelling e_corr vs e_self energies):
//   ebond   pot/ebond.f:33-72     E = k (r - r0)^2
//   etheta  pot/etheta.f:25-113   E = k (theta - theta0)^2, cos clamp,
//                                  sin floor 1e-6
//   etors   pot/ephi.f:66-277     E = sum k_n + sum phase_n k_n T_n(cos phi)
//   eimphi  pot/eimphi.f:86-249   phi0 = 0 / pi special forms, else
//                                  k (phi - phi0)^2 with signed phi
//   ener14  pot/ener14.f:38-95    p14 LJ + plain Coulomb
//   e_corr  pot/cdie_ewald.f:83-199  -q_i q_j erf(beta r)/r over the
//                                  exclusion list (atoms 1..npt-1)
// One thread per term; gradients go to the 64-bit fixed-point
// accumulators with integer atomics, energies to per-workgroup partials.
#include <hip/hip_runtime.h>

#include "ctx.h"

#define BT 256

struct BPar {
  double fix_scale;
  double epstmp14, pick;  // ener14 switches
  double qfac;            // kofdie/eps (exclusion correction)
  double beta, erftbdns;
  int do_el14, do_v14;
  int max_parts;
};

__device__ __forceinline__ void add3(long long *facc, int atom, double fx, double fy, double fz,
                                     double s) {
  long long v;
  v = __double2ll_rn(fx * s);
  if (v) atomicAdd((unsigned long long *)&facc[3 * atom], (unsigned long long)v);
  v = __double2ll_rn(fy * s);
  if (v) atomicAdd((unsigned long long *)&facc[3 * atom + 1], (unsigned long long)v);
  v = __double2ll_rn(fz * s);
  if (v) atomicAdd((unsigned long long *)&facc[3 * atom + 2], (unsigned long long)v);
}

__device__ inline void block_partial(double e, double *out) {
  __shared__ double red[BT];
  int t = threadIdx.x;
  red[t] = e;
  __syncthreads();
  for (int off = BT / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = red[0];
}

#define XYZ(a) coor[3 * (a)], coor[3 * (a) + 1], coor[3 * (a) + 2]

__global__ __launch_bounds__(BT) void k_bonds(const double *__restrict__ coor, const int32_t *__restrict__ ib,
                                              const double *__restrict__ par, int nb, BPar P,
                                              long long *__restrict__ facc, double *__restrict__ ep) {
  int m = blockIdx.x * BT + threadIdx.x;
  double e = 0;
  if (m < nb) {
    int i = ib[2 * m] - 1, j = ib[2 * m + 1] - 1;
    double rx = coor[3 * i] - coor[3 * j], ry = coor[3 * i + 1] - coor[3 * j + 1],
           rz = coor[3 * i + 2] - coor[3 * j + 2];
    double s = sqrt(rx * rx + ry * ry + rz * rz);
    double db = s - par[2 * m + 1];
    double df = par[2 * m] * db;
    e = df * db;
    df = df * (2.0 / s);
    add3(facc, i, rx * df, ry * df, rz * df, P.fix_scale);
    add3(facc, j, -rx * df, -ry * df, -rz * df, P.fix_scale);
  }
  block_partial(e, ep);
}

__global__ __launch_bounds__(BT) void k_angles(const double *__restrict__ coor, const int32_t *__restrict__ ia,
                                               const double *__restrict__ par, int n, BPar P,
                                               long long *__restrict__ facc, double *__restrict__ ep) {
  int m = blockIdx.x * BT + threadIdx.x;
  double e = 0;
  if (m < n) {
    int i = ia[3 * m] - 1, j = ia[3 * m + 1] - 1, k = ia[3 * m + 2] - 1;
    double di[3], dj[3];
    for (int q = 0; q < 3; q++) {
      di[q] = coor[3 * i + q] - coor[3 * j + q];
      dj[q] = coor[3 * k + q] - coor[3 * j + q];
    }
    double ri = sqrt(di[0] * di[0] + di[1] * di[1] + di[2] * di[2]);
    double rj = sqrt(dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2]);
    double rir = 1.0 / ri, rjr = 1.0 / rj;
    for (int q = 0; q < 3; q++) {
      di[q] *= rir;
      dj[q] *= rjr;
    }
    double cst = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
    cst = fmin(1.0, fmax(-1.0, cst));
    double at = acos(cst);
    double da = at - par[2 * m + 1];
    double df = par[2 * m] * da;
    e = df * da;
    df = df + df;
    double s = sin(at);
    if (s < 1e-6 && s > 0.0) s = 1e-6;
    else if (s < 0.0 && s > -1e-6) s = -1e-6;
    df = -df / s;
    double t1[3], t2[3];
    for (int q = 0; q < 3; q++) {
      t1[q] = rir * (dj[q] - cst * di[q]);
      t2[q] = rjr * (di[q] - cst * dj[q]);
    }
    add3(facc, i, df * t1[0], df * t1[1], df * t1[2], P.fix_scale);
    add3(facc, j, -df * (t1[0] + t2[0]), -df * (t1[1] + t2[1]), -df * (t1[2] + t2[2]), P.fix_scale);
    add3(facc, k, df * t2[0], df * t2[1], df * t2[2], P.fix_scale);
  }
  block_partial(e, ep);
}

// cos(phi) of the dihedral and its gradient (chain rule on
// u=(a.a)(b.b)-(a.b)^2, v=(b.b)(c.c)-(b.c)^2, w=(a.b)(b.c)-(a.c)(b.b),
// a=x_j-x_i, b=x_k-x_j, c=x_l-x_k; pot/ephi.f:88-109 defines cos phi)
__device__ inline double dihedral(const double *xi, const double *xj, const double *xk,
                                  const double *xl, double g[4][3], double *sgn) {
  double a[3], b[3], c[3];
  for (int q = 0; q < 3; q++) {
    a[q] = xj[q] - xi[q];
    b[q] = xk[q] - xj[q];
    c[q] = xl[q] - xk[q];
  }
  double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
  double ac = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
  double u = aa * bb - ab * ab, v = bb * cc - bc * bc, w = ab * bc - ac * bb;
  double inv = 1.0 / sqrt(u * v);
  double co = w * inv;
  for (int q = 0; q < 3; q++) {
    double dwa = bc * b[q] - bb * c[q];
    double dwb = ab * c[q] + bc * a[q] - 2.0 * ac * b[q];
    double dwc = ab * b[q] - bb * a[q];
    double dua = 2.0 * (bb * a[q] - ab * b[q]);
    double dub = 2.0 * (aa * b[q] - ab * a[q]);
    double dvb = 2.0 * (cc * b[q] - bc * c[q]);
    double dvc = 2.0 * (bb * c[q] - bc * b[q]);
    double gA = dwa * inv - 0.5 * co * (dua / u);
    double gB = dwb * inv - 0.5 * co * (dub / u + dvb / v);
    double gC = dwc * inv - 0.5 * co * (dvc / v);
    g[0][q] = -gA;
    g[1][q] = gA - gB;
    g[2][q] = gB - gC;
    g[3][q] = gC;
  }
  if (sgn) {
    double ux = a[1] * b[2] - a[2] * b[1], uy = a[2] * b[0] - a[0] * b[2], uz = a[0] * b[1] - a[1] * b[0];
    double vx = b[1] * c[2] - b[2] * c[1], vy = b[2] * c[0] - b[0] * c[2], vz = b[0] * c[1] - b[1] * c[0];
    double dx = uy * vz - uz * vy, dy = uz * vx - ux * vz, dz = ux * vy - uy * vx;
    *sgn = dx * b[0] + dy * b[1] + dz * b[2];
  }
  return co;
}

__global__ __launch_bounds__(BT) void k_tors(const double *__restrict__ coor, const int32_t *__restrict__ it,
                                             const double *__restrict__ par, int n, BPar P,
                                             long long *__restrict__ facc, double *__restrict__ ep) {
  int m = blockIdx.x * BT + threadIdx.x;
  double e = 0;
  if (m < n) {
    int id[4] = {it[4 * m] - 1, it[4 * m + 1] - 1, it[4 * m + 2] - 1, it[4 * m + 3] - 1};
    double g[4][3];
    double co = dihedral(&coor[3 * id[0]], &coor[3 * id[1]], &coor[3 * id[2]], &coor[3 * id[3]], g, nullptr);
    const double *p = par + 6 * m;  // k1 k2 k3 ph1 ph2 ph3
    e = (p[0] + p[1] + p[2]) + p[3] * (p[0] * co) + p[4] * (p[1] * (2.0 * co * co - 1.0)) +
        p[5] * (p[2] * co * (4.0 * co * co - 3.0));
    double df = p[3] * p[0] + p[4] * p[1] * 4.0 * co + p[5] * p[2] * (12.0 * co * co - 3.0);
    for (int q = 0; q < 4; q++) add3(facc, id[q], df * g[q][0], df * g[q][1], df * g[q][2], P.fix_scale);
  }
  block_partial(e, ep);
}

__global__ __launch_bounds__(BT) void k_imps(const double *__restrict__ coor, const int32_t *__restrict__ ii,
                                             const double *__restrict__ par, int n, BPar P,
                                             long long *__restrict__ facc, double *__restrict__ ep) {
  int m = blockIdx.x * BT + threadIdx.x;
  double e = 0;
  if (m < n) {
    const double pi = 3.14159265358979323846;
    int id[4] = {ii[4 * m] - 1, ii[4 * m + 1] - 1, ii[4 * m + 2] - 1, ii[4 * m + 3] - 1};
    double g[4][3], sgn;
    double co = dihedral(&coor[3 * id[0]], &coor[3 * id[1]], &coor[3 * id[2]], &coor[3 * id[3]], g, &sgn);
    co = fmin(1.0, fmax(-1.0, co));
    double k = par[2 * m], eq = par[2 * m + 1], df;
    if (fabs(eq) < 1e-6) {
      df = -2.0 * k;
      e = -df * (1.0 - co);
    } else if (fabs(fabs(eq) - pi) < 1e-6) {
      df = 2.0 * k;
      e = df * (1.0 + co);
    } else {
      double phi = acos(co);
      if (sgn < 0) phi = -phi;
      double delta = phi - eq;
      if (delta > pi) delta -= 2.0 * pi;
      if (delta < -pi) delta += 2.0 * pi;
      df = k * delta;
      e = df * delta;
      double yy = sin(phi);
      if (fabs(yy) > 1e-6) df = -2.0 * df / yy;
      else if ((phi > 0.0 && phi < pi / 2.0) || (phi < 0.0 && phi > -pi / 2.0)) df = df * 1e6;
      else df = -df * 1e6;
    }
    for (int q = 0; q < 4; q++) add3(facc, id[q], df * g[q][0], df * g[q][1], df * g[q][2], P.fix_scale);
  }
  block_partial(e, ep);
}

__global__ __launch_bounds__(BT) void k_e14(const double *__restrict__ coor, const int32_t *__restrict__ sp,
                                            const double *__restrict__ p14, int n, BPar P,
                                            long long *__restrict__ facc, double *__restrict__ epv,
                                            double *__restrict__ epe) {
  int m = blockIdx.x * BT + threadIdx.x;
  double e1 = 0, e2 = 0;
  if (m < n) {
    int j = sp[2 * m] - 1, k = sp[2 * m + 1] - 1;
    double rx = coor[3 * j] - coor[3 * k], ry = coor[3 * j + 1] - coor[3 * k + 1],
           rz = coor[3 * j + 2] - coor[3 * k + 2];
    double r2 = rx * rx + ry * ry + rz * rz, s2 = 1.0 / r2, df1 = 0, df2 = 0;
    const double *p = p14 + 3 * m;
    if (P.do_v14 && p[0] > 1e-12) {
      double s6 = s2 * s2 * s2, A = p[0] * s6 * s6 * P.pick, B = p[1] * s6 * P.pick;
      e1 = A - B;
      df1 = -6.0 * s2 * (A + e1);
    }
    if (P.do_el14 && fabs(p[2]) > 1e-12) {
      double s = sqrt(s2);
      e2 = p[2] * s * P.epstmp14;
      df2 = -e2 * s2;
    }
    double df = df1 + df2;
    add3(facc, j, df * rx, df * ry, df * rz, P.fix_scale);
    add3(facc, k, -df * rx, -df * ry, -df * rz, P.fix_scale);
  }
  block_partial(e1, epv);
  block_partial(e2, epe);
}

__global__ __launch_bounds__(BT) void k_excl(const double *__restrict__ coor, const int32_t *__restrict__ ei,
                                             const int32_t *__restrict__ ej, const double *__restrict__ q,
                                             int n, BPar P, const double *__restrict__ erf_arr,
                                             long long *__restrict__ facc, double *__restrict__ ep) {
  int m = blockIdx.x * BT + threadIdx.x;
  double e = 0;
  if (m < n) {
    int i = ei[m], j = ej[m];
    double rx = coor[3 * i] - coor[3 * j], ry = coor[3 * i + 1] - coor[3 * j + 1],
           rz = coor[3 * i + 2] - coor[3 * j + 2];
    double r2 = rx * rx + ry * ry + rz * rz, s2 = 1.0 / r2, rij = sqrt(r2), s = 1.0 / rij;
    double qq = q[i] * P.qfac * q[j];
    double xerfc = P.beta * rij;
    int idx = (int)(P.erftbdns * xerfc);
    double dx = xerfc - idx * (1.0 / P.erftbdns);
    const double *t = erf_arr + 4 * idx;
    double derfc = -t[1] - dx * (t[2] + 0.5 * dx * t[3]);
    double erfcc = dx * (t[1] + 0.5 * dx * (t[2] + dx * t[3] * (1.0 / 3.0))) + t[0];
    double erf = 1.0 - erfcc;
    e = -qq * s * erf;
    double df = qq * s * (s2 * erf - s * P.beta * derfc);
    add3(facc, i, df * rx, df * ry, df * rz, P.fix_scale);
    add3(facc, j, -df * rx, -df * ry, -df * rz, P.fix_scale);
  }
  block_partial(e, ep);
}

int32_t bonded_run(moil_ctx *c, const double *d_coor, hipStream_t st) {
  BPar P;
  const uint32_t T = c->par.terms;
  P.fix_scale = c->fix_scale;
  P.epstmp14 = (T & MOIL_TERM_EL) ? 1.0 : 0.0;
  P.pick = (T & MOIL_TERM_VDW) ? 1.0 : 0.0;
  P.qfac = (T & MOIL_TERM_EL) ? KOFDIE / c->par.eps : 0.0;
  P.beta = c->beta;
  P.erftbdns = c->erftbdns;
  P.do_el14 = (T & MOIL_TERM_EL14) != 0;
  P.do_v14 = (T & MOIL_TERM_VDW14) != 0;
  P.max_parts = c->max_parts;
  double *ep = c->d_epart;
  const size_t MP = c->max_parts;
  if ((T & MOIL_TERM_BOND) && c->nb > 0)
    k_bonds<<<(c->nb + BT - 1) / BT, BT, 0, st>>>(d_coor, c->d_ib, c->d_bpar, c->nb, P, c->d_facc,
                                                  ep + EN_BOND * MP);
  if ((T & MOIL_TERM_ANGLE) && c->nangl > 0)
    k_angles<<<(c->nangl + BT - 1) / BT, BT, 0, st>>>(d_coor, c->d_ia, c->d_apar, c->nangl, P,
                                                      c->d_facc, ep + EN_THETA * MP);
  if ((T & MOIL_TERM_TORS) && c->ntors > 0)
    k_tors<<<(c->ntors + BT - 1) / BT, BT, 0, st>>>(d_coor, c->d_it, c->d_tpar, c->ntors, P,
                                                    c->d_facc, ep + EN_TORS * MP);
  if ((T & MOIL_TERM_IMP) && c->nimp > 0)
    k_imps<<<(c->nimp + BT - 1) / BT, BT, 0, st>>>(d_coor, c->d_ii, c->d_ipar, c->nimp, P,
                                                   c->d_facc, ep + EN_IMP * MP);
  if ((T & (MOIL_TERM_VDW | MOIL_TERM_EL))) {
    if (c->nexcl > 0)
      k_excl<<<(c->nexcl + BT - 1) / BT, BT, 0, st>>>(d_coor, c->d_exc_i, c->d_exc_j, c->d_ptchg,
                                                      c->nexcl, P, c->d_erf, c->d_facc,
                                                      ep + EN_CORR * MP);
    if (c->totspe > 0 && (P.do_el14 || P.do_v14))
      k_e14<<<(c->totspe + BT - 1) / BT, BT, 0, st>>>(d_coor, c->d_sp, c->d_p14, c->totspe, P,
                                                      c->d_facc, ep + EN_VDW14 * MP,
                                                      ep + EN_EL14 * MP);
  }
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

int bonded_max_blocks(const moil_ctx *c) {
  int m = 1;
  int ns[6] = {c->nb, c->nangl, c->ntors, c->nimp, c->nexcl, c->totspe};
  for (int k = 0; k < 6; k++) m = max(m, (ns[k] + BT - 1) / BT);
  return m;
}
