// Bonded terms, scaled 1-4 pairs and the Ewald exclusion correction, in
// FP64 for both precision modes (these are O(N) and carry the large
// cancelling e_corr vs e_self energies):
//   ebond   pot/ebond.f:33-72     E = k (r - r0)^2
//   etheta  pot/etheta.f:25-113   E = k (theta - theta0)^2, cos clamp,
//                                  sin floor 1e-6
//   etors   pot/ephi.f:66-277     E = sum k_n + sum phase_n k_n T_n(cos phi)
//   eimphi  pot/eimphi.f:86-249   phi0 = 0 / pi special forms, else
//                                  k (phi - phi0)^2 with signed phi
//   arith forms (MOIL_TERM_ARITH): torsion k2 + phase2 k2 cos(n phi),
//   n = 1..4 (pot/ephi.f:113-138); improper k - k cos(2 phi) with MOIL's
//   atom order (pot/eimphi.f:98-115,251-262)
//   ener14  pot/ener14.f:38-95    p14 LJ + plain Coulomb
//   e_corr  pot/cdie_ewald.f:83-199  -q_i q_j erf(beta r)/r over the
//                                  exclusion list (atoms 1..npt-1)
// Intra-water terms (bonds, angle, exclusion corrections inside a 3-site
// water monomer) run one thread per water: the thread owns the water's
// three atoms, sums their per-term fixed-point gradients in registers and
// stores them (this kernel runs first in an evaluation, on the cleared
// accumulators).  All other terms run one thread per term with integer
// atomics.  Each term's gradient is rounded to fixed point on its own in
// both kernels, so the integer sums do not depend on the grouping.
// Energies go to per-workgroup partials.
#include <hip/hip_runtime.h>

#include "multi.h"
#include "pairmath.h"

#define BT 256


// gradient sinks of the term functions: the registers of a thread that
// owns atoms a0 .. a0+2, or the term's own gradient slots
// (one 3-vector per atom position of the term, written whole -- zeros
// included -- so they need no clearing; the assembly kernel sums each
// atom's slots through a per-atom list built at set-up)
struct SlotSink {
  long long acc[4][3];
  int id[4];
  int n;
  double s;
  int *ovf;
  bool bad = false;
  __device__ SlotSink(int n_, int i0, int i1, int i2, int i3, double s_, int *ovf_)
      : n(n_), s(s_), ovf(ovf_) {
    id[0] = i0; id[1] = i1; id[2] = i2; id[3] = i3;
#pragma unroll
    for (int p = 0; p < 4; p++) acc[p][0] = acc[p][1] = acc[p][2] = 0;
  }
  __device__ void add(int atom, double fx, double fy, double fz) {
    const int p = atom == id[0] ? 0 : (atom == id[1] ? 1 : (atom == id[2] ? 2 : 3));
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (q == p) {
        acc[q][0] += fixd(fx, s, bad);
        acc[q][1] += fixd(fy, s, bad);
        acc[q][2] += fixd(fz, s, bad);
      }
  }
  __device__ void flush(long long *__restrict__ out) const {
    fix_report(bad, ovf);
#pragma unroll
    for (int p = 0; p < 4; p++)
      if (p < n) {
        out[3 * p] = acc[p][0];
        out[3 * p + 1] = acc[p][1];
        out[3 * p + 2] = acc[p][2];
      }
  }
};
struct WaterSink {
  long long acc[9];
  int a0;
  double s;
  bool bad;
  __device__ void add(int atom, double fx, double fy, double fz) {
    const int l = atom - a0;
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (k == l) {
        acc[3 * k] += fixd(fx, s, bad);
        acc[3 * k + 1] += fixd(fy, s, bad);
        acc[3 * k + 2] += fixd(fz, s, bad);
      }
  }
};

#define XYZ(a) coor[3 * (a)], coor[3 * (a) + 1], coor[3 * (a) + 2]

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


template <class Sink>
__device__ inline void bond_calc(const double *coor, int i, int j, double kb, double r0, Sink &sk,
                                 double &e) {
  double rx = coor[3 * i] - coor[3 * j], ry = coor[3 * i + 1] - coor[3 * j + 1],
         rz = coor[3 * i + 2] - coor[3 * j + 2];
  double s = sqrt(rx * rx + ry * ry + rz * rz);
  double db = s - r0;
  double df = kb * db;
  e = df * db;
  df = df * (2.0 / s);
  sk.add(i, rx * df, ry * df, rz * df);
  sk.add(j, -rx * df, -ry * df, -rz * df);
}

__device__ inline void term_bond(const double *coor, const BSeg &S, int m, const BPar &P,
                                 long long *out, double &e) {
  const int i = S.ib[2 * m] - 1, j = S.ib[2 * m + 1] - 1;
  SlotSink sk(2, i, j, -1, -1, P.fix_scale, P.ovf);
  bond_calc(coor, i, j, S.bp[2 * m], S.bp[2 * m + 1], sk, e);
  sk.flush(out);
}

template <class Sink>
__device__ inline void angle_calc(const double *coor, int i, int j, int k, double ka, double th0,
                                  Sink &sk, double &e) {
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
  double da = at - th0;
  double df = ka * da;
  e = df * da;
  df = df + df;
  double sn = sin(at);
  if (sn < 1e-6 && sn > 0.0) sn = 1e-6;
  else if (sn < 0.0 && sn > -1e-6) sn = -1e-6;
  df = -df / sn;
  double t1[3], t2[3];
  for (int q = 0; q < 3; q++) {
    t1[q] = rir * (dj[q] - cst * di[q]);
    t2[q] = rjr * (di[q] - cst * dj[q]);
  }
  sk.add(i, df * t1[0], df * t1[1], df * t1[2]);
  sk.add(j, -df * (t1[0] + t2[0]), -df * (t1[1] + t2[1]), -df * (t1[2] + t2[2]));
  sk.add(k, df * t2[0], df * t2[1], df * t2[2]);
}

__device__ inline void term_angle(const double *coor, const BSeg &S, int m, const BPar &P,
                                  long long *out, double &e) {
  const int i = S.ia[3 * m] - 1, j = S.ia[3 * m + 1] - 1, k = S.ia[3 * m + 2] - 1;
  SlotSink sk(3, i, j, k, -1, P.fix_scale, P.ovf);
  angle_calc(coor, i, j, k, S.ap[2 * m], S.ap[2 * m + 1], sk, e);
  sk.flush(out);
}

__device__ inline void term_tors(const double *coor, const BSeg &S, int m, const BPar &P,
                                 long long *out, double &e) {
  int id[4] = {S.it[4 * m] - 1, S.it[4 * m + 1] - 1, S.it[4 * m + 2] - 1, S.it[4 * m + 3] - 1};
  double g[4][3];
  double co = dihedral(&coor[3 * id[0]], &coor[3 * id[1]], &coor[3 * id[2]], &coor[3 * id[3]], g, nullptr);
  const double *p = S.tp + 6 * m;  // k1 k2 k3 ph1 ph2 ph3 (arith: period in k1's place)
  double df;
  if (P.arith) {
    // k2 + phase2 k2 T_n(cos phi), n = 1..4 (pot/ephi.f:113-138,242-259)
    const int n = (int)p[0];
    const double k2 = p[1], h2 = p[4], c2 = 2.0 * co * co - 1.0;
    if (n == 1) {
      e = k2 + h2 * k2 * co;
      df = k2 * h2;
    } else if (n == 2) {
      e = k2 + h2 * k2 * c2;
      df = 4.0 * k2 * h2 * co;
    } else if (n == 3) {
      e = k2 + h2 * k2 * co * (4.0 * co * co - 3.0);
      df = k2 * h2 * (12.0 * co * co - 3.0);
    } else {
      e = k2 + h2 * k2 * (2.0 * c2 * c2 - 1.0);
      df = 16.0 * k2 * h2 * co * c2;
    }
  } else {
    e = (p[0] + p[1] + p[2]) + p[3] * (p[0] * co) + p[4] * (p[1] * (2.0 * co * co - 1.0)) +
        p[5] * (p[2] * co * (4.0 * co * co - 3.0));
    df = p[3] * p[0] + p[4] * p[1] * 4.0 * co + p[5] * p[2] * (12.0 * co * co - 3.0);
  }
  SlotSink sk(4, id[0], id[1], id[2], id[3], P.fix_scale, P.ovf);
  for (int q = 0; q < 4; q++) sk.add(id[q], df * g[q][0], df * g[q][1], df * g[q][2]);
  sk.flush(out);
}

__device__ inline void term_imp(const double *coor, const BSeg &S, int m, const BPar &P,
                                long long *out, double &e) {
  const double pi = 3.14159265358979323846;
  int id[4] = {S.ii[4 * m] - 1, S.ii[4 * m + 1] - 1, S.ii[4 * m + 2] - 1, S.ii[4 * m + 3] - 1};
  double g[4][3], sgn;
  double co = dihedral(&coor[3 * id[0]], &coor[3 * id[1]], &coor[3 * id[2]], &coor[3 * id[3]], g, &sgn);
  co = fmin(1.0, fmax(-1.0, co));
  double k = S.ip[2 * m], eq = S.ip[2 * m + 1], df;
  if (P.arith) {
    // k - k cos(2 phi) over the atoms (iimp4, iimp2, iimp1, iimp3), put in
    // that order at set-up (pot/eimphi.f:110-115,251-262)
    e = k - k * (2.0 * co * co - 1.0);
    df = -4.0 * k * co;
  } else if (fabs(eq) < 1e-6) {
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
  SlotSink sk(4, id[0], id[1], id[2], id[3], P.fix_scale, P.ovf);
  for (int q = 0; q < 4; q++) sk.add(id[q], df * g[q][0], df * g[q][1], df * g[q][2]);
  sk.flush(out);
}

template <class Sink>
__device__ inline void excl_calc(const double *coor, int i, int j, const double *q,
                                 const double *erf_arr, const BPar &P, Sink &sk, double &e) {
  double rx = coor[3 * i] - coor[3 * j], ry = coor[3 * i + 1] - coor[3 * j + 1],
         rz = coor[3 * i + 2] - coor[3 * j + 2];
  double r2 = rx * rx + ry * ry + rz * rz, s2 = 1.0 / r2, rij = sqrt(r2), s = 1.0 / rij;
  double qq = q[i] * P.qfac * q[j];
  double erfcc, derfc;
  erfc_table(erf_arr, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
  double erf = 1.0 - erfcc;
  e = -qq * s * erf;
  double df = qq * s * (s2 * erf - s * P.beta * derfc);
  sk.add(i, df * rx, df * ry, df * rz);
  sk.add(j, -df * rx, -df * ry, -df * rz);
}

__device__ inline void term_excl(const double *coor, const BSeg &S, int m, const BPar &P,
                                 long long *out, double &e) {
  SlotSink sk(2, S.ei[m], S.ej[m], -1, -1, P.fix_scale, P.ovf);
  excl_calc(coor, S.ei[m], S.ej[m], S.q, S.erf, P, sk, e);
  sk.flush(out);
}

__device__ inline void term_14(const double *coor, const BSeg &S, int m, const BPar &P,
                               long long *out, double &e1, double &e2) {
  int j = S.sp[2 * m] - 1, k = S.sp[2 * m + 1] - 1;
  double rx = coor[3 * j] - coor[3 * k], ry = coor[3 * j + 1] - coor[3 * k + 1],
         rz = coor[3 * j + 2] - coor[3 * k + 2];
  double r2 = rx * rx + ry * ry + rz * rz, s2 = 1.0 / r2, df1 = 0, df2 = 0;
  const double *p = S.p14 + 3 * m;
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
  SlotSink sk(2, j, k, -1, -1, P.fix_scale, P.ovf);
  sk.add(j, df * rx, df * ry, df * rz);
  sk.add(k, -df * rx, -df * ry, -df * rz);
  sk.flush(out);
}

// deterministic workgroup sums of N per-thread values (wavefront shuffle
// tree, then the BT/64 wavefront totals in order); result in red[0..N-1]
template <int N, int NT = BT>
__device__ inline void block_sums(const double *v, double *red) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int q = 0; q < N; q++) {
    double x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[q * (NT / 64) + wv] = x;
  }
  __syncthreads();
  double tot = 0.0;
  if (t < N)
    for (int k = 0; k < NT / 64; k++) tot += red[t * (NT / 64) + k];
  __syncthreads();
  if (t < N) red[t] = tot;
  __syncthreads();
}

#define WIT 64  // threads per workgroup of the per-water kernel


// all O(N) terms in one launch over a flat index space
// [bonds | angles | torsions | impropers | exclusions | 1-4 pairs]; term
// m of a type writes the gradient slots bslot_base(type) + arity * m
// (arities 2, 3, 4, 4, 2, 2), the layout of the per-atom slot lists
__device__ __forceinline__ void bonded_impl(const double *__restrict__ coor, const BSeg &S,
                                            const BPar &P, long long *__restrict__ bslot,
                                            double *__restrict__ ep, int mp, int bid) {
  __shared__ double red[BT];
  int m = bid * BT + threadIdx.x;
  double e[7] = {0, 0, 0, 0, 0, 0, 0};  // bond theta tors imp corr vdw14 el14
  int o1 = S.nb, o2 = o1 + S.nangl, o3 = o2 + S.ntors, o4 = o3 + S.nimp, o5 = o4 + S.nexcl,
      o6 = o5 + S.n14;
  const size_t b1 = 2 * (size_t)S.nb, b2 = b1 + 3 * (size_t)S.nangl, b3 = b2 + 4 * (size_t)S.ntors,
               b4 = b3 + 4 * (size_t)S.nimp, b5 = b4 + 2 * (size_t)S.nexcl;
  if (m < o1) term_bond(coor, S, m, P, bslot + 3 * (2 * (size_t)m), e[0]);
  else if (m < o2) term_angle(coor, S, m - o1, P, bslot + 3 * (b1 + 3 * (size_t)(m - o1)), e[1]);
  else if (m < o3) term_tors(coor, S, m - o2, P, bslot + 3 * (b2 + 4 * (size_t)(m - o2)), e[2]);
  else if (m < o4) term_imp(coor, S, m - o3, P, bslot + 3 * (b3 + 4 * (size_t)(m - o3)), e[3]);
  else if (m < o5) term_excl(coor, S, m - o4, P, bslot + 3 * (b4 + 2 * (size_t)(m - o4)), e[4]);
  else if (m < o6)
    term_14(coor, S, m - o5, P, bslot + 3 * (b5 + 2 * (size_t)(m - o5)), e[5], e[6]);
  const int slot[7] = {EN_BOND, EN_THETA, EN_TORS, EN_IMP, EN_CORR, EN_VDW14, EN_EL14};
  block_sums<7>(e, red);
  if (threadIdx.x < 7) ep[(size_t)slot[threadIdx.x] * mp + bid] = red[threadIdx.x];
}
__device__ __forceinline__ void bonded_body(const BondedArgs &A, int bid, int nblk) {
  bonded_impl(A.coor, A.S, A.P, A.bslot, A.ep, A.mp, bid);
}
DEFINE_KERNEL(k_bonded, BondedArgs, BT, bonded_body)

// one thread per 3-site water: its intra-water bonds, angles and
// exclusion corrections (terms listed per water in CSR form, local atom
// numbers 0..2); the water's gradients are stored, not added
__device__ __forceinline__ void water_intra_impl(const double *__restrict__ coor, const WIntra &W,
                                                 const BPar &P, long long *__restrict__ facc,
                                                 double *__restrict__ ep, int mp, int part0,
                                                 int bid) {
  // it runs on the second stream beside the water pair kernel, ahead of
  // the reciprocal sum: its waves take the SIMDs' issue slots first
  __builtin_amdgcn_s_setprio(2);
  __shared__ double red[WIT];
  const int w = bid * WIT + threadIdx.x;
  double e[3] = {0, 0, 0};  // bond theta corr
  if (w < W.nwat) {
    const int a0 = W.first[w];
    WaterSink sk;
    sk.a0 = a0;
    sk.s = P.fix_scale;
    sk.bad = false;
#pragma unroll
    for (int k = 0; k < 9; k++) sk.acc[k] = 0;
    double et;
    for (int m = W.boff[w]; m < W.boff[w + 1]; m++) {
      bond_calc(coor, a0 + W.bidx[2 * m], a0 + W.bidx[2 * m + 1], W.bpar[2 * m],
                W.bpar[2 * m + 1], sk, et);
      e[0] += et;
    }
    for (int m = W.aoff[w]; m < W.aoff[w + 1]; m++) {
      angle_calc(coor, a0 + W.aidx[3 * m], a0 + W.aidx[3 * m + 1], a0 + W.aidx[3 * m + 2],
                 W.apar[2 * m], W.apar[2 * m + 1], sk, et);
      e[1] += et;
    }
    for (int m = W.eoff[w]; m < W.eoff[w + 1]; m++) {
      excl_calc(coor, a0 + W.eidx[2 * m], a0 + W.eidx[2 * m + 1], W.q, W.erf, P, sk, et);
      e[2] += et;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) facc[3 * a0 + k] = sk.acc[k];
    fix_report(sk.bad, P.ovf);
  }
  block_sums<3, WIT>(e, red);
  const int slot[3] = {EN_BOND, EN_THETA, EN_CORR};
  if (threadIdx.x < 3) ep[(size_t)slot[threadIdx.x] * mp + part0 + bid] = red[threadIdx.x];
}
__device__ __forceinline__ void water_intra_body(const IntraArgs &A, int bid, int nblk) {
  water_intra_impl(A.coor, A.W, A.P, A.facc, A.ep, A.mp, A.part0, bid);
}
DEFINE_KERNEL(k_water_intra, IntraArgs, WIT, water_intra_body)

static BSeg make_seg(const moil_ctx *c) {
  const uint32_t T = c->par.terms;
  BSeg S;
  const bool nbon = (T & (MOIL_TERM_VDW | MOIL_TERM_EL)) != 0;
  S.nb = (T & MOIL_TERM_BOND) ? c->gnb : 0;
  S.nangl = (T & MOIL_TERM_ANGLE) ? c->gnangl : 0;
  S.ntors = (T & MOIL_TERM_TORS) ? c->ntors : 0;
  S.nimp = (T & MOIL_TERM_IMP) ? c->nimp : 0;
  S.nexcl = nbon ? c->gnexcl : 0;
  S.n14 = (nbon && (T & (MOIL_TERM_EL14 | MOIL_TERM_VDW14))) ? c->totspe : 0;
  S.ib = c->d_ib; S.ia = c->d_ia; S.it = c->d_it; S.ii = c->d_ii;
  S.ei = c->d_exc_i; S.ej = c->d_exc_j; S.sp = c->d_sp;
  S.bp = c->d_bpar; S.ap = c->d_apar; S.tp = c->d_tpar; S.ip = c->d_ipar;
  S.p14 = c->d_p14; S.q = c->d_ptchg; S.erf = c->d_erf;
  return S;
}

int bonded_blocks(const moil_ctx *c) {
  BSeg S = make_seg(c);
  int n = S.nb + S.nangl + S.ntors + S.nimp + S.nexcl + S.n14;
  return (n + BT - 1) / BT;
}

static bool water_intra_on(const moil_ctx *c) {
  return c->nwat > 0 && c->wi_nterms > 0;
}

int water_intra_blocks(const moil_ctx *c) {
  return water_intra_on(c) ? (c->nwat + WIT - 1) / WIT : 0;
}

// the bonded launches of one evaluation: the intra-water terms first
// (they store into the cleared accumulators), then the generic terms
void bonded_plan(const moil_ctx *c, const double *d_coor, EvalPlan *p) {
  BPar P;
  const uint32_t T = c->par.terms;
  P.fix_scale = c->fix_scale;
  P.ovf = c->d_ovf;
  P.epstmp14 = (T & MOIL_TERM_EL) ? 1.0 : 0.0;
  P.pick = (T & MOIL_TERM_VDW) ? 1.0 : 0.0;
  P.qfac = (T & MOIL_TERM_EL) ? KOFDIE / c->par.eps : 0.0;
  P.beta = c->beta;
  P.erftbdns = c->erftbdns;
  P.ntab = c->ntab;
  P.do_el14 = (T & MOIL_TERM_EL14) != 0;
  P.do_v14 = (T & MOIL_TERM_VDW14) != 0;
  P.arith = (T & MOIL_TERM_ARITH) ? 1 : 0;
  P.max_parts = c->max_parts;
  const int nblk = bonded_blocks(c);
  p->n_intra = 0;
  if (water_intra_on(c)) {
    const bool nbon = (T & (MOIL_TERM_VDW | MOIL_TERM_EL)) != 0;
    IntraArgs &A = p->intra;
    WIntra &W = A.W;
    W.nwat = c->nwat;
    W.first = c->d_wat_o;
    W.boff = (T & MOIL_TERM_BOND) ? c->d_wi_boff : c->d_wi_zero;
    W.aoff = (T & MOIL_TERM_ANGLE) ? c->d_wi_aoff : c->d_wi_zero;
    W.eoff = nbon ? c->d_wi_eoff : c->d_wi_zero;
    W.bidx = c->d_wi_bidx; W.aidx = c->d_wi_aidx; W.eidx = c->d_wi_eidx;
    W.bpar = c->d_wi_bpar; W.apar = c->d_wi_apar;
    W.q = c->d_ptchg; W.erf = c->d_erf;
    A.coor = d_coor;
    A.P = P;
    A.facc = c->d_facc;
    A.ep = c->d_epart;
    A.mp = c->max_parts;
    A.part0 = nblk;
    p->n_intra = water_intra_blocks(c);
  }
  BondedArgs &B = p->bonded;
  B.coor = d_coor;
  B.S = make_seg(c);
  B.P = P;
  B.bslot = c->d_bslot;
  B.ep = c->d_epart;
  B.mp = c->max_parts;
  p->n_bonded = nblk;
}

int bonded_max_blocks(const moil_ctx *c) { return max(1, bonded_blocks(c) + water_intra_blocks(c)); }
