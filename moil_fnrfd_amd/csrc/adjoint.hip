// Parameter gradients of the energy: dE/dq, dE/d(epsgm12) and
// dE/d(epsgm6) per atom (MOIL's ptchg, epsgm12, epsgm6 of CONNECT.BLOCK),
// for force-field fitting.  MOIL has no counterpart; the energy being
// differentiated is exactly the one moil_gpu_eforce evaluates on the same
// lists:
//   real space (watwat_ewald / symwat_ewald, cdie_ewald / symcdie_ewald):
//     E = sum qfac q_i q_j erfc(beta r)/r + sum (A_ij/r^12 - B_ij/r^6),
//     A_ij = pick e12_i e12_j, B_ij = pick e6_i e6_j, with MOIL's list
//     rules for which pairs carry which term;
//   exclusion correction  -qfac q_i q_j erf(beta r)/r;
//   scaled 1-4 pairs (ener14) with p14 = (v14 e12_i e12_j, v14 e6_i e6_j,
//     el14 eps kofdie q_i q_j), the LJ part zero where an atom has
//     epsgm12 < 1e-3 (generic/wconn.f:296-325);
//   self energy -beta/sqrt(pi) qfac sum q_i^2;
//   reciprocal sum: dE/dq_i = fc sum_p theta_i(p) psi(p), psi the
//     convolved grid (pme.hip).
// Water pairs: MOIL evaluates every water pair with the parameters of the
// last water (pot/watwat_ewald.f:40-72); the gradient here is that of the
// same energy written with each water's own parameters (equal to the last
// water's), i.e. every water atom gets its own share.
//
// The kernels walk the same work items as the force kernels and recompute
// every pair distance (nothing is stored by the forward pass).  Pair
// arithmetic follows the force kernels (FP32 with the fitted erfc in mixed
// mode, FP64 with MOIL's erfc table in FP64 mode); each item's sums are
// added to the per-atom FP64 outputs with device atomics (block members
// through LDS first).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "kargs.h"
#include "pairmath.h"

#define ADJ_THREADS 256

struct AdjWater {
  double qO, qH;            // charges of the last water
  double qfac;              // kofdie/eps (0 without electrostatics)
  double e12O, e6O, pick;   // LJ factors of the last water's O, 1 or 0
  double beta, cutvdw2, cutele2, erftbdns;
  int ntab;
  double a, b, c;
};

// ---------------------------------------------------------------------------
// water groups: per chunk of a 64-water block, one lane per item
template <bool F64>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_water(
    const int32_t *__restrict__ chunks, const WItem *__restrict__ items,
    const void *__restrict__ wpos_, const int32_t *__restrict__ slot_o, int nwat, AdjWater P,
    const double *__restrict__ erf_arr, double *__restrict__ dq, double *__restrict__ de12,
    double *__restrict__ de6) {
  __shared__ float4 iA[WBLOCK], iB[WBLOCK], iC[WBLOCK];  // mixed mode positions
  __shared__ double ipos[F64 ? WBLOCK * 9 : 1];          // FP64 positions
  __shared__ double acc[WBLOCK * 5];                     // dq O H1 H2, de12 O, de6 O
  const int ch = blockIdx.x, t = threadIdx.x;
  const int blk = chunks[3 * ch], beg = chunks[3 * ch + 1], end = chunks[3 * ch + 2];
  const int s0 = blk * WBLOCK;
  if (F64) {
    const double *wpos = (const double *)wpos_;
    for (int u = t; u < WBLOCK * 9; u += ADJ_THREADS)
      ipos[u] = (s0 + u / 9 < nwat) ? wpos[9 * (size_t)s0 + u] : 0.0;
  } else if (t < WBLOCK && s0 + t < nwat) {
    const float4 *wpos = (const float4 *)wpos_;
    iA[t] = wpos[3 * (s0 + t)];
    iB[t] = wpos[3 * (s0 + t) + 1];
    iC[t] = wpos[3 * (s0 + t) + 2];
  }
  for (int u = t; u < WBLOCK * 5; u += ADJ_THREADS) acc[u] = 0.0;
  __syncthreads();
  const double qs[3] = {P.qO, P.qH, P.qH};
  // FP32 constants as in the force kernel: x' = beta sqrt(log2 e) r
  const double sl2e = sqrt(1.4426950408889634);
  const float betap = (float)(P.beta * sl2e), pe = (float)(ERFC_P / sl2e);
  const float cutvdw2f = (float)P.cutvdw2, cutele2f = (float)P.cutele2;
  for (int it = beg + t; it < end; it += ADJ_THREADS) {
    // the mask walk starts at a member given by the item, so the order of
    // an item's sums does not depend on the lane or chunk that runs it
    const int rot = it & 63;
    const WItem item = items[it];
    const int js = item.j;
    int ia, jb, kc;
    symop_of(item.s, &ia, &jb, &kc);
    double pj[3] = {0, 0, 0}, lj12 = 0, lj6 = 0;  // partner sums
    uint64_t mask = rotr64(item.m1 | item.m2, rot);
    if (F64) {
      const double *wpos = (const double *)wpos_;
      double xj[9];
#pragma unroll
      for (int k = 0; k < 9; k++) xj[k] = wpos[9 * (size_t)js + k];
      const double T[3] = {ia * P.a, jb * P.b, kc * P.c};
      while (mask) {
        const int i = (__ffsll((long long)mask) - 1 + rot) & 63;
        const bool k2 = (item.m2 >> i) & 1ull;
        mask &= mask - 1;
        double xi[9];
#pragma unroll
        for (int sa = 0; sa < 3; sa++)
#pragma unroll
          for (int k = 0; k < 3; k++) xi[3 * sa + k] = ipos[i * 9 + 3 * sa + k] + T[k];
        const double ox = xi[0] - xj[0], oy = xi[1] - xj[1], oz = xi[2] - xj[2];
        const double roo2 = ox * ox + oy * oy + oz * oz;
        if (k2 && roo2 > P.cutele2) continue;
        double pi[3] = {0, 0, 0};
#pragma unroll
        for (int sa = 0; sa < 3; sa++)
#pragma unroll
          for (int sb = 0; sb < 3; sb++) {
            const double rx = xi[3 * sa] - xj[3 * sb], ry = xi[3 * sa + 1] - xj[3 * sb + 1],
                         rz = xi[3 * sa + 2] - xj[3 * sb + 2];
            const double r2 = rx * rx + ry * ry + rz * rz;
            const double rij = sqrt(r2);
            double erfcc, derfc;
            erfc_table(erf_arr, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
            const double phi = P.qfac * erfcc / rij;
            pi[sa] += qs[sb] * phi;
            pj[sb] += qs[sa] * phi;
          }
        double l12 = 0, l6 = 0;
        if (!k2 && roo2 <= P.cutvdw2) {
          const double s2 = 1.0 / roo2, s6 = s2 * s2 * s2;
          l12 = P.pick * s6 * s6;
          l6 = -P.pick * s6;
        }
        lj12 += l12;
        lj6 += l6;
#pragma unroll
        for (int k = 0; k < 3; k++) atomicAdd(&acc[i * 5 + k], pi[k]);
        if (l12 != 0.0) {
          atomicAdd(&acc[i * 5 + 3], l12 * P.e12O);
          atomicAdd(&acc[i * 5 + 4], l6 * P.e6O);
        }
      }
    } else {
      const float4 *wpos = (const float4 *)wpos_;
      const float4 jA = wpos[3 * js], jB = wpos[3 * js + 1], jC = wpos[3 * js + 2];
      const double ojx = (double)jA.x + (double)jB.x - ia * P.a;
      const double ojy = (double)jA.y + (double)jB.y - jb * P.b;
      const double ojz = (double)jA.z + (double)jB.z - kc * P.c;
      const float jhx = (float)ojx, jhy = (float)ojy, jhz = (float)ojz;
      const float jlx = (float)(ojx - (double)jhx), jly = (float)(ojy - (double)jhy),
                  jlz = (float)(ojz - (double)jhz);
      const float vx[3] = {0.f, jA.w, jC.y}, vy[3] = {0.f, jB.w, jC.z}, vz[3] = {0.f, jC.x, jC.w};
      const float qsf[3] = {(float)(P.qfac * P.qO), (float)(P.qfac * P.qH),
                            (float)(P.qfac * P.qH)};
      float pjf[3] = {0.f, 0.f, 0.f};
      while (mask) {
        const int i = (__ffsll((long long)mask) - 1 + rot) & 63;
        const bool k2 = (item.m2 >> i) & 1ull;
        mask &= mask - 1;
        const float4 A = iA[i], B = iB[i], C = iC[i];
        const float dx = (A.x - jhx) + (B.x - jlx);
        const float dy = (A.y - jhy) + (B.y - jly);
        const float dz = (A.z - jhz) + (B.z - jlz);
        const float roo2 = dx * dx + dy * dy + dz * dz;
        if (k2 && roo2 > cutele2f) continue;
        const float ux[3] = {dx, dx + A.w, dx + C.y}, uy[3] = {dy, dy + B.w, dy + C.z},
                    uz[3] = {dz, dz + C.x, dz + C.w};
        float pif[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int sa = 0; sa < 3; sa++)
#pragma unroll
          for (int sb = 0; sb < 3; sb++) {
            const float rx = ux[sa] - vx[sb], ry = uy[sa] - vy[sb], rz = uz[sa] - vz[sb];
            const float r2 = (sa == 0 && sb == 0) ? roo2 : (rx * rx + ry * ry + rz * rz);
            const float s = __builtin_amdgcn_rsqf(r2);
            const float xp = betap * (r2 * s);
            const float tt = __builtin_amdgcn_rcpf(fmaf(pe, xp, 1.0f));
            const float ex = __builtin_amdgcn_exp2f(-(xp * xp));
            const float phi = ex * erfc_poly(tt) * s;
            pif[sa] = fmaf(qsf[sb], phi, pif[sa]);
            pjf[sb] = fmaf(qsf[sa], phi, pjf[sb]);
          }
        float l12 = 0.f, l6 = 0.f;
        if (!k2 && roo2 <= cutvdw2f) {
          const float s2 = __builtin_amdgcn_rcpf(roo2), s6 = s2 * s2 * s2;
          l12 = (float)P.pick * s6 * s6;
          l6 = -(float)P.pick * s6;
        }
        lj12 += (double)l12;
        lj6 += (double)l6;
#pragma unroll
        for (int k = 0; k < 3; k++) atomicAdd(&acc[i * 5 + k], (double)pif[k]);
        if (l12 != 0.f) {
          atomicAdd(&acc[i * 5 + 3], (double)l12 * P.e12O);
          atomicAdd(&acc[i * 5 + 4], (double)l6 * P.e6O);
        }
      }
#pragma unroll
      for (int k = 0; k < 3; k++) pj[k] = (double)pjf[k];
    }
    const int o = slot_o[js];
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (pj[k] != 0.0) atomicAdd(&dq[o + k], pj[k]);
    if (lj12 != 0.0) {
      atomicAdd(&de12[o], lj12 * P.e12O);
      atomicAdd(&de6[o], lj6 * P.e6O);
    }
  }
  __syncthreads();
  for (int u = t; u < WBLOCK * 5; u += ADJ_THREADS) {
    const int w = u / 5, k = u - 5 * w;
    if (s0 + w >= nwat || acc[u] == 0.0) continue;
    const int o = slot_o[s0 + w];
    if (k < 3) atomicAdd(&dq[o + k], acc[u]);
    else if (k == 3) atomicAdd(&de12[o], acc[u]);
    else atomicAdd(&de6[o], acc[u]);
  }
}

// ---------------------------------------------------------------------------
// atom lists: per chunk of a 64-atom solute block, one lane per item
struct AdjAtom {
  double qfac, pick;
  double beta, cutvdw2, cutele2, erftbdns;
  int ntab;
  double a, b, c;
};

__device__ inline void dir_split(int d, int &s, bool &rev) {
  rev = d > 13;
  s = rev ? d - 13 : d;
}

template <bool F64>
__global__ __launch_bounds__(ADJ_THREADS) void k_adj_atoms(
    const int32_t *__restrict__ chunks, const AItem *__restrict__ items,
    const double *__restrict__ coor, const int32_t *__restrict__ sslot2atom, int nsol,
    const double *__restrict__ q, const double *__restrict__ e12, const double *__restrict__ e6,
    const double *__restrict__ erf_arr, AdjAtom P, double *__restrict__ dq,
    double *__restrict__ de12, double *__restrict__ de6) {
  __shared__ double ix[WBLOCK][3];
  __shared__ double iq[WBLOCK], ia12[WBLOCK], ib6[WBLOCK];
  __shared__ double acc[WBLOCK * 3];  // dq, de12, de6
  const int ch = blockIdx.x, t = threadIdx.x;
  const int blk = chunks[3 * ch], beg = chunks[3 * ch + 1], end = chunks[3 * ch + 2];
  const int s0 = blk * WBLOCK;
  if (t < WBLOCK) {
    const int a = (s0 + t < nsol) ? sslot2atom[s0 + t] : -1;
    if (a >= 0) {
      ix[t][0] = coor[3 * a];
      ix[t][1] = coor[3 * a + 1];
      ix[t][2] = coor[3 * a + 2];
      iq[t] = q[a];
      ia12[t] = e12[a];
      ib6[t] = e6[a];
    } else {
      ix[t][0] = ix[t][1] = ix[t][2] = 0.0;
      iq[t] = ia12[t] = ib6[t] = 0.0;
    }
  }
  for (int u = t; u < WBLOCK * 3; u += ADJ_THREADS) acc[u] = 0.0;
  __syncthreads();
  const float betaf = (float)P.beta;
  const float cutvdw2f = (float)P.cutvdw2, cutele2f = (float)P.cutele2;
  const float log2e = 1.4426950408889634f;
  for (int it = beg + t; it < end; it += ADJ_THREADS) {
    // the mask walk starts at a member given by the item, so the order of
    // an item's sums does not depend on the lane or chunk that runs it
    const int rot = it & 63;
    const AItem item = items[it];
    const int ja = item.j;
    int s;
    bool rev;
    dir_split(item.s, s, rev);
    int sa, sb, sc;
    symop_of(s, &sa, &sb, &sc);
    const double T[3] = {sa * P.a, sb * P.b, sc * P.c};
    double xj[3] = {coor[3 * ja], coor[3 * ja + 1], coor[3 * ja + 2]};
    if (rev)
      for (int k = 0; k < 3; k++) xj[k] = xj[k] + T[k];
    const double qj = q[ja], aj = e12[ja], bj = e6[ja];
    double pj = 0, lj12 = 0, lj6 = 0;
    const uint64_t m2 = item.m2, m3 = item.m3;
    uint64_t mask = rotr64(item.m1 | m2 | m3, rot);
    while (mask) {
      const int i = (__ffsll((long long)mask) - 1 + rot) & 63;
      mask &= mask - 1;
      const int kind = ((m2 >> i) & 1ull) ? 2 : (((m3 >> i) & 1ull) ? 3 : 1);
      double d[3];
      if (s > 0 && !rev) {
#pragma clang fp contract(off)
        for (int k = 0; k < 3; k++) d[k] = (ix[i][k] + T[k]) - xj[k];
      } else {
        for (int k = 0; k < 3; k++) d[k] = ix[i][k] - xj[k];
      }
      double phi = 0.0, s12 = 0.0, s6v = 0.0;  // erfc/r, 1/r^12, 1/r^6 where active
      if (F64) {
#pragma clang fp contract(off)
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double s2 = 1.0 / r2;
        const bool lj = (kind != 3) && r2 <= P.cutvdw2;
        const bool el = (kind == 1) || (kind == 3 && r2 <= P.cutele2);
        if (el) {
          const double rij = sqrt(r2);
          double erfcc, derfc;
          erfc_table(erf_arr, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
          phi = erfcc / rij;
        }
        if (lj) {
          const double s6 = s2 * s2 * s2;
          s12 = s6 * s6;
          s6v = s6;
        }
      } else {
        const float dx = (float)d[0], dy = (float)d[1], dz = (float)d[2];
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float sv = __builtin_amdgcn_rsqf(r2);
        const float s2 = sv * sv;
        const bool lj = (kind != 3) && r2 <= cutvdw2f;
        const bool el = (kind == 1) || (kind == 3 && r2 <= cutele2f);
        if (el) {
          const float x = betaf * (r2 * sv);
          const float tt = __builtin_amdgcn_rcpf(fmaf(ERFC_P, x, 1.0f));
          const float ex = __builtin_amdgcn_exp2f(-(x * x) * log2e);
          phi = (double)(ex * erfc_poly(tt) * sv);
        }
        if (lj) {
          const float s6 = s2 * s2 * s2;
          s12 = (double)(s6 * s6);
          s6v = (double)s6;
        }
      }
      if (phi != 0.0) {
        atomicAdd(&acc[3 * i], P.qfac * qj * phi);
        pj += P.qfac * iq[i] * phi;
      }
      if (s12 != 0.0) {
        atomicAdd(&acc[3 * i + 1], P.pick * aj * s12);
        atomicAdd(&acc[3 * i + 2], -P.pick * bj * s6v);
        lj12 += P.pick * ia12[i] * s12;
        lj6 -= P.pick * ib6[i] * s6v;
      }
    }
    if (pj != 0.0) atomicAdd(&dq[ja], pj);
    if (lj12 != 0.0) atomicAdd(&de12[ja], lj12);
    if (lj6 != 0.0) atomicAdd(&de6[ja], lj6);
  }
  __syncthreads();
  for (int u = t; u < WBLOCK * 3; u += ADJ_THREADS) {
    const int w = u / 3, k = u - 3 * w;
    if (s0 + w >= nsol || acc[u] == 0.0) continue;
    const int a = sslot2atom[s0 + w];
    if (a < 0) continue;
    double *out = k == 0 ? dq : (k == 1 ? de12 : de6);
    atomicAdd(&out[a], acc[u]);
  }
}

// ---------------------------------------------------------------------------
// flat pass over [exclusion pairs | 1-4 pairs | atoms (self energy)]
struct AdjTerms {
  int nx, n14, npt;
  const int32_t *xpair;   // [2 nx] 0-based
  const int32_t *sp;      // [2 n14] 1-based
  const double *q, *e12, *e6, *erf;
  double qfac_excl;       // kofdie/eps with electrostatics on, else 0
  double beta, erftbdns;
  int ntab;
  double v14, k14;        // 1-4 scalings: v14, el14 eps kofdie
  double pick, epst14;    // LJ on (1/0), electrostatics on (1/0)
  int do_v14, do_el14;
  double selfc;           // -2 beta/sqrt(pi) kofdie/eps (0 without nonbonded terms)
};

__global__ __launch_bounds__(ADJ_THREADS) void k_adj_terms(const double *__restrict__ coor,
                                                          AdjTerms P, double *__restrict__ dq,
                                                          double *__restrict__ de12,
                                                          double *__restrict__ de6) {
  const size_t m = (size_t)blockIdx.x * ADJ_THREADS + threadIdx.x;
  if (m < (size_t)P.nx) {
    const int i = P.xpair[2 * m], j = P.xpair[2 * m + 1];
    const double rx = coor[3 * i] - coor[3 * j], ry = coor[3 * i + 1] - coor[3 * j + 1],
                 rz = coor[3 * i + 2] - coor[3 * j + 2];
    const double rij = sqrt(rx * rx + ry * ry + rz * rz);
    double erfcc, derfc;
    erfc_table(P.erf, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
    const double phi = -P.qfac_excl * (1.0 - erfcc) / rij;
    atomicAdd(&dq[i], phi * P.q[j]);
    atomicAdd(&dq[j], phi * P.q[i]);
    return;
  }
  if (m < (size_t)P.nx + P.n14) {
    const size_t n = m - P.nx;
    const int j = P.sp[2 * n] - 1, k = P.sp[2 * n + 1] - 1;
    const double rx = coor[3 * j] - coor[3 * k], ry = coor[3 * j + 1] - coor[3 * k + 1],
                 rz = coor[3 * j + 2] - coor[3 * k + 2];
    const double r2 = rx * rx + ry * ry + rz * rz, s2 = 1.0 / r2;
    if (P.do_v14 && P.e12[j] >= 1e-3 && P.e12[k] >= 1e-3 && P.v14 != 0.0) {
      const double s6 = s2 * s2 * s2, w12 = P.pick * P.v14 * s6 * s6, w6 = -P.pick * P.v14 * s6;
      atomicAdd(&de12[j], w12 * P.e12[k]);
      atomicAdd(&de12[k], w12 * P.e12[j]);
      atomicAdd(&de6[j], w6 * P.e6[k]);
      atomicAdd(&de6[k], w6 * P.e6[j]);
    }
    if (P.do_el14) {
      const double w = P.k14 * P.epst14 * sqrt(s2);
      atomicAdd(&dq[j], w * P.q[k]);
      atomicAdd(&dq[k], w * P.q[j]);
    }
    return;
  }
  if (m < (size_t)P.nx + P.n14 + P.npt) {
    const size_t a = m - P.nx - P.n14;
    atomicAdd(&dq[a], P.selfc * P.q[a]);
  }
}

__global__ void k_zero3(double *__restrict__ a, double *__restrict__ b, double *__restrict__ c,
                        size_t n) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) a[g] = b[g] = c[g] = 0.0;
}

// ---------------------------------------------------------------------------
int32_t pme_potential(moil_ctx *c, const EvalPlan &p, double *d_dq, hipStream_t st);

int32_t adjoint_run(moil_ctx *c, const double *d_coor, double *d_dq, double *d_de12,
                    double *d_de6, hipStream_t st) {
  const uint32_t T = c->par.terms;
  const bool nb = (T & (MOIL_TERM_VDW | MOIL_TERM_EL)) != 0;
  const bool f64 = c->par.precision != MOIL_PREC_MIXED;
  const size_t n = (size_t)c->npt;
  k_zero3<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(d_dq, d_de12, d_de6, n);
  EvalPlan p;
  if (nb) {
    pme_plan(c, d_coor, &p);
    p.f64 = f64;
    // the reciprocal potential needs the convolved grid of these
    // coordinates and charges: the spreading and transforms run again
    int32_t rc = pme_potential(c, p, d_dq, st);
    if (rc != MOIL_GPU_OK) return rc;
    water_plan(c, d_coor, &p);
    if (p.n_pack > 0) {
      if (f64) launch_k_pack_water_f64(p.pack, p.n_pack, st);
      else launch_k_pack_water_f32(p.pack, p.n_pack, st);
    }
    const double epst = (T & MOIL_TERM_EL) ? 1.0 : 0.0;
    const double pick = (T & MOIL_TERM_VDW) ? 1.0 : 0.0;
    const double qfac = epst * KOFDIE / c->par.eps;
    if (p.n_water > 0) {
      AdjWater W;
      const int io = c->h_wat_o[c->nwat - 1];
      W.qO = c->h_q[io];
      W.qH = c->h_q[io + 2];
      W.qfac = qfac;
      W.e12O = c->h_e12[io];
      W.e6O = c->h_e6[io];
      W.pick = pick;
      W.beta = c->beta;
      W.cutvdw2 = c->par.cutvdw2;
      W.cutele2 = c->par.cutele2;
      W.erftbdns = c->erftbdns;
      W.ntab = c->ntab;
      W.a = c->par.a; W.b = c->par.b; W.c = c->par.c;
      if (f64)
        k_adj_water<true><<<p.n_water, ADJ_THREADS, 0, st>>>(
            c->d_wchunk, c->d_witems, c->d_wpos, c->d_slot_o, c->nwat, W, c->d_erf, d_dq, d_de12,
            d_de6);
      else
        k_adj_water<false><<<p.n_water, ADJ_THREADS, 0, st>>>(
            c->d_wchunk, c->d_witems, c->d_wpos, c->d_slot_o, c->nwat, W, c->d_erf, d_dq, d_de12,
            d_de6);
    }
    if (c->nachunk > 0) {
      AdjAtom A;
      A.qfac = qfac;
      A.pick = pick;
      A.beta = c->beta;
      A.cutvdw2 = c->par.cutvdw2;
      A.cutele2 = c->par.cutele2;
      A.erftbdns = c->erftbdns;
      A.ntab = c->ntab;
      A.a = c->par.a; A.b = c->par.b; A.c = c->par.c;
      if (f64)
        k_adj_atoms<true><<<c->nachunk, ADJ_THREADS, 0, st>>>(
            c->d_achunk, c->d_aitems, d_coor, c->d_sslot2atom, c->nsslot, c->d_ptchg, c->d_eps12,
            c->d_eps6, c->d_erf, A, d_dq, d_de12, d_de6);
      else
        k_adj_atoms<false><<<c->nachunk, ADJ_THREADS, 0, st>>>(
            c->d_achunk, c->d_aitems, d_coor, c->d_sslot2atom, c->nsslot, c->d_ptchg, c->d_eps12,
            c->d_eps6, c->d_erf, A, d_dq, d_de12, d_de6);
    }
  }
  AdjTerms P;
  P.nx = nb ? c->nxall : 0;
  P.n14 = (nb && (T & (MOIL_TERM_EL14 | MOIL_TERM_VDW14))) ? c->totspe : 0;
  P.npt = c->npt;
  P.xpair = c->d_xall;
  P.sp = c->d_sp;
  P.q = c->d_ptchg;
  P.e12 = c->d_eps12;
  P.e6 = c->d_eps6;
  P.erf = c->d_erf;
  P.qfac_excl = (T & MOIL_TERM_EL) ? KOFDIE / c->par.eps : 0.0;
  P.beta = c->beta;
  P.erftbdns = c->erftbdns;
  P.ntab = c->ntab;
  P.v14 = c->v14;
  P.k14 = c->el14 * c->par.eps * KOFDIE;
  P.pick = (T & MOIL_TERM_VDW) ? 1.0 : 0.0;
  P.epst14 = (T & MOIL_TERM_EL) ? 1.0 : 0.0;
  P.do_v14 = (T & MOIL_TERM_VDW14) != 0;
  P.do_el14 = (T & MOIL_TERM_EL14) != 0;
  P.selfc = nb ? -2.0 * c->beta / sqrt(M_PI) * KOFDIE / c->par.eps : 0.0;
  const size_t tot = (size_t)P.nx + P.n14 + P.npt;
  k_adj_terms<<<(unsigned)((tot + ADJ_THREADS - 1) / ADJ_THREADS), ADJ_THREADS, 0, st>>>(
      d_coor, P, d_dq, d_de12, d_de6);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}
