// The molecular virial of MOIL's constant-pressure path (pressON): the
// terms MOIL's eforce adds up beside the energy when pressON is set,
//   virial, virXX..ZZ   LJ pairs: -df r_a R_a          (pot/cdie_ewald.f:440-460)
//   V_PIdirXX..ZZ       Coulomb pairs: V_PIdir r_a R_a (pot/cdie_ewald.f:351-368)
//   V_PIrecXX..ZZ       reciprocal sum                 (pot/ewald_receip.f:799-807)
//   V_PIrelXX..ZZ       reciprocal force . relative position
//                                                      (pot/ewald_receip.f:929-933)
// with R the separation of the two atoms' molecules' centres of mass
// (translated with the atom for an image pair: symm/symcdie_ewald.f:150-200,
// symm/symwat_ewald.f:160-203 write it as M_j - M_i - T with the signs
// folded in, the same quantity).
//
// moil_gpu_eforce_virial_dev runs one evaluation (the forces and energies
// of moil_gpu_eforce_dev) with the two reciprocal-space kernels added to
// its reciprocal chain, and the real-space terms in a pass of their own
// over the same work items (the force kernels stay as they are): one
// workgroup per chunk, FP64 pair arithmetic with MOIL's erfc table and
// V_PIdir's own exp, as MOIL evaluates them, whatever the context's
// precision.  Partial sums per chunk / block, summed in a fixed order:
// the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "kargs.h"
#include "pairmath.h"

#define VT 256
#define VIR_REAL 7  // virial, virXX, virYY, virZZ, V_PIdirXX, YY, ZZ

namespace {

int32_t fail_status(int32_t code, const std::string &msg) {
  set_error(msg);
  return code;
}

// centres of mass, one thread per molecule: the coordinate part of
// atom2molecule (analysis/pressure.f:419-505) -- per component the
// mass-weighted sum over the molecule's atoms in ordlist order, over the
// molecular mass -- written to every atom of the molecule
__global__ void k_mol_com(const double *__restrict__ coor, const int32_t *__restrict__ patom,
                          const int32_t *__restrict__ ordlist, const double *__restrict__ ptms,
                          const double *__restrict__ massmol, int nmol,
                          double *__restrict__ acom) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmol) return;
  const int k0 = m ? patom[m - 1] : 0, k1 = patom[m];
  double x[3];
  for (int d = 0; d < 3; d++) {
    double s = 0.0;
    for (int k = k0; k < k1; k++) {
      const int a = ordlist[k] - 1;
      s = s + coor[3 * a + d] * ptms[a];
    }
    x[d] = s / massmol[m];
  }
  for (int k = k0; k < k1; k++) {
    const int a = ordlist[k] - 1;
    acom[3 * a] = x[0];
    acom[3 * a + 1] = x[1];
    acom[3 * a + 2] = x[2];
  }
}

// one pair's terms: v[0] virial, v[1..3] virXX..ZZ (LJ part dfv), v[4..6]
// V_PIdirXX..ZZ (vpd); r = x_i - x_j, R the molecular separation
__device__ __forceinline__ void pair_terms(double v[VIR_REAL], const double r[3], const double R[3],
                                           double dfv, double vpd) {
#pragma clang fp contract(off)
  if (dfv != 0.0) {
    v[0] = v[0] - R[0] * dfv * r[0] - R[1] * dfv * r[1] - R[2] * dfv * r[2];
#pragma unroll
    for (int k = 0; k < 3; k++) v[1 + k] = v[1 + k] - dfv * r[k] * R[k];
  }
#pragma unroll
  for (int k = 0; k < 3; k++) v[4 + k] = v[4 + k] + vpd * r[k] * R[k];
}

// MOIL's V_PIdir of a Coulomb pair, qq (erfc s^3 + 2 beta/sqrt(pi)
// exp(-x^2) s^2), x = beta r, s = 1/r (pot/cdie_ewald.f:351-353)
__device__ __forceinline__ double v_pidir(double qq, double erfcc, double s, double x,
                                          double c2b) {
#pragma clang fp contract(off)
  return qq * (erfcc * s * s * s + c2b * exp(-(x * x)) * s * s);
}

// the chunk's seven sums into its column of vpart
__device__ __forceinline__ void store_chunk(double v[VIR_REAL], double *ws, double *vpart,
                                            int cap, int col) {
  const int t = threadIdx.x;
  double a = wg_sum4(v[0], v[1], v[2], v[3], ws);
  if (t < 4) vpart[(size_t)t * cap + col] = a;
  __syncthreads();
  a = wg_sum4(v[4], v[5], v[6], 0.0, ws);
  if (t < 3) vpart[(size_t)(4 + t) * cap + col] = a;
}

// water groups (pot/watwat_ewald.f, symm/symwat_ewald.f): per chunk of a
// 64-water block, one lane per item, every listed pair's nine site pairs
// with the two waters' molecular separation
__global__ __launch_bounds__(VT) void k_vir_water(WaterArgs W, const int32_t *__restrict__ slot_o,
                                                   const double *__restrict__ coor,
                                                   const double *__restrict__ acom,
                                                   double *__restrict__ vpart, int cap) {
#pragma clang fp contract(off)
  __shared__ double ipos[WBLOCK * 9], icom[WBLOCK * 3];
  __shared__ double ws[16];
  const int ch = blockIdx.x, t = threadIdx.x;
  const int blk = W.chunks[3 * ch], beg = W.chunks[3 * ch + 1], end = W.chunks[3 * ch + 2];
  const int s0 = blk * WBLOCK;
  for (int u = t; u < WBLOCK * 9; u += VT) {
    const int slot = s0 + u / 9;
    ipos[u] = slot < W.nwat ? coor[3 * (size_t)slot_o[slot] + u % 9] : 0.0;
  }
  for (int u = t; u < WBLOCK * 3; u += VT) {
    const int slot = s0 + u / 3;
    icom[u] = slot < W.nwat ? acom[3 * (size_t)slot_o[slot] + u % 3] : 0.0;
  }
  __syncthreads();
  const WaterRS &P = W.P;
  const double c2b = 2.0 * P.beta / sqrt(M_PI);
  double v[VIR_REAL] = {0, 0, 0, 0, 0, 0, 0};
  for (int it = beg + t; it < end; it += VT) {
    const WItem item = W.items[it];
    const int o = slot_o[item.j];
    double xj[9], jc[3];
#pragma unroll
    for (int k = 0; k < 9; k++) xj[k] = coor[3 * (size_t)o + k];
#pragma unroll
    for (int k = 0; k < 3; k++) jc[k] = acom[3 * (size_t)o + k];
    int ia, jb, kc;
    symop_of(item.s, &ia, &jb, &kc);
    const double T[3] = {ia * P.a, jb * P.b, kc * P.c};
    for (uint64_t mask = item.m1 | item.m2; mask; mask &= mask - 1) {
      const int i = __ffsll((long long)mask) - 1;
      const bool k2 = (item.m2 >> i) & 1ull;
      double xi[9];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int k = 0; k < 3; k++) xi[3 * a + k] = ipos[i * 9 + 3 * a + k] + T[k];
      const double ox = xi[0] - xj[0], oy = xi[1] - xj[1], oz = xi[2] - xj[2];
      if (k2 && ox * ox + oy * oy + oz * oz > P.cutele2) continue;
      double R[3];
#pragma unroll
      for (int k = 0; k < 3; k++) R[k] = (icom[i * 3 + k] + T[k]) - jc[k];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
          const double r[3] = {xi[3 * a] - xj[3 * b], xi[3 * a + 1] - xj[3 * b + 1],
                               xi[3 * a + 2] - xj[3 * b + 2]};
          const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          const double qq = (a == 0 && b == 0) ? P.qq_oo : ((a == 0 || b == 0) ? P.qq_oh : P.qq_hh);
          const double s2 = 1.0 / r2, rij = sqrt(r2), s = 1.0 / rij;
          double dfv = 0.0;
          if (a == 0 && b == 0 && !k2 && r2 <= P.cutvdw2) {
            const double s6 = s2 * s2 * s2;
            const double A12 = P.oo12 * s6 * s6, B6 = P.oo6 * s6;
            dfv = -6.0 * s2 * (A12 + (A12 - B6));
          }
          double erfcc, derfc;
          erfc_table(W.erf, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
          pair_terms(v, r, R, dfv, v_pidir(qq, erfcc, s, P.beta * rij, c2b));
        }
    }
  }
  store_chunk(v, ws, vpart, cap, ch);
}

// atom lists (pot/cdie_ewald.f lists 1-3, symm/symcdie_ewald.f): per
// chunk of a 64-atom block, one lane per item (the force kernel's pair
// rules: list1 LJ within rvmx and Coulomb, list2 LJ within rvmx, list3
// Coulomb within relx)
__global__ __launch_bounds__(VT) void k_vir_atoms(AtomArgs A, const double *__restrict__ acom,
                                                   double *__restrict__ vpart, int cap) {
#pragma clang fp contract(off)
  __shared__ double ix[WBLOCK][3], ic[WBLOCK][3];
  __shared__ double iq[WBLOCK], ia12[WBLOCK], ib6[WBLOCK];
  __shared__ double ws[16];
  const int ch = blockIdx.x, t = threadIdx.x;
  const int blk = A.chunks[3 * ch], beg = A.chunks[3 * ch + 1], end = A.chunks[3 * ch + 2];
  const int s0 = blk * WBLOCK;
  const ARS &P = A.P;
  if (t < WBLOCK) {
    const int a = (s0 + t < A.nsol) ? A.sslot2atom[s0 + t] : -1;
    for (int k = 0; k < 3; k++) {
      ix[t][k] = a >= 0 ? A.coor[3 * a + k] : 0.0;
      ic[t][k] = a >= 0 ? acom[3 * a + k] : 0.0;
    }
    iq[t] = a >= 0 ? A.q[a] * P.qfac : 0.0;
    ia12[t] = a >= 0 ? A.e12[a] * P.pick : 0.0;
    ib6[t] = a >= 0 ? (P.arith ? A.e6[a] : A.e6[a] * P.pick) : 0.0;
  }
  __syncthreads();
  const double c2b = 2.0 * P.beta / sqrt(M_PI);
  double v[VIR_REAL] = {0, 0, 0, 0, 0, 0, 0};
  for (int it = beg + t; it < end; it += VT) {
    const AItem item = A.items[it];
    const int ja = item.j;
    int s;
    bool rev;
    dir_of(item.s, s, rev);
    int sa, sb, sc;
    symop_of(s, &sa, &sb, &sc);
    const double T[3] = {sa * P.a, sb * P.b, sc * P.c};
    // the translated atom: the block atom (s > 0), or the partner (rev)
    const double ti = (s > 0 && !rev) ? 1.0 : 0.0, tj = rev ? 1.0 : 0.0;
    double xj[3], jc[3];
    for (int k = 0; k < 3; k++) {
      xj[k] = A.coor[3 * ja + k] + tj * T[k];
      jc[k] = acom[3 * ja + k] + tj * T[k];
    }
    const double qj = A.q[ja], aj = A.e12[ja], bj = A.e6[ja];
    const uint64_t m2 = item.m2, m3 = item.m3;
    for (uint64_t mask = item.m1 | m2 | m3; mask; mask &= mask - 1) {
      const int i = __ffsll((long long)mask) - 1;
      const int kind = ((m2 >> i) & 1ull) ? 2 : (((m3 >> i) & 1ull) ? 3 : 1);
      double r[3], R[3];
      for (int k = 0; k < 3; k++) {
        r[k] = (ix[i][k] + ti * T[k]) - xj[k];
        R[k] = (ic[i][k] + ti * T[k]) - jc[k];
      }
      const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      const double s2 = 1.0 / r2;
      const bool lj = (kind != 3) && r2 <= P.cutvdw2;
      const bool el = (kind == 1) || (kind == 3 && r2 <= P.cutele2);
      double dfv = 0.0, vpd = 0.0;
      // under the arith rule MOIL adds no LJ virial for a real list1 pair
      // (pot/cdie_ewald.f:432-473: the virial sits in the geometric branch)
      if (lj && !(P.arith && kind == 1 && s == 0)) {
        const double s6 = s2 * s2 * s2;
        double Ac, Bc;
        lj_coeffs(P.arith, ia12[i], ib6[i], aj, bj, Ac, Bc);
        const double A12 = Ac * s6 * s6, B6 = Bc * s6;
        dfv = -6.0 * s2 * (A12 + (A12 - B6));
      }
      if (el) {
        const double rij = sqrt(r2), sv = 1.0 / rij;
        double erfcc, derfc;
        erfc_table(A.erf, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
        vpd = v_pidir(iq[i] * qj, erfcc, sv, P.beta * rij, c2b);
      }
      if (lj || el) pair_terms(v, r, R, dfv, vpd);
    }
  }
  store_chunk(v, ws, vpart, cap, A.part0 + ch);
}

// V_PIrec (pot/ewald_receip.f:799-807) on the half spectrum after the
// forward transform: E(k) = 0.5 eterm |S(k)|^2 (times 2 for k1 strictly
// inside the half), times 1 - 2 (1 + fac m^2)/m^2 m_a^2
template <typename R>
__global__ __launch_bounds__(VT) void k_pme_vir(const R *__restrict__ S,
                                                 const double *__restrict__ eterm, int nc1, int n1,
                                                 int n2, int n3, double ra, double rb, double rc,
                                                 double fac, size_t H, double *__restrict__ vrec,
                                                 int nblk) {
#pragma clang fp contract(off)
  __shared__ double ws[16];
  const size_t g = (size_t)blockIdx.x * VT + threadIdx.x;
  double v[3] = {0, 0, 0};
  if (g > 0 && g < H) {
    const int k1 = (int)(g % nc1), k2 = (int)((g / nc1) % n2), k3 = (int)(g / ((size_t)nc1 * n2));
    const int h1 = n1 / 2 + (n1 & 1), h2 = n2 / 2 + (n2 & 1), h3 = n3 / 2 + (n3 & 1);
    const int m1 = k1 + 1 > h1 ? k1 - n1 : k1, m2 = k2 + 1 > h2 ? k2 - n2 : k2,
              m3 = k3 + 1 > h3 ? k3 - n3 : k3;
    const double mh[3] = {ra * m1, rb * m2, rc * m3};
    const double msq = mh[0] * mh[0] + mh[1] * mh[1] + mh[2] * mh[2];
    const double re = (double)S[2 * g], im = (double)S[2 * g + 1];
    const double w = (k1 == 0 || 2 * k1 == n1) ? 0.5 : 1.0;
    const double e = w * eterm[g] * (re * re + im * im);
    for (int a = 0; a < 3; a++) v[a] = e * (1.0 - 2.0 * (1.0 + fac * msq) / msq * mh[a] * mh[a]);
  }
  const double a = wg_sum4(v[0], v[1], v[2], 0.0, ws);
  if (threadIdx.x < 3) vrec[(size_t)threadIdx.x * nblk + blockIdx.x] = a;
}

// V_PIrel (pot/ewald_receip.f:929-933): the reciprocal force on every
// atom (minus the gather's gradient) times its position relative to its
// molecule's centre of mass
__global__ __launch_bounds__(VT) void k_vir_rel(const double *__restrict__ fpme,
                                                 const double *__restrict__ coor,
                                                 const double *__restrict__ acom, int npt,
                                                 double *__restrict__ vrel, int nblk) {
#pragma clang fp contract(off)
  __shared__ double ws[16];
  const int n = blockIdx.x * VT + threadIdx.x;
  double v[3] = {0, 0, 0};
  if (n < npt)
    for (int a = 0; a < 3; a++) v[a] = -fpme[3 * n + a] * (coor[3 * n + a] - acom[3 * n + a]);
  const double s = wg_sum4(v[0], v[1], v[2], 0.0, ws);
  if (threadIdx.x < 3) vrel[(size_t)threadIdx.x * nblk + blockIdx.x] = s;
}

// the thirteen totals: every row of partials summed in a fixed order (a
// strided sum per thread, then the workgroup's fixed tree)
__global__ __launch_bounds__(VT) void k_vir_sum(const double *__restrict__ vpart, int cap,
                                                 int nreal, const double *__restrict__ vrec,
                                                 int nrec, const double *__restrict__ vrel,
                                                 int nrel, double *__restrict__ out) {
  __shared__ double ws[16];
  for (int row = 0; row < MOIL_N_VIRIAL; row++) {
    const double *src;
    int n;
    if (row < VIR_REAL) {
      src = vpart + (size_t)row * cap;
      n = nreal;
    } else if (row < VIR_REAL + 3) {
      src = vrec + (size_t)(row - VIR_REAL) * nrec;
      n = nrec;
    } else {
      src = vrel + (size_t)(row - VIR_REAL - 3) * nrel;
      n = nrel;
    }
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += VT) s += src[k];
    const double tot = wg_sum4(s, 0.0, 0.0, 0.0, ws);
    if (threadIdx.x == 0) out[row] = tot;
    __syncthreads();
  }
}

int rec_blocks(const moil_ctx *c) {
  return (int)(((size_t)c->nc1 * c->nf2 * c->nf3 + VT - 1) / VT);
}
int rel_blocks(const moil_ctx *c) { return (c->npt + VT - 1) / VT; }

}  // namespace

void launch_pme_virial(const moil_ctx *c, const EvalPlan &p, hipStream_t st) {
  const double pi = 3.14159265358979323846;
  const double fac = pi * pi / (c->beta * c->beta);
  const size_t H = (size_t)c->nc1 * c->nf2 * c->nf3;
  const int nb = rec_blocks(c);
  if (p.f64)
    k_pme_vir<double><<<nb, VT, 0, st>>>((const double *)c->d_qcplx, c->d_eterm, c->nc1, c->nf1,
                                         c->nf2, c->nf3, 1.0 / c->par.a, 1.0 / c->par.b,
                                         1.0 / c->par.c, fac, H, p.vrec, nb);
  else
    k_pme_vir<float><<<nb, VT, 0, st>>>((const float *)c->d_qcplx, c->d_eterm, c->nc1, c->nf1,
                                        c->nf2, c->nf3, 1.0 / c->par.a, 1.0 / c->par.b,
                                        1.0 / c->par.c, fac, H, p.vrec, nb);
}

void launch_rel_virial(const moil_ctx *c, const EvalPlan &p, hipStream_t st) {
  const int nb = rel_blocks(c);
  k_vir_rel<<<nb, VT, 0, st>>>(c->d_fpme, p.gather.coor, p.acom, c->npt, p.vrel, nb);
}

void virial_free(moil_ctx *c) {
  void *ptrs[] = {c->d_patom, c->d_ordlist, c->d_ptms, c->d_massmol, c->d_acom, c->d_vpart,
                  c->d_vrec, c->d_vrel, c->d_virial};
  for (void *q : ptrs)
    if (q) hipFree(q);
  c->d_patom = c->d_ordlist = nullptr;
  c->d_ptms = c->d_massmol = c->d_acom = c->d_vpart = c->d_vrec = c->d_vrel = c->d_virial =
      nullptr;
  c->nmol = 0;
  c->vpart_cap = 0;
}

extern "C" int32_t moil_gpu_set_molecules(moil_ctx *c, int32_t nmol, const int32_t *patom,
                                          const int32_t *ordlist, const int32_t *pmol,
                                          const double *ptms) {
  if (!c) return fail_status(MOIL_GPU_EINVAL, "set_molecules: null context");
  if (nmol == 0) {
    HIP_CHECK(hipDeviceSynchronize());
    virial_free(c);
    return MOIL_GPU_OK;
  }
  const int npt = c->npt;
  if (nmol < 0 || nmol > npt || !patom || !ordlist || !pmol || !ptms)
    return fail_status(MOIL_GPU_EINVAL, "set_molecules: bad molecule count or null array");
  // every atom in exactly one molecule, ordlist and pmol agreeing, and the
  // molecular masses (compute_massmol, analysis/init_pressure.f:411-428)
  std::vector<char> seen(npt, 0);
  std::vector<double> massmol(nmol);
  int k0 = 0;
  for (int m = 0; m < nmol; m++) {
    if (patom[m] <= k0 || patom[m] > npt)
      return fail_status(MOIL_GPU_EINVAL, "set_molecules: patom not increasing within 1..npt");
    double sm = 0.0;
    for (int k = k0; k < patom[m]; k++) {
      const int a = ordlist[k];
      if (a < 1 || a > npt || seen[a - 1] || pmol[a - 1] != m + 1)
        return fail_status(MOIL_GPU_EINVAL, "set_molecules: atom " + std::to_string(a) +
                                                " listed twice or not in molecule " +
                                                std::to_string(m + 1) + " by pmol");
      seen[a - 1] = 1;
      sm = sm + ptms[a - 1];
    }
    if (!(sm > 0.0))
      return fail_status(MOIL_GPU_EINVAL, "set_molecules: molecule " + std::to_string(m + 1) +
                                              " has no positive mass");
    massmol[m] = sm;
    k0 = patom[m];
  }
  if (k0 != npt)
    return fail_status(MOIL_GPU_EINVAL, "set_molecules: " + std::to_string(npt - k0) +
                                            " atoms in no molecule");
  HIP_CHECK(hipDeviceSynchronize());
  virial_free(c);
  HIP_CHECK(hipMalloc(&c->d_patom, sizeof(int32_t) * nmol));
  HIP_CHECK(hipMalloc(&c->d_ordlist, sizeof(int32_t) * npt));
  HIP_CHECK(hipMalloc(&c->d_ptms, sizeof(double) * npt));
  HIP_CHECK(hipMalloc(&c->d_massmol, sizeof(double) * nmol));
  HIP_CHECK(hipMalloc(&c->d_acom, sizeof(double) * 3 * (size_t)npt));
  HIP_CHECK(hipMalloc(&c->d_vrec, sizeof(double) * 3 * (size_t)std::max(rec_blocks(c), 1)));
  HIP_CHECK(hipMalloc(&c->d_vrel, sizeof(double) * 3 * (size_t)std::max(rel_blocks(c), 1)));
  HIP_CHECK(hipMalloc(&c->d_virial, sizeof(double) * MOIL_N_VIRIAL));
  HIP_CHECK(hipMemcpy(c->d_patom, patom, sizeof(int32_t) * nmol, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_ordlist, ordlist, sizeof(int32_t) * npt, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_ptms, ptms, sizeof(double) * npt, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_massmol, massmol.data(), sizeof(double) * nmol, hipMemcpyHostToDevice));
  c->nmol = nmol;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_eforce_virial_dev(moil_ctx *c, const double *d_coor, double *d_dpot,
                                              double *d_energies, double *d_virial,
                                              void *stream) {
  if (!c || !d_coor || !d_dpot || !d_virial)
    return fail_status(MOIL_GPU_EINVAL, "eforce_virial: null argument");
  if (!c->have_lists)
    return fail_status(MOIL_GPU_ESTATE, "eforce_virial called before build_lists");
  if (c->nmol == 0)
    return fail_status(MOIL_GPU_ESTATE, "eforce_virial needs moil_gpu_set_molecules");
  hipStream_t st = pick_stream(c, stream);
  const int nreal = c->nwchunk + c->nachunk;
  if (nreal > c->vpart_cap || !c->d_vpart) {
    HIP_CHECK(hipStreamSynchronize(st));
    if (c->d_vpart) HIP_CHECK(hipFree(c->d_vpart));
    c->d_vpart = nullptr;
    c->vpart_cap = std::max(nreal + nreal / 4, 64);
    HIP_CHECK(hipMalloc(&c->d_vpart, sizeof(double) * VIR_REAL * (size_t)c->vpart_cap));
  }
  if (c->facc_dirty) {
    HIP_CHECK(hipMemsetAsync(c->d_facc, 0, sizeof(long long) * 3 * (size_t)c->npt, st));
    c->facc_dirty = false;
  }
  // centres of mass first: the fork inside the evaluation orders the
  // reciprocal chain after them
  k_mol_com<<<(c->nmol + VT - 1) / VT, VT, 0, st>>>(d_coor, c->d_patom, c->d_ordlist, c->d_ptms,
                                                     c->d_massmol, c->nmol, c->d_acom);
  EvalPlan p;
  eval_plan(c, d_coor, d_dpot, d_energies, &p);
  p.vir = true;
  p.acom = c->d_acom;
  p.vrec = c->d_vrec;
  p.vrel = c->d_vrel;
  c->facc_dirty = true;
  const int32_t rc = eforce_eager(c, p, st);
  if (rc != MOIL_GPU_OK) return rc;
  c->facc_dirty = false;
  if (p.n_water > 0)
    k_vir_water<<<p.n_water, VT, 0, st>>>(p.water, p.pack.slot_o, d_coor, c->d_acom, c->d_vpart,
                                           c->vpart_cap);
  if (p.n_atoms > 0)
    k_vir_atoms<<<p.n_atoms, VT, 0, st>>>(p.atoms, c->d_acom, c->d_vpart, c->vpart_cap);
  const int nrec = p.pme ? rec_blocks(c) : 0, nrel = p.pme ? rel_blocks(c) : 0;
  k_vir_sum<<<1, VT, 0, st>>>(c->d_vpart, c->vpart_cap, p.n_water + p.n_atoms, c->d_vrec, nrec,
                              c->d_vrel, nrel, d_virial);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_eforce_virial(moil_ctx *c, const double *coor, double *dpot,
                                          moil_energies *e, moil_virial *v) {
  if (!c || !coor || !dpot || !v) return fail_status(MOIL_GPU_EINVAL, "null argument");
  if (c->nmol == 0)
    return fail_status(MOIL_GPU_ESTATE, "eforce_virial needs moil_gpu_set_molecules");
  const size_t n3 = 3 * (size_t)c->npt;
  HIP_CHECK(hipMemcpyAsync(c->d_coor, coor, sizeof(double) * n3, hipMemcpyHostToDevice,
                           c->stream));
  const int32_t rc =
      moil_gpu_eforce_virial_dev(c, c->d_coor, c->d_dpot, c->d_energy, c->d_virial, c->stream);
  if (rc != MOIL_GPU_OK) return rc;
  double en[MOIL_N_ENERGIES];
  HIP_CHECK(hipMemcpyAsync(dpot, c->d_dpot, sizeof(double) * n3, hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipMemcpyAsync(en, c->d_energy, sizeof(en), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(v, c->d_virial, sizeof(moil_virial), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (e) std::memcpy(e, en, sizeof(en));
  if (std::isnan(en[0]))
    return fail_status(MOIL_GPU_ECAPACITY,
                       "eforce_virial: a force term beyond the fixed-point range; energies are "
                       "NaN and the gradient is not valid");
  return MOIL_GPU_OK;
}
