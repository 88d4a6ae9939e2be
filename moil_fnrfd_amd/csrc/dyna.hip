// Resident molecular dynamics: MOIL's velocity step loop
// (dynamics/dyna.f:105-301) with the matrix SHAKE of rigid waters
// (dynamics/mshakpt.f, dynamics/mshakvl.f) on coordinates and velocities
// that stay on the GPU between force evaluations.
//
// Every nlist steps the coordinates are squeezed, the lists rebuilt and
// the forces evaluated once more on the new lists (dyna.f:129-144); each
// step then makes the position step velo*dt - factor1*dpot, corrects it
// with the waters' matrix SHAKE, moves the atoms, turns the step back into
// a velocity, evaluates the forces, adds -factor2*dpot and corrects the
// velocities (dyna.f:199-315), with factor1 = dt^2/2 / m and factor2 =
// dt/2 / m (dynamics/init_dyna.f:279-288).  The per-atom and per-water
// kernels run in FP64 in MOIL's operation order without contraction, so a
// step given the same forces is MOIL's step to the bit.  Temperature
// control, frozen particles and bond SHAKE are not part of it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "multi.h"

void set_error(const std::string &msg);

struct moil_dyna {
  moil_ctx *ctx = nullptr;
  moil_dyna_params par;
  double *d_f1 = nullptr, *d_f2 = nullptr;  // [npt] dt^2/2 / m, dt/2 / m
  int32_t *d_status = nullptr;              // [2] SHAKE failures (position, velocity)
  double *d_en = nullptr;                   // energies of the list-update evaluations
};

#define DY_THREADS 256

// velo <- velo dt - factor1 dpot (dyna.f:199-205)
__global__ void k_dyna_posstep(double *__restrict__ velo, const double *__restrict__ dpot,
                               const double *__restrict__ f1, double dt, int npt) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * DY_THREADS + threadIdx.x;
  if (i >= npt) return;
  const double f = f1[i];
#pragma unroll
  for (int k = 0; k < 3; k++) velo[3 * i + k] = velo[3 * i + k] * dt - f * dpot[3 * i + k];
}

// coor <- coor + velo, velo <- velo / dt (dyna.f:260-269)
__global__ void k_dyna_move(double *__restrict__ coor, double *__restrict__ velo, double rdt,
                            int npt) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * DY_THREADS + threadIdx.x;
  if (i >= npt) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    coor[3 * i + k] = coor[3 * i + k] + velo[3 * i + k];
    velo[3 * i + k] = velo[3 * i + k] * rdt;
  }
}

// velo <- velo - factor2 dpot (dyna.f:297-302)
__global__ void k_dyna_kick(double *__restrict__ velo, const double *__restrict__ dpot,
                            const double *__restrict__ f2, int npt) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * DY_THREADS + threadIdx.x;
  if (i >= npt) return;
  const double f = f2[i];
#pragma unroll
  for (int k = 0; k < 3; k++) velo[3 * i + k] = velo[3 * i + k] - f * dpot[3 * i + k];
}

struct MShake {
  double c[9];  // cmat(i,j) at c[3*(j-1) + (i-1)]
  double tol, roh2, rhh2, imh, imo;
};

// Matrix SHAKE of the position step of one water per thread
// (dynamics/mshakpt.f:19-117): the squared O-H, O-H and H-H distances of
// coor + step against their equilibrium values, Lagrange multipliers from
// the inverted constraint matrix, correction along the bond vectors of the
// coordinates before the step.  MOIL rebuilds a water whose error exceeds
// 0.4 or that needs a 19th iteration (placing its hydrogens anew); here
// such a water is left as it is and counted in status[0].
__global__ void k_mshake_pos(const double *__restrict__ coor, double *__restrict__ velo,
                             const int32_t *__restrict__ wat_o, int nwat, MShake M,
                             int32_t *__restrict__ status) {
#pragma clang fp contract(off)
  const int w = blockIdx.x * DY_THREADS + threadIdx.x;
  if (w >= nwat) return;
  const int o = wat_o[w], h1 = o + 1, h2 = o + 2;
  double r1[3], r2[3], rh[3], vo[3], v1[3], v2[3];
  for (int i = 0; i < 3; i++) {
    r1[i] = coor[3 * o + i] - coor[3 * h1 + i];
    r2[i] = coor[3 * o + i] - coor[3 * h2 + i];
    rh[i] = coor[3 * h1 + i] - coor[3 * h2 + i];
    vo[i] = velo[3 * o + i];
    v1[i] = velo[3 * h1 + i];
    v2[i] = velo[3 * h2 + i];
  }
  bool ok = false;
  for (int iter = 1; iter <= 200; iter++) {
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
    for (int i = 0; i < 3; i++) {
      double t = r1[i] + vo[i] - v1[i];
      d1 = d1 + t * t;
      t = r2[i] + vo[i] - v2[i];
      d2 = d2 + t * t;
      t = rh[i] + v1[i] - v2[i];
      d3 = d3 + t * t;
    }
    const double err[3] = {d1 - M.roh2, d2 - M.roh2, d3 - M.rhh2};
    // MOIL's check, in its order: the first error past 0.4 (or the 19th
    // iteration) ends the water, the first error above tolerance asks
    // for one more correction, and no error above tolerance converges
    int outcome = 0;  // 0 converged, 1 correct again, 2 failed
    for (int i = 0; i < 3; i++) {
      if (fabs(err[i]) > 0.4 || iter == 19) {
        outcome = 2;
        break;
      }
      if (fabs(err[i]) > M.tol) {
        outcome = 1;
        break;
      }
    }
    if (outcome == 0) {
      ok = true;
      break;
    }
    if (outcome == 2) break;
    double lam[3];
    for (int i = 0; i < 3; i++) {
      lam[i] = 0.0;
      for (int j = 0; j < 3; j++) lam[i] = lam[i] - M.c[3 * j + i] * err[j];
    }
    for (int i = 0; i < 3; i++) {
      vo[i] = vo[i] + 0.5 * M.imo * (lam[0] * r1[i] + lam[1] * r2[i]);
      v1[i] = v1[i] + 0.5 * M.imh * (-lam[0] * r1[i] + lam[2] * rh[i]);
      v2[i] = v2[i] + 0.5 * M.imh * (-lam[1] * r2[i] - lam[2] * rh[i]);
    }
  }
  if (!ok) atomicAdd(&status[0], 1);
  for (int i = 0; i < 3; i++) {
    velo[3 * o + i] = vo[i];
    velo[3 * h1 + i] = v1[i];
    velo[3 * h2 + i] = v2[i];
  }
}

// Matrix SHAKE of the velocities of one water per thread
// (dynamics/mshakvl.f:16-96): the velocity components along the three bond
// vectors of the new coordinates removed by up to 20 corrections (MOIL
// compares the signed errors with the tolerance); a water not converged
// after 20 is counted in status[1].
__global__ void k_mshake_vel(const double *__restrict__ coor, double *__restrict__ velo,
                             const int32_t *__restrict__ wat_o, int nwat, MShake M,
                             int32_t *__restrict__ status) {
#pragma clang fp contract(off)
  const int w = blockIdx.x * DY_THREADS + threadIdx.x;
  if (w >= nwat) return;
  const int o = wat_o[w], h1 = o + 1, h2 = o + 2;
  double r1[3], r2[3], rh[3], vo[3], v1[3], v2[3];
  for (int i = 0; i < 3; i++) {
    r1[i] = coor[3 * o + i] - coor[3 * h1 + i];
    r2[i] = coor[3 * o + i] - coor[3 * h2 + i];
    rh[i] = coor[3 * h1 + i] - coor[3 * h2 + i];
    vo[i] = velo[3 * o + i];
    v1[i] = velo[3 * h1 + i];
    v2[i] = velo[3 * h2 + i];
  }
  double err[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; i++) {
    err[0] = err[0] + (vo[i] - v1[i]) * r1[i];
    err[1] = err[1] + (vo[i] - v2[i]) * r2[i];
    err[2] = err[2] + (v1[i] - v2[i]) * rh[i];
  }
  bool ok = false;
  for (int iter = 1; iter <= 20 && !ok; iter++) {
    double lam[3];
    for (int i = 0; i < 3; i++) {
      lam[i] = 0.0;
      for (int j = 0; j < 3; j++) lam[i] = lam[i] - M.c[3 * j + i] * err[j];
    }
    for (int i = 0; i < 3; i++) {
      vo[i] = vo[i] + M.imo * (lam[0] * r1[i] + lam[1] * r2[i]);
      v1[i] = v1[i] + M.imh * (-lam[0] * r1[i] + lam[2] * rh[i]);
      v2[i] = v2[i] + M.imh * (-lam[1] * r2[i] - lam[2] * rh[i]);
    }
    err[0] = err[1] = err[2] = 0.0;
    for (int i = 0; i < 3; i++) {
      err[0] = err[0] + (vo[i] - v1[i]) * r1[i];
      err[1] = err[1] + (vo[i] - v2[i]) * r2[i];
      err[2] = err[2] + (v1[i] - v2[i]) * rh[i];
    }
    ok = !(err[0] > M.tol || err[1] > M.tol || err[2] > M.tol);
  }
  if (!ok) atomicAdd(&status[1], 1);
  for (int i = 0; i < 3; i++) {
    velo[3 * o + i] = vo[i];
    velo[3 * h1 + i] = v1[i];
    velo[3 * h2 + i] = v2[i];
  }
}

static int32_t dyna_fail(int32_t code, const std::string &m) {
  set_error(m);
  return code;
}

extern "C" int32_t moil_gpu_dyna_create(moil_ctx *c, const double *ptms,
                                        const moil_dyna_params *par, moil_dyna **out) {
  if (!c || !ptms || !par || !out) return dyna_fail(MOIL_GPU_EINVAL, "dyna_create: null argument");
  *out = nullptr;
  if (!(par->dt > 0) || par->nlist < 1)
    return dyna_fail(MOIL_GPU_EINVAL, "dyna_create: dt must be > 0 and nlist >= 1");
  const int n = c->npt;
  for (int i = 0; i < n; i++)
    if (!(ptms[i] > 0))
      return dyna_fail(MOIL_GPU_EINVAL, "dyna_create: every mass must be > 0 (atom " +
                                            std::to_string(i + 1) + ")");
  if (par->mshake && c->nwat == 0)
    return dyna_fail(MOIL_GPU_EINVAL, "dyna_create: matrix SHAKE asked for a system without waters");
  // factor1 = dt^2/2 / m, factor2 = dt/2 / m as dynamics/init_dyna.f:281-288
  const double twodt = par->dt * 0.5, dt2 = par->dt * par->dt * 0.5;
  std::vector<double> f1(n), f2(n);
  for (int i = 0; i < n; i++) {
    const double invms = 1.0 / ptms[i];
    f1[i] = dt2 * invms;
    f2[i] = twodt * invms;
  }
  moil_dyna *d = new moil_dyna();
  d->ctx = c;
  d->par = *par;
  const bool ok = hipMalloc(&d->d_f1, sizeof(double) * n) == hipSuccess &&
                  hipMalloc(&d->d_f2, sizeof(double) * n) == hipSuccess &&
                  hipMalloc(&d->d_status, sizeof(int32_t) * 2) == hipSuccess &&
                  hipMalloc(&d->d_en, sizeof(double) * MOIL_N_ENERGIES) == hipSuccess &&
                  hipMemcpy(d->d_f1, f1.data(), sizeof(double) * n, hipMemcpyHostToDevice) ==
                      hipSuccess &&
                  hipMemcpy(d->d_f2, f2.data(), sizeof(double) * n, hipMemcpyHostToDevice) ==
                      hipSuccess &&
                  hipMemset(d->d_status, 0, sizeof(int32_t) * 2) == hipSuccess;
  if (!ok) {
    hipFree(d->d_f1); hipFree(d->d_f2); hipFree(d->d_status); hipFree(d->d_en);
    delete d;
    return dyna_fail(MOIL_GPU_EHIP, "dyna_create: device allocation failed");
  }
  *out = d;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_dyna_destroy(moil_dyna *d) {
  if (!d) return MOIL_GPU_OK;
  hipDeviceSynchronize();
  hipFree(d->d_f1);
  hipFree(d->d_f2);
  hipFree(d->d_status);
  hipFree(d->d_en);
  delete d;
  return MOIL_GPU_OK;
}

// the SHAKE failure counts since the last check (one host read) and clear
static int32_t dyna_check(moil_dyna *d, hipStream_t st) {
  int32_t s[2];
  HIP_CHECK(hipMemcpyAsync(s, d->d_status, sizeof(s), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (s[0] || s[1]) {
    HIP_CHECK(hipMemsetAsync(d->d_status, 0, sizeof(s), st));
    return dyna_fail(MOIL_GPU_ECAPACITY,
                     "dyna: matrix SHAKE did not converge for " + std::to_string(s[0]) +
                         " water(s) in a position step and " + std::to_string(s[1]) +
                         " in a velocity step (MOIL rebuilds or stops there)");
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_dyna_run_dev(moil_dyna *d, double *d_coor, double *d_velo,
                                         double *d_dpot, int32_t nstep, double *d_energies,
                                         void *stream) {
  if (!d || !d_coor || !d_velo || !d_dpot) return dyna_fail(MOIL_GPU_EINVAL, "dyna_run: null argument");
  if (nstep < 0) return dyna_fail(MOIL_GPU_EINVAL, "dyna_run: nstep < 0");
  moil_ctx *c = d->ctx;
  hipStream_t st = pick_stream(c, stream);
  const int n = c->npt, nb = (n + DY_THREADS - 1) / DY_THREADS;
  const int nw = (c->nwat + DY_THREADS - 1) / DY_THREADS;
  MShake M;
  std::memcpy(M.c, d->par.cmat, sizeof(M.c));
  M.tol = d->par.tolcons;
  M.roh2 = d->par.reqoh2;
  M.rhh2 = d->par.reqhh2;
  M.imh = d->par.invmh;
  M.imo = d->par.invmo;
  const double dt = d->par.dt, rdt = 1.0 / dt;
  const bool shk = d->par.mshake != 0;
  for (int istep = 0; istep < nstep; istep += d->par.nlist) {
    // list update: squeeze, lists, forces on the new lists (dyna.f:129-144)
    int32_t rc = moil_gpu_squeeze_dev(c, d_coor, st);
    if (rc == MOIL_GPU_OK) rc = moil_gpu_build_lists_dev(c, d_coor, st);
    if (rc == MOIL_GPU_OK) rc = dyna_check(d, st);
    if (rc == MOIL_GPU_OK) rc = moil_gpu_eforce_dev(c, d_coor, d_dpot, d->d_en, st);
    if (rc != MOIL_GPU_OK) return rc;
    const int jend = std::min(nstep, istep + d->par.nlist);
    for (int j = istep; j < jend; j++) {
      k_dyna_posstep<<<nb, DY_THREADS, 0, st>>>(d_velo, d_dpot, d->d_f1, dt, n);
      if (shk) k_mshake_pos<<<nw, DY_THREADS, 0, st>>>(d_coor, d_velo, c->d_wat_o, c->nwat, M,
                                                       d->d_status);
      k_dyna_move<<<nb, DY_THREADS, 0, st>>>(d_coor, d_velo, rdt, n);
      rc = moil_gpu_eforce_dev(c, d_coor, d_dpot,
                               d_energies ? d_energies + (size_t)MOIL_N_ENERGIES * j : d->d_en, st);
      if (rc != MOIL_GPU_OK) return rc;
      k_dyna_kick<<<nb, DY_THREADS, 0, st>>>(d_velo, d_dpot, d->d_f2, n);
      if (shk) k_mshake_vel<<<nw, DY_THREADS, 0, st>>>(d_coor, d_velo, c->d_wat_o, c->nwat, M,
                                                       d->d_status);
      HIP_CHECK(hipGetLastError());
    }
  }
  return dyna_check(d, st);
}

extern "C" int32_t moil_gpu_dyna_run(moil_dyna *d, double *coor, double *velo, double *dpot,
                                     int32_t nstep, double *energies) {
  if (!d || !coor || !velo || !dpot) return dyna_fail(MOIL_GPU_EINVAL, "dyna_run: null argument");
  moil_ctx *c = d->ctx;
  const size_t n3 = 3 * (size_t)c->npt;
  double *buf = nullptr;
  const size_t ne = energies ? (size_t)MOIL_N_ENERGIES * std::max(nstep, 1) : 0;
  HIP_CHECK(hipMalloc(&buf, sizeof(double) * (n3 + std::max<size_t>(ne, 1))));
  double *dv = buf, *de = buf + n3;  // velocities, then the energy records
  int32_t rc = MOIL_GPU_OK;
  if (hipMemcpy(c->d_coor, coor, sizeof(double) * n3, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dv, velo, sizeof(double) * n3, hipMemcpyHostToDevice) != hipSuccess)
    rc = dyna_fail(MOIL_GPU_EHIP, "dyna_run: copy to the device failed");
  if (rc == MOIL_GPU_OK)
    rc = moil_gpu_dyna_run_dev(d, c->d_coor, dv, c->d_dpot, nstep, energies ? de : nullptr,
                               c->stream);
  if (rc == MOIL_GPU_OK) {
    if (hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(coor, c->d_coor, sizeof(double) * n3, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(velo, dv, sizeof(double) * n3, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(dpot, c->d_dpot, sizeof(double) * n3, hipMemcpyDeviceToHost) != hipSuccess ||
        (energies && ne &&
         hipMemcpy(energies, de, sizeof(double) * ne, hipMemcpyDeviceToHost) != hipSuccess))
      rc = dyna_fail(MOIL_GPU_EHIP, "dyna_run: copy from the device failed");
  }
  hipFree(buf);
  return rc;
}
