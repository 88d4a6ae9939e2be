// Chain-of-replicas functional: MOIL's sds (chain/sds.f:1-257), the energy
// and gradient its chain minimiser, annealer and chain dynamics evaluate
// (chain/chmin1.f, chain/chnneal.f:85, chain/chndyn.f:112):
//
//   S = sum_{j=2}^{igrid-1} V(r_j)                       interior replicas
//     + gamma sum_{j=1}^{igrid-1} (d_j - <d>)^2           nearest springs
//     + rho/lambda sum_{j=1}^{igrid-2} exp(-lambda D_j / <d>^2)
//
// d_j = |r_{j+1} - r_j| and D_j = |r_{j+2} - r_j|^2 over the selected atoms
// (ipick > 0), <d> the mean of the d_j.
//
// One call:
//   1. copies the interior replicas into scratch (sds.f:75-80 copies each
//      into coor, so squeeze never moves the caller's replicas);
//   2. squeezes each copy and rebuilds its lists (sds.f:84-86), the
//      replicas' builds queued side by side on their own streams with one
//      host wait for all;
//   3. evaluates all interior replicas in one batched evaluation, whose
//      gradients land straight in their dv rows (sds.f:87-106);
//   4. runs three small kernels for the chain terms:
//      k_chain_dist   the 2 igrid - 3 squared distances, as partial sums
//                     per workgroup (one workgroup per atom block and
//                     replica: the pairs ending at that replica);
//      k_chain_coef   one workgroup: the sums, <d>, the energy terms and
//                     the per-distance gradient factors (sds.f:149-195);
//      k_chain_grad   one thread per (replica, atom): the chain terms of
//                     the gradient (sds.f:199-250), added to the replica's
//                     own gradient in sds.f's order, without contraction
//                     into fused multiply-adds (sds.f's separate product
//                     and sum).
// The distance sums are tree sums (sds.f sums sequentially); everything
// else follows sds.f operation by operation.
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.h"

namespace {

constexpr int CH_BLOCK = 256;
constexpr int CH_MAX_GRID = 1024;  // replicas per chain (LDS of k_chain_coef)

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// part[j][blk] = selected-atom sum of |r_{j+1} - r_j|^2 (j < igrid-1) and
// part[igrid-1+j][blk] = that of |r_{j+2} - r_j|^2 (j < igrid-2); workgroup
// (blk, s-1) handles the pairs ending at replica s >= 1 of its atom block
__global__ __launch_bounds__(CH_BLOCK) void k_chain_dist(const double *__restrict__ r,
                                                         const int32_t *__restrict__ pick,
                                                         int npt, int igrid,
                                                         double *__restrict__ part) {
  __shared__ double red[2][CH_BLOCK / 64];
  const int i = blockIdx.x * CH_BLOCK + threadIdx.x;
  const int s = blockIdx.y + 1;
  const int nblk = gridDim.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double v1 = 0.0, v2 = 0.0;
  if (i < npt && pick[i] > 0) {
    const double *x = r + 3 * ((size_t)s * npt + i);
    const double *p1 = r + 3 * ((size_t)(s - 1) * npt + i);
    const double a = x[0] - p1[0], b = x[1] - p1[1], c = x[2] - p1[2];
    v1 = a * a + b * b + c * c;
    if (s >= 2) {
      const double *p2 = r + 3 * ((size_t)(s - 2) * npt + i);
      const double e = x[0] - p2[0], f = x[1] - p2[1], g = x[2] - p2[2];
      v2 = e * e + f * f + g * g;
    }
  }
  v1 = wave_sum(v1);
  v2 = wave_sum(v2);
  if (lane == 0) {
    red[0][w] = v1;
    red[1][w] = v2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t1 = 0.0, t2 = 0.0;
    for (int k = 0; k < CH_BLOCK / 64; k++) {
      t1 += red[0][k];
      t2 += red[1][k];
    }
    part[(size_t)(s - 1) * nblk + blockIdx.x] = t1;
    if (s >= 2) part[(size_t)(igrid - 1 + s - 2) * nblk + blockIdx.x] = t2;
  }
}

// One workgroup.  d0/e1 as MOIL returns them (nearest distances, then the
// squared next-nearest ones; spring factors, then the repulsion
// exponentials over <d>^2); e0 and the energy of the interior replicas;
// coef = [spring factors e1(j) | 2 rho e1(igrid-1+j) | d2w / d0(j)].
// Everything lives in LDS (igrid <= CH_MAX_GRID); the elementwise work is
// spread over the workgroup, the sums over replicas run in one thread in
// sds.f's order.
__global__ __launch_bounds__(CH_BLOCK) void k_chain_coef(
    const double *__restrict__ part, int nblk, int igrid, const double *__restrict__ en,
    double gamma, double rho, double lambda, double *__restrict__ coef, double *__restrict__ d0,
    double *__restrict__ e0, double *__restrict__ e1, double *__restrict__ sener) {
#pragma clang fp contract(off)
  // [0, 2 igrid - 3): d0; [2 igrid - 1, 3 igrid - 3): the exponentials
  __shared__ double dd[3 * CH_MAX_GRID];
  __shared__ double ee[CH_MAX_GRID];      // interior energies
  __shared__ double sc[4];                // dave, dave2, d2w, energy
  const int nd = 2 * igrid - 3, t = threadIdx.x;
  for (int j = t; j < nd; j += CH_BLOCK) {
    double v = 0.0;
    for (int b = 0; b < nblk; b++) v += part[(size_t)j * nblk + b];
    dd[j] = j < igrid - 1 ? sqrt(v) : v;  // sds.f:149-152
  }
  for (int j = t + 1; j < igrid - 1; j += CH_BLOCK) {
    ee[j] = en[(size_t)(j - 1) * MOIL_N_ENERGIES];  // e_total
    e0[j] = ee[j];
  }
  __syncthreads();
  if (t == 0) {
    double s = 0.0;  // sds.f:71-104
    for (int j = 1; j < igrid - 1; j++) s = s + ee[j];
    double dave = 0.0;  // sds.f:149-154
    for (int j = 0; j < igrid - 1; j++) dave = dave + dd[j];
    dave = dave / (double)(igrid - 1);
    double tmp = 0.0;  // sds.f:169-174
    for (int j = 0; j < igrid - 1; j++) {
      const double t1 = dd[j] - dave;
      tmp = tmp + t1 * t1;
    }
    sc[0] = dave;
    sc[1] = 1.0 / (dave * dave);
    sc[3] = s + gamma * tmp;
  }
  __syncthreads();
  const double dave = sc[0], dave2 = sc[1];
  // sds.f:180-188: the exponentials
  for (int j = igrid - 1 + t; j < nd; j += CH_BLOCK)
    dd[igrid + j] = exp(-(lambda * dd[j] / (dave * dave)));
  __syncthreads();
  if (t == 0) {
    double tmp = 0.0, d2w = 0.0;
    for (int j = igrid - 1; j < nd; j++) {
      tmp = tmp + dd[igrid + j];
      d2w = d2w + dd[igrid + j] * dd[j];
    }
    *sener = sc[3] + rho / lambda * tmp;
    sc[2] = 2.0 * rho * dave2 / ((double)(igrid - 1) * dave) * d2w;
  }
  __syncthreads();
  const double d2w = sc[2], gam2 = 2.0 * gamma, rho2 = 2.0 * rho;
  double *cA = coef, *cB = coef + (igrid - 1), *cC = coef + (2 * igrid - 3);
  for (int j = t; j < nd; j += CH_BLOCK) {
    d0[j] = dd[j];
    if (j < igrid - 1) {
      const double f = gam2 * (1.0 - dave / dd[j]);  // sds.f:193-195
      e1[j] = f;
      cA[j] = f;
      cC[j] = d2w / dd[j];
    } else {
      const double f = dd[igrid + j] * dave2;
      e1[j] = f;
      cB[j - (igrid - 1)] = rho2 * f;
    }
  }
}

// dv(replica s, atom i): the replica's own gradient (interior) or zero
// (ends), plus the chain terms in the order sds.f adds them: springs
// (sds.f:199-214, pair s-1 then pair s), repulsion (sds.f:218-235, pair
// s-2 then pair s), the mean distance's dependence (sds.f:236-250).
__global__ __launch_bounds__(CH_BLOCK) void k_chain_grad(const double *__restrict__ r,
                                                         const int32_t *__restrict__ pick,
                                                         int npt, int igrid,
                                                         const double *__restrict__ coef,
                                                         double *__restrict__ dv) {
#pragma clang fp contract(off)
  const size_t idx = (size_t)blockIdx.x * CH_BLOCK + threadIdx.x;
  if (idx >= (size_t)igrid * npt) return;
  const int s = (int)(idx / npt), i = (int)(idx - (size_t)s * npt);
  const bool interior = s > 0 && s < igrid - 1;
  double *out = dv + 3 * idx;
  double g[3];
  for (int k = 0; k < 3; k++) g[k] = interior ? out[k] : 0.0;
  if (pick[i] > 0) {
    const double *cA = coef, *cB = coef + (igrid - 1), *cC = coef + (2 * igrid - 3);
    const double *x = r + 3 * idx;
    // atom i of replica t (only formed for 0 <= t < igrid)
    auto rep = [&](int t) { return r + 3 * ((size_t)t * npt + i); };
    const double *xm1 = s >= 1 ? rep(s - 1) : x, *xp1 = s <= igrid - 2 ? rep(s + 1) : x;
    const double *xm2 = s >= 2 ? rep(s - 2) : x, *xp2 = s <= igrid - 3 ? rep(s + 2) : x;
    if (s >= 1) {
      const double c = cA[s - 1];
      for (int k = 0; k < 3; k++) g[k] = g[k] - c * (xm1[k] - x[k]);
    }
    if (s <= igrid - 2) {
      const double c = cA[s];
      for (int k = 0; k < 3; k++) g[k] = g[k] + c * (x[k] - xp1[k]);
    }
    if (s >= 2) {
      const double c = cB[s - 2];
      for (int k = 0; k < 3; k++) g[k] = g[k] + (xm2[k] - x[k]) * c;
    }
    if (s <= igrid - 3) {
      const double c = cB[s];
      for (int k = 0; k < 3; k++) g[k] = g[k] - (x[k] - xp2[k]) * c;
    }
    if (s >= 1) {
      const double c = cC[s - 1];
      for (int k = 0; k < 3; k++) g[k] = g[k] - (xm1[k] - x[k]) * c;
    }
    if (s <= igrid - 2) {
      const double c = cC[s];
      for (int k = 0; k < 3; k++) g[k] = g[k] + (x[k] - xp1[k]) * c;
    }
  }
  for (int k = 0; k < 3; k++) out[k] = g[k];
}

}  // namespace

struct moil_chain {
  int igrid = 0, npt = 0, nin = 0, nblk = 0;
  std::vector<moil_ctx *> ctx;  // interior replicas 2..igrid-1
  moil_batch *batch = nullptr;
  double *d_scr = nullptr;   // (nin, npt, 3) squeezed copies
  double *d_en = nullptr;    // (nin, MOIL_N_ENERGIES)
  double *d_part = nullptr;  // (2 igrid - 3, nblk)
  double *d_coef = nullptr;  // 3 igrid
  // staging of the host entry point
  double *d_r = nullptr, *d_dv = nullptr, *d_out = nullptr;
  int32_t *d_pick = nullptr;
};

static void chain_free(moil_chain *ch) {
  if (!ch) return;
  hipDeviceSynchronize();
  if (ch->batch) moil_gpu_batch_destroy(ch->batch);
  hipFree(ch->d_scr);
  hipFree(ch->d_en);
  hipFree(ch->d_part);
  hipFree(ch->d_coef);
  hipFree(ch->d_r);
  hipFree(ch->d_dv);
  hipFree(ch->d_out);
  hipFree(ch->d_pick);
  delete ch;
}

extern "C" int32_t moil_gpu_chain_create(moil_ctx **slices, int32_t igrid, moil_chain **out) {
  if (!slices || !out || igrid < 3 || igrid > CH_MAX_GRID) {
    set_error("chain_create: null argument, or fewer than 3 or more than " +
              std::to_string(CH_MAX_GRID) + " replicas");
    return MOIL_GPU_EINVAL;
  }
  const int nin = igrid - 2;
  for (int k = 0; k < nin; k++)
    if (!slices[k] || slices[k]->npt != slices[0]->npt) {
      set_error("chain_create: null replica context or replicas of different sizes");
      return MOIL_GPU_EINVAL;
    }
  moil_chain *ch = new moil_chain();
  ch->igrid = igrid;
  ch->nin = nin;
  ch->npt = slices[0]->npt;
  ch->nblk = (ch->npt + CH_BLOCK - 1) / CH_BLOCK;
  ch->ctx.assign(slices, slices + nin);
  int32_t rc = moil_gpu_batch_create(slices, nin, 1, &ch->batch);
  if (rc != MOIL_GPU_OK) {
    ch->batch = nullptr;
    chain_free(ch);
    return rc;
  }
  // every call rebuilds the lists, so graphs would be recaptured each time
  rc = moil_gpu_batch_set_graph(ch->batch, 0);
  if (rc != MOIL_GPU_OK) {
    chain_free(ch);
    return rc;
  }
  const size_t n3 = 3 * (size_t)ch->npt;
  const bool ok =
      hipMalloc(&ch->d_scr, sizeof(double) * n3 * nin) == hipSuccess &&
      hipMalloc(&ch->d_en, sizeof(double) * MOIL_N_ENERGIES * nin) == hipSuccess &&
      hipMalloc(&ch->d_part, sizeof(double) * (2 * igrid - 3) * (size_t)ch->nblk) == hipSuccess &&
      hipMalloc(&ch->d_coef, sizeof(double) * 3 * igrid) == hipSuccess;
  if (!ok) {
    chain_free(ch);
    set_error("chain_create: device allocation failed");
    return MOIL_GPU_EHIP;
  }
  *out = ch;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_chain_destroy(moil_chain *ch) {
  chain_free(ch);
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_chain_sds_dev(moil_chain *ch, const double *d_r,
                                          const int32_t *d_pick, double gamma, double rho,
                                          double lambda, double *d_dv, double *d_d0,
                                          double *d_e0, double *d_e1, double *d_sener,
                                          void *stream) {
  if (!ch || !d_r || !d_pick || !d_dv || !d_d0 || !d_e0 || !d_e1 || !d_sener) {
    set_error("chain_sds: null argument");
    return MOIL_GPU_EINVAL;
  }
  if (d_dv == d_r) {
    set_error("chain_sds: dv must not alias r");
    return MOIL_GPU_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int igrid = ch->igrid, npt = ch->npt, nin = ch->nin;
  const size_t n3 = 3 * (size_t)npt;
  HIP_CHECK(hipMemcpyAsync(ch->d_scr, d_r + n3, sizeof(double) * n3 * nin,
                           hipMemcpyDeviceToDevice, st));
  std::vector<double *> xm(nin);
  std::vector<const double *> xs(nin);
  std::vector<double *> gs(nin), es(nin);
  for (int k = 0; k < nin; k++) {
    xm[k] = ch->d_scr + n3 * k;
    xs[k] = xm[k];
    gs[k] = d_dv + n3 * (k + 1);
    es[k] = ch->d_en + (size_t)MOIL_N_ENERGIES * k;
  }
  // every replica squeezed and its lists rebuilt, the builds side by side
  // on the replicas' own streams, one host wait for all
  int32_t rc = moil_gpu_batch_build_lists_dev(ch->batch, xm.data(), 1, st);
  if (rc != MOIL_GPU_OK) return rc;
  rc = moil_gpu_batch_eforce_dev(ch->batch, xs.data(), gs.data(), es.data(), st);
  if (rc != MOIL_GPU_OK) return rc;
  k_chain_dist<<<dim3(ch->nblk, igrid - 1), CH_BLOCK, 0, st>>>(d_r, d_pick, npt, igrid,
                                                              ch->d_part);
  k_chain_coef<<<1, CH_BLOCK, 0, st>>>(ch->d_part, ch->nblk, igrid, ch->d_en, gamma, rho, lambda,
                                       ch->d_coef, d_d0, d_e0, d_e1, d_sener);
  const size_t nt = (size_t)igrid * npt;
  k_chain_grad<<<(unsigned)((nt + CH_BLOCK - 1) / CH_BLOCK), CH_BLOCK, 0, st>>>(
      d_r, d_pick, npt, igrid, ch->d_coef, d_dv);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_chain_sds(moil_chain *ch, const double *r, const int32_t *ipick,
                                      double gamma, double rho, double lambda, double *sener,
                                      double *dv, double *d0, double *e0, double *e1) {
  if (!ch || !r || !ipick || !sener || !dv || !d0 || !e0 || !e1) {
    set_error("chain_sds: null argument");
    return MOIL_GPU_EINVAL;
  }
  const int igrid = ch->igrid, nd = 2 * igrid - 3;
  const size_t nr = 3 * (size_t)ch->npt * igrid;
  if (!ch->d_r) {
    const bool ok = hipMalloc(&ch->d_r, sizeof(double) * nr) == hipSuccess &&
                    hipMalloc(&ch->d_dv, sizeof(double) * nr) == hipSuccess &&
                    hipMalloc(&ch->d_out, sizeof(double) * (2 * nd + igrid + 1)) == hipSuccess &&
                    hipMalloc(&ch->d_pick, sizeof(int32_t) * ch->npt) == hipSuccess;
    if (!ok) {
      // all four or none: the next call allocates again
      void *bufs[4] = {ch->d_r, ch->d_dv, ch->d_out, ch->d_pick};
      for (void *b : bufs)
        if (b) hipFree(b);
      ch->d_r = ch->d_dv = ch->d_out = nullptr;
      ch->d_pick = nullptr;
      set_error("chain_sds: device allocation failed");
      return MOIL_GPU_EHIP;
    }
  }
  HIP_CHECK(hipMemcpy(ch->d_r, r, sizeof(double) * nr, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(ch->d_pick, ipick, sizeof(int32_t) * ch->npt, hipMemcpyHostToDevice));
  double *o_d0 = ch->d_out, *o_e1 = o_d0 + nd, *o_e0 = o_e1 + nd, *o_s = o_e0 + igrid;
  const int32_t rc = moil_gpu_chain_sds_dev(ch, ch->d_r, ch->d_pick, gamma, rho, lambda, ch->d_dv,
                                            o_d0, o_e0, o_e1, o_s, nullptr);
  if (rc != MOIL_GPU_OK) return rc;
  HIP_CHECK(hipMemcpy(dv, ch->d_dv, sizeof(double) * nr, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(d0, o_d0, sizeof(double) * nd, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(e1, o_e1, sizeof(double) * nd, hipMemcpyDeviceToHost));
  // MOIL sets e0 of the interior replicas only
  HIP_CHECK(hipMemcpy(e0 + 1, o_e0 + 1, sizeof(double) * (igrid - 2), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(sener, o_s, sizeof(double), hipMemcpyDeviceToHost));
  return MOIL_GPU_OK;
}
