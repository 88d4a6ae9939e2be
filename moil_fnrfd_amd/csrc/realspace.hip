// Real-space Lennard-Jones + Ewald-erfc Coulomb for MOIL's water groups
// (pot/watwat_ewald.f:131-612,624-1085 and symm/symwat_ewald.f:96-620,
// 640-1150): per listed water pair, O-O LJ when r_OO^2 <= rvmx^2 (listwt1
// only) and all nine site-pair erfc Coulomb terms with no cutoff check
// (listwt1) or only when r_OO^2 <= relx^2 (listwt2).  Charges and LJ
// parameters are those of the last water (pot/watwat_ewald.f:40-72).
//
// Work decomposition (see lists.hip): a workgroup takes a chunk of the
// items of one 64-water block, sized by its pair count.  The block's
// waters (the translated pair members) sit in LDS; each lane owns an item
// (partner water j plus a mask of at most WSPLIT = 32 bits over the block) and
// walks the mask.  Items come sorted by popcount, so the lanes of a
// wavefront run loops of equal length.  j-forces accumulate in registers
// and are stored once per item, in fixed point, into the item's row of a
// partial buffer (rows grouped by partner water, so each water's rows are
// contiguous); i-forces accumulate in LDS with 64-bit integer atomics and are
// stored once per workgroup into the chunk's row.  No global atomics: on
// this GPU they are performed at the memory side and cost far more than
// the stores plus one gather pass (k_water_reduce) that sums each water's
// rows.  All partials are 64-bit fixed point, so the forces are
// independent of execution order (bit-reproducible).
//
// MOIL_PREC_MIXED: FP32 pair arithmetic.  Positions are carried as the O
// atom (float hi + float lo) and the H offsets from their O, so every
// site-site vector is formed from the O-O difference plus small offsets
// (relative error ~1e-7 of the O-O distance instead of the box size).
// erfc(x) = exp(-x^2) t P(t), t = 1/(1+p x): the functional form of
// Abramowitz & Stegun 7.1.26 with a degree-8 polynomial fitted here for
// relative accuracy 3e-8 on [0, 6].  Forces are rounded per term to
// 64-bit fixed point with FIX_SHIFT_MIXED = 20 fractional bits (nearest)
// before any sum: the partner's rows as FP32 sums converted by the
// reduction, the block members' terms in the LDS sums.
// MOIL_PREC_FP64: FP64 arithmetic in MOIL's operation order with MOIL's
// cubic-spline erfc table (pot/cdie_ewald.f:152-160).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "multi.h"
#include "pairmath.h"

#define RS_THREADS 256

// the bit pattern of 1.5 2^32 (the water kernel's i-side sums, see there)
#define kBias32 FIX_BIAS20

// ---------------------------------------------------------------------------
// pack water positions per slot
__device__ __forceinline__ void pack_f32_body(const PackArgs &A, int bid, int nblk) {
  int s = bid * RS_THREADS + threadIdx.x;
  if (s >= A.nwat) return;
  const double *coor = A.coor;
  float4 *wpos = (float4 *)A.wpos;
  int o = A.slot_o[s];
  double ox = coor[3 * o], oy = coor[3 * o + 1], oz = coor[3 * o + 2];
  float hx = (float)ox, hy = (float)oy, hz = (float)oz;
  float lx = (float)(ox - (double)hx), ly = (float)(oy - (double)hy), lz = (float)(oz - (double)hz);
  float h1x = (float)(coor[3 * o + 3] - ox), h1y = (float)(coor[3 * o + 4] - oy),
        h1z = (float)(coor[3 * o + 5] - oz);
  float h2x = (float)(coor[3 * o + 6] - ox), h2y = (float)(coor[3 * o + 7] - oy),
        h2z = (float)(coor[3 * o + 8] - oz);
  wpos[3 * s] = make_float4(hx, hy, hz, h1x);
  wpos[3 * s + 1] = make_float4(lx, ly, lz, h1y);
  wpos[3 * s + 2] = make_float4(h1z, h2x, h2y, h2z);
}
DEFINE_KERNEL(k_pack_water_f32, PackArgs, RS_THREADS, pack_f32_body)

__device__ __forceinline__ void pack_f64_body(const PackArgs &A, int bid, int nblk) {
  int s = bid * RS_THREADS + threadIdx.x;
  if (s >= A.nwat) return;
  int o = A.slot_o[s];
  double *wpos = (double *)A.wpos;
#pragma unroll
  for (int k = 0; k < 9; k++) wpos[9 * s + k] = A.coor[3 * o + k];
}
DEFINE_KERNEL(k_pack_water_f64, PackArgs, RS_THREADS, pack_f64_body)

// the energy terms of a water chunk's four sums (real LJ, real Coulomb,
// image LJ, image Coulomb), in the order wg_sum4 returns them
__constant__ const int kWaterEn[4] = {EN_VDW, EN_EL, EN_VSYM, EN_LSYM};

// One listed water pair in FP32: the partner's O (hi + lo, translated)
// and H offsets, and the constants of the pair terms
struct WPartner {
  float jhx, jhy, jhz, jlx, jly, jlz;
  float j1x, j1y, j1z, j2x, j2y, j2z;
};
struct WConst {
  float nb2, pb, qoo, qoh, qhh, qcoo, qcoh, qchh, a12, b6, cutvdw2, cutele2;
};
// The nine site pairs of block member (A, B, C) and partner J, whose O-O
// distance is within the pair's list cutoff (relx for a listwt2 member,
// which the caller has tested): the member's gradient components in fi,
// the partner's added to fj and the pair energies to elj / eel, O-O LJ for
// a listwt1 member (lj) within rvmx.  Every sum of a product is an
// explicit fmaf (no add of a multiply is left for the compiler to fuse),
// so every call site computes the same values.
__device__ __forceinline__ void water_entry(const float4 &A, const float4 &B, const float4 &C,
                                            const WPartner &J, bool lj, const WConst &K,
                                            float fi[9], float fj[9], float &elj, float &eel) {
  const float dx = (A.x - J.jhx) + (B.x - J.jlx);
  const float dy = (A.y - J.jhy) + (B.y - J.jly);
  const float dz = (A.z - J.jhz) + (B.z - J.jlz);
  const float roo2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  // site vectors r_ab = x_ia - x_jb
  float ux[3], uy[3], uz[3];
  ux[0] = dx; uy[0] = dy; uz[0] = dz;
  ux[1] = dx + A.w; uy[1] = dy + B.w; uz[1] = dz + C.x;
  ux[2] = dx + C.y; uy[2] = dy + C.z; uz[2] = dz + C.w;
  const float vx[3] = {0.f, J.j1x, J.j2x}, vy[3] = {0.f, J.j1y, J.j2y},
              vz[3] = {0.f, J.j1z, J.j2z};
#pragma unroll
  for (int k = 0; k < 9; k++) fi[k] = 0.f;
#pragma unroll
  for (int sa = 0; sa < 3; sa++) {
#pragma unroll
    for (int sb = 0; sb < 3; sb++) {
      const float rx = ux[sa] - vx[sb], ry = uy[sa] - vy[sb], rz = uz[sa] - vz[sb];
      const float r2 = (sa == 0 && sb == 0) ? roo2 : fmaf(rz, rz, fmaf(ry, ry, rx * rx));
      const float qq = (sa == 0 && sb == 0) ? K.qoo : ((sa == 0 || sb == 0) ? K.qoh : K.qhh);
      const float qc = (sa == 0 && sb == 0) ? K.qcoo : ((sa == 0 || sb == 0) ? K.qcoh : K.qchh);
      const float s = __builtin_amdgcn_rsqf(r2);
      const float s2 = s * s;
      const float tt = __builtin_amdgcn_rcpf(fmaf(K.pb, r2 * s, 1.0f));
      const float ex = __builtin_amdgcn_exp2f(r2 * K.nb2);
      const float x = ex * erfc_poly(tt) * s;
      const float e = qq * x;
      eel = fmaf(qq, x, eel);
      float df = -s2 * fmaf(qc, ex, e);
      if (sa == 0 && sb == 0 && lj && roo2 <= K.cutvdw2) {
        const float s6 = s2 * s2 * s2;
        const float a6 = K.a12 * s6;  // A12 = a6 s6
        const float e1 = fmaf(a6, s6, -(K.b6 * s6));
        elj += e1;
        df = fmaf(-6.0f * s2, fmaf(a6, s6, e1), df);
      }
      fi[3 * sa] = fmaf(df, rx, fi[3 * sa]);
      fi[3 * sa + 1] = fmaf(df, ry, fi[3 * sa + 1]);
      fi[3 * sa + 2] = fmaf(df, rz, fi[3 * sa + 2]);
      fj[3 * sb] = fmaf(-df, rx, fj[3 * sb]);
      fj[3 * sb + 1] = fmaf(-df, ry, fj[3 * sb + 1]);
      fj[3 * sb + 2] = fmaf(-df, rz, fj[3 * sb + 2]);
    }
  }
}

// An entry's i-side components into the block's LDS sums as the bit
// patterns of (double)f + 1.5 2^32 (exact for |f| < 2^31, see
// water_f32_impl); the result tells whether the entry has a larger
// component, whose pattern the caller replaces afterwards by the exact
// value (the entry recomputed bit for bit).  A NaN component, which the
// maximum passes over, comes from a NaN or infinite site-pair term that
// reaches the partner's row as well, where the reduction's conversion
// reports it.
__device__ __forceinline__ bool add_i_side(const float fi[9], int i, unsigned long long *iacc,
                                           unsigned int *icnt) {
  const float fmx = fmaxf(fmaxf(fmaxf(fabsf(fi[0]), fabsf(fi[1])), fabsf(fi[2])),
                          fmaxf(fmaxf(fmaxf(fabsf(fi[3]), fabsf(fi[4])), fabsf(fi[5])),
                                fmaxf(fmaxf(fabsf(fi[6]), fabsf(fi[7])), fabsf(fi[8]))));
#pragma unroll
  for (int k = 0; k < 9; k++)
    atomicAdd(&iacc[k * WBLOCK + i],
              (unsigned long long)__double_as_longlong((double)fi[k] + 6442450944.0));
  atomicAdd(&icnt[i], 1u);
  return !(fmx < 2147483648.0f);
}

// FP32 (mixed precision) water kernel
__device__ __forceinline__ void water_f32_impl(
    const int32_t *__restrict__ chunks, const int32_t *__restrict__ order,
    const WItem *__restrict__ items,
    const int32_t *__restrict__ jpos, const float4 *__restrict__ wpos, int nwat, const WaterRS &P,
    long long *__restrict__ jpart, long long *__restrict__ ipart, double *__restrict__ epart,
    int max_parts, int bid, int nblk) {
  // i-force partials in LDS, component-major (member i's component k at
  // k * WBLOCK + i): sums of the bit patterns of (double)f + 1.5 2^32,
  // which for |f| < 2^31 is the 64-bit integer round(f 2^20) plus the
  // pattern of 1.5 2^32 (the fixed point of FIX_SHIFT_MIXED = 20), and
  // each member's number of added terms, which takes the patterns'
  // constant part out again at the end (two instructions per value in
  // place of a float-to-fixed-point conversion).
  __shared__ float4 iA[WBLOCK], iB[WBLOCK], iC[WBLOCK];
  __shared__ unsigned long long iacc[WBLOCK * 9];
  __shared__ unsigned int icnt[WBLOCK];
  __shared__ double wsum[16];
  // each thread's FP64 energy sums (LJ / Coulomb, real / image items), in
  // LDS: in registers they cost the entry loop its registers
  __shared__ double eacc[4][RS_THREADS];
  const int ch = order[bid], t = threadIdx.x;  // largest first
  const int blk = chunks[3 * ch], beg = chunks[3 * ch + 1], end = chunks[3 * ch + 2];
  const int s0 = blk * WBLOCK;
  if (t < WBLOCK) {
    if (s0 + t < nwat) {
      iA[t] = wpos[3 * (s0 + t)];
      iB[t] = wpos[3 * (s0 + t) + 1];
      iC[t] = wpos[3 * (s0 + t) + 2];
    }
  }
  for (int u = t; u < WBLOCK * 9; u += RS_THREADS) iacc[u] = 0;
  if (t < WBLOCK) icnt[t] = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) eacc[k][t] = 0.0;
  __syncthreads();

  // exp(-beta^2 r^2) = exp2(-beta^2 log2(e) r^2), t = 1 / (1 + p beta r)
  WConst K;
  K.nb2 = (float)(-P.beta * P.beta * 1.4426950408889634);
  K.pb = (float)(ERFC_P * P.beta);
  K.qoo = (float)P.qq_oo; K.qoh = (float)P.qq_oh; K.qhh = (float)P.qq_hh;
  const float c2b = (float)(2.0 / sqrt(M_PI) * P.beta);
  K.qcoo = K.qoo * c2b; K.qcoh = K.qoh * c2b; K.qchh = K.qhh * c2b;
  K.a12 = (float)P.oo12; K.b6 = (float)P.oo6;
  K.cutvdw2 = (float)P.cutvdw2; K.cutele2 = (float)P.cutele2;
  bool bad = false;  // a term beyond the LDS sums' range (or a NaN)

  for (int it = beg + t; it < end; it += RS_THREADS) {
    // the mask walk starts at a member given by the item, so the order of
    // an item's sums does not depend on the lane or chunk that runs it
    const int rot = it & 63;
    const WItem item = items[it];
    const int js = item.j;
    const float4 jA = wpos[3 * js], jB = wpos[3 * js + 1], jC = wpos[3 * js + 2];
    int ia, jb, kc;
    symop_of(item.s, &ia, &jb, &kc);
    // partner O translated by -T (the i member carries +T), split hi/lo
    const double ojx = (double)jA.x + (double)jB.x - ia * P.a;
    const double ojy = (double)jA.y + (double)jB.y - jb * P.b;
    const double ojz = (double)jA.z + (double)jB.z - kc * P.c;
    WPartner J;
    J.jhx = (float)ojx; J.jhy = (float)ojy; J.jhz = (float)ojz;
    J.jlx = (float)(ojx - (double)J.jhx); J.jly = (float)(ojy - (double)J.jhy);
    J.jlz = (float)(ojz - (double)J.jhz);
    J.j1x = jA.w; J.j1y = jB.w; J.j1z = jC.x;
    J.j2x = jC.y; J.j2y = jC.z; J.j2z = jC.w;
    float fj[9];
#pragma unroll
    for (int k = 0; k < 9; k++) fj[k] = 0.f;
    float elj = 0.f, eel = 0.f;
    uint64_t big = 0;  // members whose entry is beyond the bit-pattern conversion
    // listwt2 members take part only within relx (pot/watwat_ewald.f:624+):
    // their O-O distances first, in a cheap loop of their own, so that the
    // full entries below run only for members that contribute (a member
    // skipped inside the entry loop would idle its lane for a whole entry)
    uint64_t act = item.m1;
    for (uint64_t m2 = item.m2; m2; m2 &= m2 - 1) {
      const int i = __ffsll((long long)m2) - 1;
      const float4 A = iA[i], B = iB[i];
      const float dx = (A.x - J.jhx) + (B.x - J.jlx);
      const float dy = (A.y - J.jhy) + (B.y - J.jly);
      const float dz = (A.z - J.jhz) + (B.z - J.jlz);
      if (fmaf(dz, dz, fmaf(dy, dy, dx * dx)) <= K.cutele2) act |= 1ull << i;
    }
    uint64_t mask = rotr64(act, rot);
    while (mask) {
      const int i = (__ffsll((long long)mask) - 1 + rot) & 63;
      mask &= mask - 1;
      float fi[9];
      water_entry(iA[i], iB[i], iC[i], J, !((item.m2 >> i) & 1ull), K, fi, fj, elj, eel);
      big |= (uint64_t)add_i_side(fi, i, iacc, icnt) << i;
    }
    if (big) {
      // rare (forces beyond 2^31 kcal/mol/A): these entries recomputed
      // (the same values bit for bit), the bit patterns added above
      // replaced by the exact fixed-point values (range-checked up to
      // FIX_TERM_UNITS); the partner's side and the energies were added in
      // the walk above
      float fjd[9], ed = 0.f;
#pragma unroll
      for (int k = 0; k < 9; k++) fjd[k] = 0.f;
      while (big) {
        const int i = __ffsll((long long)big) - 1;
        big &= big - 1;
        float fi[9];
        water_entry(iA[i], iB[i], iC[i], J, !((item.m2 >> i) & 1ull), K, fi, fjd, ed, ed);
#pragma unroll
        for (int k = 0; k < 9; k++)
          atomicAdd(&iacc[k * WBLOCK + i],
                    (unsigned long long)fixd((double)fi[k], 1048576.0, bad) + kBias32 -
                        (unsigned long long)__double_as_longlong((double)fi[k] + 6442450944.0));
      }
    }
    // the partner's row holds the FP32 sums; the reduction converts them
    // to fixed point (the same conversion of the same value, half the bytes)
    float *jrow = reinterpret_cast<float *>(jpart) + (size_t)jpos[it] * 9;
#pragma unroll
    for (int k = 0; k < 9; k++) jrow[k] = fj[k];
    const int es = item.s == 0 ? 0 : 2;
    eacc[es][t] += (double)elj;
    eacc[es + 1][t] += (double)eel;
  }
  fix_report(bad, P.ovf);
  __syncthreads();
  // the chunk's row in fixed point (member-major as the reduction reads it)
  for (int u = t; u < WBLOCK * 9; u += RS_THREADS) {
    const int i = u / 9, k = u - 9 * i;
    ipart[(size_t)ch * (WBLOCK * 9) + u] =
        (long long)(iacc[k * WBLOCK + i] - (unsigned long long)icnt[i] * kBias32);
  }
  const double v = wg_sum4(eacc[0][t], eacc[1][t], eacc[2][t], eacc[3][t], wsum);
  if (t < 4) epart[kWaterEn[t] * max_parts + ch] = v;
}

__device__ __forceinline__ void water_f32_body(const WaterArgs &A, int bid, int nblk) {
  water_f32_impl(A.chunks, A.order, A.items, A.jpos, (const float4 *)A.wpos, A.nwat, A.P, A.jpart, A.ipart,
                 A.epart, A.max_parts, bid, nblk);
}
DEFINE_KERNEL(k_water_f32, WaterArgs, RS_THREADS, water_f32_body)

// FP64 water kernel (MOIL parity mode)
__device__ __forceinline__ void water_f64_impl(
    const int32_t *__restrict__ chunks, const int32_t *__restrict__ order,
    const WItem *__restrict__ items,
    const int32_t *__restrict__ jpos, const double *__restrict__ wpos, int nwat, const WaterRS &P,
    const double *__restrict__ erf_arr, long long *__restrict__ jpart, long long *__restrict__ ipart,
    double *__restrict__ epart, int max_parts, int bid, int nblk) {
  __shared__ double ipos[WBLOCK * 9];
  __shared__ long long iacc[WBLOCK * 9];
  __shared__ double wsum[16];
  const int ch = order[bid], t = threadIdx.x;  // largest first
  const int blk = chunks[3 * ch], beg = chunks[3 * ch + 1], end = chunks[3 * ch + 2];
  const int s0 = blk * WBLOCK;
  for (int u = t; u < WBLOCK * 9; u += RS_THREADS) {
    ipos[u] = (s0 + u / 9 < nwat) ? wpos[9 * s0 + u] : 0.0;
    iacc[u] = 0;
  }
  __syncthreads();
  const double fs = P.fix_scale;
  bool bad = false;  // a term beyond the fixed-point range (or a NaN)
  double e_lj_r = 0, e_el_r = 0, e_lj_i = 0, e_el_i = 0;
  for (int it = beg + t; it < end; it += RS_THREADS) {
    // the mask walk starts at a member given by the item, so the order of
    // an item's sums does not depend on the lane or chunk that runs it
    const int rot = it & 63;
    const WItem item = items[it];
    const int js = item.j;
    double xj[9];
#pragma unroll
    for (int k = 0; k < 9; k++) xj[k] = wpos[9 * js + k];
    int ia, jb, kc;
    symop_of(item.s, &ia, &jb, &kc);
    const double T[3] = {ia * P.a, jb * P.b, kc * P.c};
    double fj[9];
#pragma unroll
    for (int k = 0; k < 9; k++) fj[k] = 0.0;
    double elj = 0, eel = 0;
    uint64_t mask = rotr64(item.m1 | item.m2, rot);
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
      double fi[9];
#pragma unroll
      for (int k = 0; k < 9; k++) fi[k] = 0.0;
#pragma unroll
      for (int sa = 0; sa < 3; sa++) {
#pragma unroll
        for (int sb = 0; sb < 3; sb++) {
          const double rx = xi[3 * sa] - xj[3 * sb], ry = xi[3 * sa + 1] - xj[3 * sb + 1],
                       rz = xi[3 * sa + 2] - xj[3 * sb + 2];
          const double r2 = rx * rx + ry * ry + rz * rz;
          const double qq = (sa == 0 && sb == 0) ? P.qq_oo : ((sa == 0 || sb == 0) ? P.qq_oh : P.qq_hh);
          const double s2 = 1.0 / r2;
          const double rij = sqrt(r2);
          const double s = 1.0 / rij;
          double df = 0.0;
          if (sa == 0 && sb == 0 && !k2 && r2 <= P.cutvdw2) {
            const double s6 = s2 * s2 * s2;
            const double A12 = P.oo12 * s6 * s6, B6 = P.oo6 * s6;
            const double e1 = A12 - B6;
            elj += e1;
            df = -6.0 * s2 * (A12 + e1);
          }
          double erfcc, derfc;
          erfc_table(erf_arr, P.erftbdns, P.ntab, P.beta * rij, erfcc, derfc);
          eel += qq * s * erfcc;
          df += -qq * s * (s2 * erfcc + s * P.beta * derfc);
          fi[3 * sa] += df * rx;
          fi[3 * sa + 1] += df * ry;
          fi[3 * sa + 2] += df * rz;
          fj[3 * sb] -= df * rx;
          fj[3 * sb + 1] -= df * ry;
          fj[3 * sb + 2] -= df * rz;
        }
      }
#pragma unroll
      for (int k = 0; k < 9; k++) {
        long long v = fixd(fi[k], fs, bad);
        if (v != 0) atomicAdd((unsigned long long *)&iacc[i * 9 + k], (unsigned long long)v);
      }
    }
    long long *jrow = jpart + (size_t)jpos[it] * 9;
#pragma unroll
    for (int k = 0; k < 9; k++) jrow[k] = fixd(fj[k], fs, bad);
    if (item.s == 0) {
      e_lj_r += elj;
      e_el_r += eel;
    } else {
      e_lj_i += elj;
      e_el_i += eel;
    }
  }
  fix_report(bad, P.ovf);
  __syncthreads();
  for (int u = t; u < WBLOCK * 9; u += RS_THREADS) ipart[(size_t)ch * (WBLOCK * 9) + u] = iacc[u];
  const double v = wg_sum4(e_lj_r, e_el_r, e_lj_i, e_el_i, wsum);
  if (t < 4) epart[kWaterEn[t] * max_parts + ch] = v;
}

__device__ __forceinline__ void water_f64_body(const WaterArgs &A, int bid, int nblk) {
  water_f64_impl(A.chunks, A.order, A.items, A.jpos, (const double *)A.wpos, A.nwat, A.P, A.erf, A.jpart,
                 A.ipart, A.epart, A.max_parts, bid, nblk);
}
DEFINE_KERNEL(k_water_f64, WaterArgs, RS_THREADS, water_f64_body)

// one thread per (water slot, gradient component): the slot's partials --
// its jpart rows (contiguous at the slot's CSR range; FP32 values turned
// into 64-bit fixed point here in mixed precision, fixed point already in
// FP64) and its entries in the ipart rows of its block's chunks -- summed
// as integers (order independent) and added to the accumulator.  The
// slot's atoms are touched by no other thread, so no atomics.
__device__ __forceinline__ void water_reduce_impl(
    const int32_t *__restrict__ jstart, const long long *__restrict__ jpart, int jf32,
    const int32_t *__restrict__ cfirst, const long long *__restrict__ ipart,
    const int32_t *__restrict__ slot_o, int nwat, long long *__restrict__ facc, int *ovf,
    int bid) {
  const int g = bid * 256 + threadIdx.x;
  if (g >= 9 * nwat) return;
  const int s = g / 9, k = g - 9 * s;
  long long a0 = 0, a1 = 0;
  const int q0 = jstart[s], q1 = jstart[s + 1];
  int q = q0;
  if (jf32) {
    // four rows in flight per iteration (integer sums: any grouping gives
    // the same total)
    const float *jp = reinterpret_cast<const float *>(jpart);
    bool bad = false;
    long long a2 = 0, a3 = 0;
    for (; q + 3 < q1; q += 4) {
      const float v0 = jp[(size_t)q * 9 + k], v1 = jp[(size_t)(q + 1) * 9 + k],
                  v2 = jp[(size_t)(q + 2) * 9 + k], v3 = jp[(size_t)(q + 3) * 9 + k];
      a0 += fixf(v0, bad);
      a1 += fixf(v1, bad);
      a2 += fixf(v2, bad);
      a3 += fixf(v3, bad);
    }
    for (; q < q1; q++) a0 += fixf(jp[(size_t)q * 9 + k], bad);
    a0 += a2;
    a1 += a3;
    fix_report(bad, ovf);
  } else {
    for (; q + 1 < q1; q += 2) {
      a0 += jpart[(size_t)q * 9 + k];
      a1 += jpart[(size_t)(q + 1) * 9 + k];
    }
    if (q < q1) a0 += jpart[(size_t)q * 9 + k];
  }
  const int b = s / WBLOCK, i = s % WBLOCK;
  for (int c = cfirst[b]; c < cfirst[b + 1]; c++) a1 += ipart[(size_t)c * (WBLOCK * 9) + i * 9 + k];
  facc[3 * (size_t)slot_o[s] + k] += a0 + a1;
}
__device__ __forceinline__ void water_reduce_body(const WReduceArgs &A, int bid, int nblk) {
  water_reduce_impl(A.jstart, A.jpart, A.jf32, A.cfirst, A.ipart, A.slot_o, A.nwat, A.facc,
                    A.ovf, bid);
}
DEFINE_KERNEL(k_water_reduce, WReduceArgs, 256, water_reduce_body)

// ---------------------------------------------------------------------------
// the water real-space launches of one evaluation: pack, pair kernel,
// reduction (none with fewer than two waters or an empty list)
void water_plan(const moil_ctx *c, const double *d_coor, EvalPlan *p) {
  p->n_pack = p->n_water = p->n_wred = 0;
  if (c->nwat < 2 || c->nwchunk == 0) return;
  p->pack.coor = d_coor;
  p->pack.slot_o = c->d_slot_o;
  p->pack.nwat = c->nwat;
  p->pack.wpos = c->d_wpos;
  p->n_pack = (c->nwat + RS_THREADS - 1) / RS_THREADS;
  WaterRS P;
  P.qq_oo = c->qq_oo; P.qq_oh = c->qq_oh; P.qq_hh = c->qq_hh;
  P.oo12 = c->oo12; P.oo6 = c->oo6;
  P.beta = c->beta; P.cutvdw2 = c->par.cutvdw2; P.cutele2 = c->par.cutele2;
  P.erftbdns = c->erftbdns; P.a = c->par.a; P.b = c->par.b; P.c = c->par.c;
  P.fix_scale = c->fix_scale; P.ntab = c->ntab;
  P.ovf = c->d_ovf;
  WaterArgs &W = p->water;
  W.chunks = c->d_wchunk;
  W.order = c->d_worder;
  W.items = c->d_witems;
  W.jpos = c->d_jpos;
  W.wpos = c->d_wpos;
  W.nwat = c->nwat;
  W.P = P;
  W.erf = c->d_erf;
  W.jpart = c->d_jpart;
  W.ipart = c->d_ipart;
  W.epart = c->d_epart;
  W.max_parts = c->max_parts;
  p->n_water = c->nwchunk;
  WReduceArgs &R = p->wred;
  R.jstart = c->d_jstart;
  R.jpart = c->d_jpart;
  R.jf32 = c->par.precision == MOIL_PREC_MIXED;
  R.cfirst = c->d_cfirst;
  R.ipart = c->d_ipart;
  R.slot_o = c->d_slot_o;
  R.nwat = c->nwat;
  R.facc = c->d_facc;
  R.ovf = c->d_ovf;
  p->n_wred = (9 * c->nwat + 255) / 256;
}

// Single-system form on a short grid (mixed precision): fewer workgroups
// than chunks -- a fixed number per CU, below the CU's capacity for this
// kernel (six at its budget of 80 VGPRs, which leaves room on every SIMD
// for the reciprocal-space kernels' waves; the few values it spills are
// reloaded once per chunk, outside the entry loop) -- walk the chunks in
// rounds of
// gridDim.x, alternating direction every round (the chunks are in
// dispatch order, largest first, so a workgroup that takes a large chunk
// in one round takes a small one in the next).  The slots left free on
// every CU let the reciprocal-space kernels of the same evaluation, on
// the second stream, run beside it instead of after it.  The chunk index
// depends only on blockIdx.x, gridDim.x and the round, so the loop and
// its barriers are uniform over the workgroup.  Which workgroup runs a
// chunk does not matter: each chunk writes its own partial rows and
// energy partials.
__global__ __launch_bounds__(RS_THREADS) __attribute__((amdgpu_waves_per_eu(6)))
void k_water_rounds(WaterArgs A, int nchunk) {
  const int G = gridDim.x, b = blockIdx.x;
  for (int r = 0; r * G < nchunk; r++) {
    const int k = r * G + ((r & 1) ? G - 1 - b : b);
    if (k < nchunk) water_f32_body(A, k, nchunk);
  }
}

// launched through hipExtLaunchKernelGGL so that optional start/stop
// events are recorded by the dispatch itself; mixed precision on nwg
// workgroups (k_water_rounds) when 0 < nwg < nblk, else one workgroup per
// chunk
void launch_k_water_ev(const WaterArgs &a, bool f64, int nblk, hipStream_t st, hipEvent_t e0,
                       hipEvent_t e1, int nwg) {
  if (nblk <= 0) return;
  if (!f64 && nwg > 0 && nwg < nblk) {
    hipExtLaunchKernelGGL(k_water_rounds, dim3(nwg), dim3(RS_THREADS), 0, st, e0, e1, 0, a, nblk);
    return;
  }
  if (f64)
    hipExtLaunchKernelGGL(k_water_f64, dim3(nblk), dim3(RS_THREADS), 0, st, e0, e1, 0, a);
  else
    hipExtLaunchKernelGGL(k_water_f32, dim3(nblk), dim3(RS_THREADS), 0, st, e0, e1, 0, a);
}
