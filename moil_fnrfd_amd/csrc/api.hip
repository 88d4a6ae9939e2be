// C ABI of the MI355X MOIL energy library (declared in include/moil_gpu.h):
// context set-up (the ewald_init equivalent), squeeze, list build and the
// eforce step that chains the bonded, PME and real-space kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "ewald_setup.h"
#include "multi.h"

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

extern "C" int32_t moil_gpu_abi_version(void) { return MOIL_GPU_ABI_VERSION; }
extern "C" const char *moil_gpu_last_error(void) { return g_err.c_str(); }


#define CHECK_RC(x)          \
  do {                       \
    int32_t r_ = (x);        \
    if (r_ != MOIL_GPU_OK) return r_; \
  } while (0)

static int32_t fail(int32_t code, const std::string &msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// ewald_init (symm/ewald_init.f:8-242) on the host, FP64.

template <typename T>
static int32_t upload(T **dst, const T *src, size_t n) {
  size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
  HIP_CHECK(hipMalloc((void **)dst, bytes));
  if (n > 0 && src) HIP_CHECK(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
  return MOIL_GPU_OK;
}

void egraph_drop(moil_ctx *c);
void virial_free(moil_ctx *c);

static void ctx_free(moil_ctx *c) {
  if (!c) return;
  egraph_drop(c);
  virial_free(c);
  lists_free(c);
  atoms_free(c);
  pme_free(c);
  void *ptrs[] = {c->d_poipt, c->d_monoflag, c->d_flagchr, c->d_ptchg, c->d_eps6, c->d_eps12,
                  c->d_exc_i, c->d_exc_j, c->d_ib, c->d_bpar, c->d_ia, c->d_apar, c->d_it,
                  c->d_tpar, c->d_ii, c->d_ipar, c->d_sp, c->d_p14, c->d_wat_o, c->d_wat_mono,
                  c->d_erf, c->d_facc, c->d_epart, c->d_energy, c->d_coor, c->d_dpot, c->d_wpos,
                  c->d_wi_boff, c->d_wi_aoff, c->d_wi_eoff, c->d_wi_zero, c->d_wi_bidx,
                  c->d_wi_aidx, c->d_wi_eidx, c->d_wi_bpar, c->d_wi_apar, c->d_etot, c->d_ticket,
                  c->d_xall, c->d_adj, c->d_scan_ws, c->d_brstart, c->d_brslot, c->d_bslot};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (c->gev_in) hipEventDestroy(c->gev_in);
  if (c->gev_out) hipEventDestroy(c->gev_out);
  if (c->gstream) hipStreamDestroy(c->gstream);
  for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->ev_intra) hipEventDestroy(c->ev_intra);
  if (c->stream2) hipStreamDestroy(c->stream2);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

static int32_t ensure_parts(moil_ctx *c, int need) {
  if (need <= c->max_parts && c->d_epart) return MOIL_GPU_OK;
  if (c->d_epart) HIP_CHECK(hipFree(c->d_epart));
  c->max_parts = std::max(need, 64);
  HIP_CHECK(hipMalloc(&c->d_epart, sizeof(double) * EN_COUNT * c->max_parts));
  return MOIL_GPU_OK;
}

// Every index array of the topology against the sizes it indexes, before
// any of it is used: atom numbers in 1..npt, monomer numbers in 1..totmon,
// the monomer and exclusion CSR pointers non-decreasing and in range.  A
// .con record with a blank index field (read as 0) ends here as EINVAL
// instead of as an access outside an array, on the host or the device.
// The exclusion pointers are returned as a complete CSR in exc1n: MOIL's
// rconn sets exc1 only up to the last atom that has an exclusion record
// (generic/rconn.f:313-338), so in MOIL's own COMMON the entries after it
// are still 0; they are read as empty ranges, as MOIL's loops over
// exc1(i-1)+1..exc1(i) read them.
static int32_t check_topology(const moil_topology *t, std::vector<int32_t> &exc1n) {
  const int32_t npt = t->npt, nmo = t->totmon;
  if (npt <= 0 || nmo <= 0) return fail(MOIL_GPU_EINVAL, "empty system");
  if (t->nb < 0 || t->nangl < 0 || t->ntors < 0 || t->nimp < 0 || t->totspe < 0 ||
      t->nwaters < 0)
    return fail(MOIL_GPU_EINVAL, "negative term count");
  if (!t->poipt || !t->monoflag || !t->ptchg || !t->epsgm6 || !t->epsgm12 || !t->flagchr ||
      !t->exc1)
    return fail(MOIL_GPU_EINVAL, "null topology array");
  if (t->poipt[0] != 0 || t->poipt[nmo] != npt)
    return fail(MOIL_GPU_EINVAL, "poipt must start at 0 and end at npt");
  for (int32_t m = 0; m < nmo; m++)
    if (t->poipt[m + 1] < t->poipt[m])
      return fail(MOIL_GPU_EINVAL, "poipt decreases at monomer " + std::to_string(m + 1));
  if (t->exc1[0] != 0) return fail(MOIL_GPU_EINVAL, "exc1 must start at 0");
  exc1n.assign((size_t)npt + 1, 0);
  {
    int32_t top = 0;
    bool tail = false;
    for (int32_t a = 1; a <= npt; a++) {
      const int32_t v = t->exc1[a];
      if (!tail && v >= top) {
        top = v;
      } else if (v == 0) {
        tail = true;
      } else {
        return fail(MOIL_GPU_EINVAL, "exc1 decreases at atom " + std::to_string(a));
      }
      exc1n[a] = top;
    }
  }
  const int64_t nexc = exc1n[npt];
  if (nexc > 0 && !t->exc2) return fail(MOIL_GPU_EINVAL, "null exc2");
  for (int64_t k = 0; k < nexc; k++)
    if (t->exc2[k] < 1 || t->exc2[k] > npt)
      return fail(MOIL_GPU_EINVAL, "exclusion partner " + std::to_string(t->exc2[k]) +
                                       " outside 1.." + std::to_string(npt));
  struct Idx {
    const char *name;
    const int32_t *v;
    int32_t n;
  };
  const Idx ids[] = {{"ib1", t->ib1, t->nb},         {"ib2", t->ib2, t->nb},
                     {"iangl1", t->iangl1, t->nangl}, {"iangl2", t->iangl2, t->nangl},
                     {"iangl3", t->iangl3, t->nangl}, {"itor1", t->itor1, t->ntors},
                     {"itor2", t->itor2, t->ntors},   {"itor3", t->itor3, t->ntors},
                     {"itor4", t->itor4, t->ntors},   {"iimp1", t->iimp1, t->nimp},
                     {"iimp2", t->iimp2, t->nimp},    {"iimp3", t->iimp3, t->nimp},
                     {"iimp4", t->iimp4, t->nimp},    {"spec1", t->spec1, t->totspe},
                     {"spec2", t->spec2, t->totspe}};
  for (const Idx &x : ids) {
    if (x.n > 0 && !x.v) return fail(MOIL_GPU_EINVAL, std::string("null ") + x.name);
    for (int32_t m = 0; m < x.n; m++)
      if (x.v[m] < 1 || x.v[m] > npt)
        return fail(MOIL_GPU_EINVAL, std::string(x.name) + "(" + std::to_string(m + 1) +
                                         ") = " + std::to_string(x.v[m]) + " outside 1.." +
                                         std::to_string(npt));
  }
  const void *par[] = {t->kbond, t->req, t->kangl, t->angleq, t->ktors1, t->ktors2,
                       t->ktors3, t->phase1, t->phase2, t->phase3, t->kimp, t->impeq, t->p14};
  const int32_t npar[] = {t->nb, t->nb, t->nangl, t->nangl, t->ntors, t->ntors, t->ntors,
                          t->ntors, t->ntors, t->ntors, t->nimp, t->nimp, t->totspe};
  for (int q = 0; q < 13; q++)
    if (npar[q] > 0 && !par[q]) return fail(MOIL_GPU_EINVAL, "null term parameter array");
  if (t->nwaters > 0 && !t->idxtip3) return fail(MOIL_GPU_EINVAL, "null idxtip3");
  for (int32_t w = 0; w < t->nwaters; w++)
    if (t->idxtip3[w] < 1 || t->idxtip3[w] > nmo)
      return fail(MOIL_GPU_EINVAL, "idxtip3(" + std::to_string(w + 1) + ") = " +
                                       std::to_string(t->idxtip3[w]) + " outside 1.." +
                                       std::to_string(nmo));
  return MOIL_GPU_OK;
}

// O-O LJ coefficients of the water pair kernels, pot/watwat_ewald.f:62-68:
// geometric A = pick e12^2, B = pick e6^2; arith A = pick e12^2 sigma^12,
// B = pick e12^2 sigma^6 (the integer powers by repeated squaring)
static void water_oo_lj(double pick, double e12, double e6, bool arith, double &oo12,
                        double &oo6) {
  if (!arith) {
    oo6 = pick * e6 * e6;
    oo12 = pick * e12 * e12;
    return;
  }
  const double s2 = e6 * e6, s4 = s2 * s2, s8 = s4 * s4;
  oo6 = pick * (e12 * e12) * (s2 * s4);
  oo12 = pick * (e12 * e12) * (s4 * s8);
}

extern "C" int32_t moil_gpu_init(const moil_topology *t, const moil_params *p, moil_ctx **out) {
  if (!t || !p || !out) return fail(MOIL_GPU_EINVAL, "null argument");
  *out = nullptr;
  std::vector<int32_t> exc1n;
  CHECK_RC(check_topology(t, exc1n));
  moil_topology tn = *t;  // the same arrays, the exclusion pointers complete
  tn.exc1 = exc1n.data();
  t = &tn;
  if (p->order != 4) return fail(MOIL_GPU_EINVAL, "PME interpolation order must be 4");
  if (!(p->a > 0 && p->b > 0 && p->c > 0)) return fail(MOIL_GPU_EINVAL, "box edges must be > 0");
  if (p->nfft1 < 4 || p->nfft2 < 4 || p->nfft3 < 4) return fail(MOIL_GPU_EINVAL, "PME grid too small");
  if (p->precision != MOIL_PREC_FP64 && p->precision != MOIL_PREC_MIXED)
    return fail(MOIL_GPU_EINVAL, "unknown precision mode");
  if (p->cutele2 <= 0 || p->eps <= 0 || p->dtol <= 0) return fail(MOIL_GPU_EINVAL, "bad cutoffs/eps/dtol");
  if (p->terms & ~(MOIL_TERM_DEFAULT | MOIL_TERM_ARITH))
    return fail(MOIL_GPU_EINVAL, "unknown term bits");
  // the arith torsion form exists for periods 1..4 only (pot/ephi.f:113-138)
  if ((p->terms & MOIL_TERM_ARITH) && (p->terms & MOIL_TERM_TORS))
    for (int32_t m = 0; m < t->ntors; m++)
      if (t->period[m] < 1 || t->period[m] > 4)
        return fail(MOIL_GPU_EINVAL, "torsion " + std::to_string(m + 1) + " has period " +
                                         std::to_string(t->period[m]) +
                                         ": the arith torsion form needs 1..4");
  // A list reach above half a box edge makes MOIL's half-shell images
  // double count pairs (symm/syminit.f:20-21); MOIL does not refuse it, and
  // neither does this library: the pair set is whatever MOIL's predicates
  // give.
  double qtot = 0;
  for (int i = 0; i < t->npt; i++) qtot += t->ptchg[i];
  if (std::fabs(qtot) > 1e-7)  // symm/ewald_init.f:62-72
    return fail(MOIL_GPU_ECHARGE, "The charge of the system is " + std::to_string(qtot) +
                                      ": add counterions and try again");
  {
    std::string why;
    if (!pme_charges_fit(t->ptchg, (size_t)t->npt, p->eps, &why)) return fail(MOIL_GPU_EINVAL, why);
  }

  moil_ctx *c = new moil_ctx();
  c->par = *p;
  c->npt = t->npt;
  c->totmon = t->totmon;
  c->nb = t->nb; c->nangl = t->nangl; c->ntors = t->ntors; c->nimp = t->nimp;
  c->totspe = t->totspe;
  c->fix_shift = (p->precision == MOIL_PREC_MIXED) ? FIX_SHIFT_MIXED : FIX_SHIFT_FP64;
  c->fix_scale = std::ldexp(1.0, c->fix_shift);
  c->fix_inv = std::ldexp(1.0, -c->fix_shift);
  int32_t rc;
#define TRY(x)                 \
  do {                         \
    rc = (x);                  \
    if (rc != MOIL_GPU_OK) {   \
      ctx_free(c);             \
      return rc;               \
    }                          \
  } while (0)
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    ctx_free(c);
    return fail(MOIL_GPU_EHIP, "no HIP device available");
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->gev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->gev_out, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_intra, hipEventDisableTiming) != hipSuccess) {
    ctx_free(c);
    return fail(MOIL_GPU_EHIP, "hipStreamCreate/hipEventCreate failed");
  }
  {
    const char *e = std::getenv("MOIL_GPU_PME_CONCURRENT");
    c->pme_concurrent = !(e && e[0] == '0');
    const char *gr = std::getenv("MOIL_GPU_GRAPH");
    c->use_graph = gr && gr[0] == '1';
  }
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      ctx_free(c);
      return fail(MOIL_GPU_EHIP, "device query failed");
    }
    c->device = dev;
    c->ncu = ncu;
    // the device address of the no-allocation evaluation's scratch on this
    // device (looked up once per device)
    TRY(scratch_attach(dev));
    const char *w = std::getenv("MOIL_GPU_WATER_WG_PER_CU");
    if (w && w[0]) c->water_wg_per_cu = std::max(0, std::atoi(w));
    const char *cp = std::getenv("MOIL_GPU_CHUNK_PAIRS");
    if (cp && cp[0]) c->wchunk_pairs = std::max(256, std::atoi(cp));
    const char *ap = std::getenv("MOIL_GPU_ATOM_CHUNK_PAIRS");
    if (ap && ap[0]) c->achunk_pairs = std::max(256, std::atoi(ap));
    c->water_wgs = c->water_wg_per_cu * ncu;
  }
  // ---- Ewald set-up ----
  double cutoff = std::sqrt(p->cutele2);
  c->beta = moil_ewald_beta(cutoff, moil_ewald_dtol(p->dtol));
  c->erftbdns = 5000.0;  // 0.5*perft, symm/ewald_init.f:103-104
  c->ntab = (int)(c->beta * c->erftbdns * cutoff * 1.5);
  moil_erfc_table(c->erftbdns, c->ntab, c->h_erf);
  {
    double pi = 4.0 * std::atan(1.0), spi = std::sqrt(1.0 / pi), s = 0;
    for (int i = 0; i < t->npt; i++) s += t->ptchg[i] * t->ptchg[i];
    c->e_self = -(c->beta * spi * (KOFDIE / p->eps) * s);  // symm/ewald_init.f:131-150
  }
  TRY(upload(&c->d_erf, c->h_erf.data(), c->h_erf.size()));
  // ---- topology ----
  TRY(upload(&c->d_poipt, t->poipt, (size_t)t->totmon + 1));
  TRY(upload(&c->d_monoflag, t->monoflag, (size_t)t->totmon));
  TRY(upload(&c->d_flagchr, t->flagchr, (size_t)t->npt));
  TRY(upload(&c->d_ptchg, t->ptchg, (size_t)t->npt));
  c->h_q.assign(t->ptchg, t->ptchg + t->npt);
  c->h_e12.assign(t->epsgm12, t->epsgm12 + t->npt);
  c->h_e6.assign(t->epsgm6, t->epsgm6 + t->npt);
  TRY(upload(&c->d_eps6, t->epsgm6, (size_t)t->npt));
  TRY(upload(&c->d_eps12, t->epsgm12, (size_t)t->npt));
  // ---- waters (monomers named TIP3/SPC*) ----
  for (int m = 0; m < t->totmon; m++) {
    if (t->monoflag[m] & MOIL_MONO_WATER) {
      int nat = t->poipt[m + 1] - t->poipt[m];
      if (nat != 3) {
        ctx_free(c);
        return fail(MOIL_GPU_EINVAL, "water monomer without exactly 3 atoms (O, H, H)");
      }
      c->h_wat_mono.push_back(m + 1);
      c->h_wat_o.push_back(t->poipt[m + 1] - 3);  // pot/watwat_ewald.f:103
    }
  }
  c->nwat = (int)c->h_wat_mono.size();
  if (c->nwat != t->nwaters) {
    ctx_free(c);
    return fail(MOIL_GPU_EINVAL, "nwaters does not match the number of TIP3/SPC monomers");
  }
  for (int w = 0; w < t->nwaters; w++)
    if (!(t->monoflag[t->idxtip3[w] - 1] & MOIL_MONO_WATER)) {
      ctx_free(c);
      return fail(MOIL_GPU_EINVAL, "idxtip3(" + std::to_string(w + 1) +
                                       ") is not a TIP3/SPC monomer");
    }
  // MOIL's water routines assume the three intra-water pairs are excluded
  // (the atom lists then never hold an intra-water pair)
  for (int w = 0; w < c->nwat; w++) {
    const int o = c->h_wat_o[w];
    int found = 0;
    for (int k = t->exc1[o]; k < t->exc1[o + 1]; k++)
      found += (t->exc2[k] - 1 == o + 1) + (t->exc2[k] - 1 == o + 2);
    for (int k = t->exc1[o + 1]; k < t->exc1[o + 2]; k++) found += (t->exc2[k] - 1 == o + 2);
    if (found != 3) {
      ctx_free(c);
      return fail(MOIL_GPU_EINVAL, "water monomer whose O-H, O-H and H-H pairs are not all excluded");
    }
  }
  // water index of every atom (-1 outside waters) for the intra-water split
  std::vector<int32_t> wat_of(t->npt, -1);
  for (int w = 0; w < c->nwat; w++)
    for (int k = 0; k < 3; k++) wat_of[c->h_wat_o[w] + k] = w;
  auto same_water = [&](std::initializer_list<int> atoms0) -> int {
    int w = -2;
    for (int a : atoms0) {
      int x = wat_of[a];
      if (x < 0 || (w != -2 && x != w)) return -1;
      w = x;
    }
    return w;
  };
  // per-water CSR of intra-water terms (original order kept within a water)
  std::vector<std::vector<int>> wb(c->nwat), wa(c->nwat), we(c->nwat);
  std::vector<int32_t> wi_eidx;
  std::vector<int32_t> ei, ej;  // generic exclusion pairs (0-based)
  {
    std::vector<std::pair<int, int>> all;
    for (int i = 1; i <= t->npt - 1; i++)  // pot/cdie_ewald.f:68,83 (atom npt ignored)
      for (int k = t->exc1[i - 1]; k < t->exc1[i]; k++) all.push_back({i - 1, t->exc2[k] - 1});
    c->nexcl = (int)all.size();
    {
      std::vector<int32_t> xp(2 * all.size());
      for (size_t m = 0; m < all.size(); m++) {
        xp[2 * m] = all[m].first;
        xp[2 * m + 1] = all[m].second;
      }
      c->nxall = (int)all.size();
      TRY(upload(&c->d_xall, xp.data(), xp.size()));
    }
    for (size_t m = 0; m < all.size(); m++) {
      int w = same_water({all[m].first, all[m].second});
      if (w >= 0) {
        we[w].push_back((int)m);
      } else {
        ei.push_back(all[m].first);
        ej.push_back(all[m].second);
      }
    }
    c->gnexcl = (int)ei.size();
    TRY(upload(&c->d_exc_i, ei.data(), ei.size()));
    TRY(upload(&c->d_exc_j, ej.data(), ej.size()));
    for (int w = 0; w < c->nwat; w++)
      for (int m : we[w]) {
        wi_eidx.push_back(all[m].first - c->h_wat_o[w]);
        wi_eidx.push_back(all[m].second - c->h_wat_o[w]);
      }
  }
  {
    std::vector<int32_t> ib, wi_bidx;
    std::vector<double> bp, wi_bpar;
    for (int m = 0; m < c->nb; m++) {
      int w = same_water({t->ib1[m] - 1, t->ib2[m] - 1});
      if (w >= 0) {
        wb[w].push_back(m);
      } else {
        ib.push_back(t->ib1[m]); ib.push_back(t->ib2[m]);
        bp.push_back(t->kbond[m]); bp.push_back(t->req[m]);
      }
    }
    c->gnb = (int)ib.size() / 2;
    TRY(upload(&c->d_ib, ib.data(), ib.size()));
    TRY(upload(&c->d_bpar, bp.data(), bp.size()));
    std::vector<int32_t> ia, wi_aidx;
    std::vector<double> ap, wi_apar;
    for (int m = 0; m < c->nangl; m++) {
      int w = same_water({t->iangl1[m] - 1, t->iangl2[m] - 1, t->iangl3[m] - 1});
      if (w >= 0) {
        wa[w].push_back(m);
      } else {
        ia.push_back(t->iangl1[m]); ia.push_back(t->iangl2[m]); ia.push_back(t->iangl3[m]);
        ap.push_back(t->kangl[m]); ap.push_back(t->angleq[m]);
      }
    }
    c->gnangl = (int)ia.size() / 3;
    TRY(upload(&c->d_ia, ia.data(), ia.size()));
    TRY(upload(&c->d_apar, ap.data(), ap.size()));
    // intra-water CSR
    std::vector<int32_t> boff(c->nwat + 1, 0), aoff(c->nwat + 1, 0), eoff(c->nwat + 1, 0);
    for (int w = 0; w < c->nwat; w++) {
      const int a0 = c->h_wat_o[w];
      for (int m : wb[w]) {
        wi_bidx.push_back(t->ib1[m] - 1 - a0); wi_bidx.push_back(t->ib2[m] - 1 - a0);
        wi_bpar.push_back(t->kbond[m]); wi_bpar.push_back(t->req[m]);
      }
      for (int m : wa[w]) {
        wi_aidx.push_back(t->iangl1[m] - 1 - a0); wi_aidx.push_back(t->iangl2[m] - 1 - a0);
        wi_aidx.push_back(t->iangl3[m] - 1 - a0);
        wi_apar.push_back(t->kangl[m]); wi_apar.push_back(t->angleq[m]);
      }
      boff[w + 1] = boff[w] + (int)wb[w].size();
      aoff[w + 1] = aoff[w] + (int)wa[w].size();
      eoff[w + 1] = eoff[w] + (int)we[w].size();
    }
    c->wi_nterms = boff[c->nwat] + aoff[c->nwat] + eoff[c->nwat];
    std::vector<int32_t> zero(c->nwat + 1, 0);
    TRY(upload(&c->d_wi_boff, boff.data(), boff.size()));
    TRY(upload(&c->d_wi_aoff, aoff.data(), aoff.size()));
    TRY(upload(&c->d_wi_eoff, eoff.data(), eoff.size()));
    TRY(upload(&c->d_wi_zero, zero.data(), zero.size()));
    TRY(upload(&c->d_wi_bidx, wi_bidx.data(), wi_bidx.size()));
    TRY(upload(&c->d_wi_aidx, wi_aidx.data(), wi_aidx.size()));
    TRY(upload(&c->d_wi_eidx, wi_eidx.data(), wi_eidx.size()));
    TRY(upload(&c->d_wi_bpar, wi_bpar.data(), wi_bpar.size()));
    TRY(upload(&c->d_wi_apar, wi_apar.data(), wi_apar.size()));
    std::vector<int32_t> it(4 * (size_t)c->ntors);
    std::vector<double> tp(6 * (size_t)c->ntors);
    const bool arith = (p->terms & MOIL_TERM_ARITH) != 0;
    for (int m = 0; m < c->ntors; m++) {
      it[4 * m] = t->itor1[m]; it[4 * m + 1] = t->itor2[m];
      it[4 * m + 2] = t->itor3[m]; it[4 * m + 3] = t->itor4[m];
      tp[6 * m] = t->ktors1[m]; tp[6 * m + 1] = t->ktors2[m]; tp[6 * m + 2] = t->ktors3[m];
      tp[6 * m + 3] = t->phase1[m]; tp[6 * m + 4] = t->phase2[m]; tp[6 * m + 5] = t->phase3[m];
      // the arith form reads k2, phase2 and the period, kept in k1's place
      if (arith) tp[6 * m] = (double)t->period[m];
    }
    TRY(upload(&c->d_it, it.data(), it.size()));
    TRY(upload(&c->d_tpar, tp.data(), tp.size()));
    std::vector<int32_t> ii(4 * (size_t)c->nimp);
    std::vector<double> ip(2 * (size_t)c->nimp);
    for (int m = 0; m < c->nimp; m++) {
      ii[4 * m] = t->iimp1[m]; ii[4 * m + 1] = t->iimp2[m];
      ii[4 * m + 2] = t->iimp3[m]; ii[4 * m + 3] = t->iimp4[m];
      if (arith) {  // the dihedral of (iimp4, iimp2, iimp1, iimp3), pot/eimphi.f:110-115
        ii[4 * m] = t->iimp4[m];
        ii[4 * m + 2] = t->iimp1[m];
        ii[4 * m + 3] = t->iimp3[m];
      }
      ip[2 * m] = t->kimp[m]; ip[2 * m + 1] = t->impeq[m];
    }
    TRY(upload(&c->d_ii, ii.data(), ii.size()));
    TRY(upload(&c->d_ipar, ip.data(), ip.size()));
    std::vector<int32_t> sp(2 * (size_t)c->totspe);
    for (int m = 0; m < c->totspe; m++) {
      sp[2 * m] = t->spec1[m]; sp[2 * m + 1] = t->spec2[m];
    }
    TRY(upload(&c->d_sp, sp.data(), sp.size()));
    c->h_sp = sp;
    TRY(upload(&c->d_p14, t->p14, 3 * (size_t)c->totspe));
    // gradient slots of the bonded terms and each atom's list of them:
    // the terms the bonded kernel evaluates (the enabled types), in its
    // flat order, one slot per atom position (arities 2, 3, 4, 4, 2, 2)
    {
      const uint32_t T = p->terms;
      const bool nbon = (T & (MOIL_TERM_VDW | MOIL_TERM_EL)) != 0;
      std::vector<int32_t> slot_atom;  // atom (0-based) of every slot
      if (T & MOIL_TERM_BOND)
        for (int32_t a : ib) slot_atom.push_back(a - 1);
      if (T & MOIL_TERM_ANGLE)
        for (int32_t a : ia) slot_atom.push_back(a - 1);
      if (T & MOIL_TERM_TORS)
        for (int32_t a : it) slot_atom.push_back(a - 1);
      if (T & MOIL_TERM_IMP)
        for (int32_t a : ii) slot_atom.push_back(a - 1);
      if (nbon)
        for (size_t m = 0; m < ei.size(); m++) {
          slot_atom.push_back(ei[m]);
          slot_atom.push_back(ej[m]);
        }
      if (nbon && (T & (MOIL_TERM_EL14 | MOIL_TERM_VDW14)))
        for (int32_t a : sp) slot_atom.push_back(a - 1);
      const size_t ns = slot_atom.size();
      std::vector<int32_t> rstart(c->npt + 1, 0), rslot(std::max<size_t>(ns, 1));
      for (size_t q = 0; q < ns; q++) rstart[slot_atom[q] + 1]++;
      for (int a = 0; a < c->npt; a++) rstart[a + 1] += rstart[a];
      std::vector<int32_t> fill(rstart.begin(), rstart.end() - 1);
      for (size_t q = 0; q < ns; q++) rslot[fill[slot_atom[q]]++] = (int32_t)q;
      c->nbslot = ns;
      TRY(upload(&c->d_brstart, rstart.data(), rstart.size()));
      TRY(upload(&c->d_brslot, rslot.data(), rslot.size()));
      if (hipMalloc(&c->d_bslot, sizeof(long long) * 3 * std::max<size_t>(ns, 1)) != hipSuccess) {
        ctx_free(c);
        return fail(MOIL_GPU_EHIP, "hipMalloc of the bonded gradient slots failed");
      }
    }
  }
  if (c->nwat >= 2) {
    // parameters of the last water, pot/watwat_ewald.f:40-72
    int m = t->idxtip3[t->nwaters - 1];
    int io = t->poipt[m] - 2 - 1, ih = t->poipt[m] - 1;
    double epst = (p->terms & MOIL_TERM_EL) ? 1.0 : 0.0;
    double pick = (p->terms & MOIL_TERM_VDW) ? 1.0 : 0.0;
    water_oo_lj(pick, t->epsgm12[io], t->epsgm6[io], (p->terms & MOIL_TERM_ARITH) != 0, c->oo12,
                c->oo6);
    c->qq_oo = epst * KOFDIE / p->eps * t->ptchg[io] * t->ptchg[io];
    c->qq_oh = epst * KOFDIE / p->eps * t->ptchg[io] * t->ptchg[ih];
    c->qq_hh = epst * KOFDIE / p->eps * t->ptchg[ih] * t->ptchg[ih];
  }
  if (p->precision == MOIL_PREC_FP64) {
    // the FP64 pair kernels read MOIL's erfc table (1.5 relx long,
    // symm/ewald_init.f:107-126) at every listed distance: pairs reach
    // max(rvbg, rebg), water site pairs up to ~2 A more
    const double reach = std::sqrt(std::max(p->cutvbig2, p->cutebig2)) + (c->nwat > 0 ? 2.0 : 0.0);
    if (c->beta * c->erftbdns * reach >= (double)(c->ntab - 1)) {
      ctx_free(c);
      return fail(MOIL_GPU_EINVAL, "list cutoffs reach beyond MOIL's erfc table (1.5 relx)");
    }
  }
  TRY(upload(&c->d_wat_mono, c->h_wat_mono.data(), c->h_wat_mono.size()));
  TRY(upload(&c->d_wat_o, c->h_wat_o.data(), c->h_wat_o.size()));
  // ---- buffers ----
  auto mal = [&](void **ptr, size_t bytes) -> int32_t {
    HIP_CHECK(hipMalloc(ptr, std::max<size_t>(bytes, 16)));
    return MOIL_GPU_OK;
  };
  TRY(mal((void **)&c->d_facc, sizeof(long long) * 3 * c->npt));
  if (hipMemset(c->d_facc, 0, sizeof(long long) * 3 * c->npt) != hipSuccess) {
    ctx_free(c);
    return fail(MOIL_GPU_EHIP, "hipMemset failed");
  }
  TRY(mal((void **)&c->d_energy, sizeof(double) * 32));
  TRY(mal((void **)&c->d_etot, sizeof(double) * EN_COUNT));
  TRY(mal((void **)&c->d_ticket, sizeof(unsigned int) * 4));
  c->d_ovf = (int *)(c->d_ticket + 1);
  if (hipMemset(c->d_ticket, 0, sizeof(unsigned int) * 4) != hipSuccess) {
    ctx_free(c);
    return fail(MOIL_GPU_EHIP, "hipMemset failed");
  }
  TRY(mal((void **)&c->d_coor, sizeof(double) * 3 * c->npt));
  TRY(mal((void **)&c->d_dpot, sizeof(double) * 3 * c->npt));
  TRY(mal(&c->d_wpos, (p->precision == MOIL_PREC_MIXED ? 48 : 72) * (size_t)std::max(c->nwat, 1)));
  TRY(lists_init(c));
  TRY(atoms_init(c, t));
  TRY(pme_init(c));
  TRY(ensure_parts(c, std::max(pme_conv_blocks(c), bonded_max_blocks(c))));
#undef TRY
  *out = c;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_finalize(moil_ctx *c) {
  if (c) {
    hipStreamSynchronize(c->stream);
    ctx_free(c);
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_ewald_info(moil_ctx *c, double *ewaldcof, double *e_self,
                                       int32_t *ntab, double *erftbdns) {
  if (!c) return fail(MOIL_GPU_EINVAL, "null context");
  if (ewaldcof) *ewaldcof = c->beta;
  if (e_self) *e_self = c->e_self;
  if (ntab) *ntab = c->ntab;
  if (erftbdns) *erftbdns = c->erftbdns;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_ewald_setup(double cutele2, double dtol, double *ewaldcof,
                                    int32_t *ntab, double *erftbdns, double *erf_arr,
                                    int32_t cap_rows) {
  if (!(cutele2 > 0) || !(dtol > 0)) return fail(MOIL_GPU_EINVAL, "bad cutoff or dtol");
  const double cutoff = std::sqrt(cutele2), dens = 5000.0;
  const double beta = moil_ewald_beta(cutoff, moil_ewald_dtol(dtol));
  const int n = (int)(beta * dens * cutoff * 1.5);
  if (ewaldcof) *ewaldcof = beta;
  if (ntab) *ntab = n;
  if (erftbdns) *erftbdns = dens;
  if (erf_arr && cap_rows > 0) {
    std::vector<double> tab;
    moil_erfc_table(dens, n, tab);
    std::memcpy(erf_arr, tab.data(), sizeof(double) * 4 * (size_t)std::min(cap_rows, n));
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_erf_table(moil_ctx *c, double *erf_arr, int32_t cap_rows) {
  if (!c || !erf_arr) return fail(MOIL_GPU_EINVAL, "null argument");
  int n = std::min(cap_rows, c->ntab);
  // the device copy the FP64 kernels read
  if (n > 0) HIP_CHECK(hipMemcpy(erf_arr, c->d_erf, sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
  return MOIL_GPU_OK;
}

// ---------------------------------------------------------------------------
// squeeze, symm/squeeze.f:64-145 (one wrap per axis, y first, MOIL order)
__global__ void k_squeeze(double *__restrict__ coor, const int32_t *__restrict__ poipt,
                          const int32_t *__restrict__ monoflag, int totmon, double a, double b,
                          double c) {
#pragma clang fp contract(off)
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= totmon) return;
  int f = monoflag[m];
  bool cm = (f & MOIL_MONO_SQZ_CENT) != 0;
  if (!(cm || (f & MOIL_MONO_SQZ_FIRST))) return;
  int is = poipt[m], ie = poipt[m + 1];
  double xcm, ycm, zcm;
  if (cm) {
    xcm = ycm = zcm = 0.0;
    for (int k = is; k < ie; k++) {
      xcm = xcm + coor[3 * k];
      ycm = ycm + coor[3 * k + 1];
      zcm = zcm + coor[3 * k + 2];
    }
    double xn = 1.0 / (double)(ie - is);
    xcm = xcm * xn;
    ycm = ycm * xn;
    zcm = zcm * xn;
  } else {
    xcm = coor[3 * is];
    ycm = coor[3 * is + 1];
    zcm = coor[3 * is + 2];
  }
  const double xlow = -a * 0.5, xhigh = a * 0.5, ylow = -b * 0.5, yhigh = b * 0.5;
  const double zlow = -c * 0.5, zhigh = c * 0.5;
  for (int k = is; k < ie; k++) {
    if (ycm < ylow) coor[3 * k + 1] = coor[3 * k + 1] + b;
    if (ycm > yhigh) coor[3 * k + 1] = coor[3 * k + 1] - b;
    if (xcm < xlow) coor[3 * k] = coor[3 * k] + a;
    if (zcm < zlow) coor[3 * k + 2] = coor[3 * k + 2] + c;
    if (xcm > xhigh) coor[3 * k] = coor[3 * k] - a;
    if (zcm > zhigh) coor[3 * k + 2] = coor[3 * k + 2] - c;
  }
}

// the caller's stream; NULL is the default (null) stream, so device-pointer
// calls stay ordered with the caller's own work on that stream
hipStream_t pick_stream(moil_ctx *c, void *s) {
  (void)c;
  return (hipStream_t)s;
}

extern "C" int32_t moil_gpu_squeeze_dev(moil_ctx *c, double *d_coor, void *stream) {
  if (!c || !d_coor) return fail(MOIL_GPU_EINVAL, "null argument");
  hipStream_t st = pick_stream(c, stream);
  k_squeeze<<<(c->totmon + 255) / 256, 256, 0, st>>>(d_coor, c->d_poipt, c->d_monoflag, c->totmon,
                                                      c->par.a, c->par.b, c->par.c);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_squeeze(moil_ctx *c, double *coor) {
  if (!c || !coor) return fail(MOIL_GPU_EINVAL, "null argument");
  HIP_CHECK(hipMemcpyAsync(c->d_coor, coor, sizeof(double) * 3 * c->npt, hipMemcpyHostToDevice, c->stream));
  CHECK_RC(moil_gpu_squeeze_dev(c, c->d_coor, c->stream));
  HIP_CHECK(hipMemcpyAsync(coor, c->d_coor, sizeof(double) * 3 * c->npt, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return MOIL_GPU_OK;
}

// The water and atom list builds are enqueued without waiting; their sizes
// and any capacity overflow come back in one small copy at the end.  An
// overflow (the first build, or lists that grew past earlier sizes) grows
// the buffers and redoes the build.
extern "C" int32_t moil_gpu_build_lists_dev(moil_ctx *c, const double *d_coor, void *stream) {
  if (!c || !d_coor) return fail(MOIL_GPU_EINVAL, "null argument");
  hipStream_t st = pick_stream(c, stream);
  int32_t status[ST_COUNT];
  for (int attempt = 0;; attempt++) {
    CHECK_RC(lists_build(c, d_coor, st));  // clears the status words first
    CHECK_RC(atoms_build(c, d_coor, st));
    CHECK_RC(pme_bin(c, d_coor, st));      // spreading tiles, kept until the next build
    HIP_CHECK(hipMemcpyAsync(status, c->d_status, sizeof(status), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const bool w = lists_grow(c, status), a = atoms_grow(c, status);
    if (!w && !a) break;
    if (attempt >= 2) return fail(MOIL_GPU_EHIP, "list build did not fit after growing its buffers");
  }
  CHECK_RC(lists_finish(c, status));
  CHECK_RC(atoms_finish(c, status));
  CHECK_RC(ensure_parts(c, c->nwchunk + c->nachunk));
  c->have_lists = true;
  c->list_gen++;
  return MOIL_GPU_OK;
}

// Lists of n systems at once: each system's squeeze (optional) and build
// run on its own context stream after `ready` on st, with no host wait
// between systems; the status words of all builds come back in one copy
// (h_stat: n * ST_COUNT pinned words; d_stat: as many device words).
// A system whose lists did not fit is rebuilt alone after its buffers
// grew.  st waits for every build.
int32_t build_lists_many(moil_ctx *const *cs, double *const *d_coor, int n, bool squeeze,
                         hipStream_t st, int32_t *d_stat, int32_t *h_stat, hipEvent_t ready,
                         const hipEvent_t *done) {
  HIP_CHECK(hipEventRecord(ready, st));
  for (int k = 0; k < n; k++) {
    moil_ctx *c = cs[k];
    hipStream_t s = c->stream;
    HIP_CHECK(hipStreamWaitEvent(s, ready, 0));
    if (squeeze) CHECK_RC(moil_gpu_squeeze_dev(c, d_coor[k], s));
    CHECK_RC(lists_build(c, d_coor[k], s));
    CHECK_RC(atoms_build(c, d_coor[k], s));
    CHECK_RC(pme_bin(c, d_coor[k], s));
    HIP_CHECK(hipMemcpyAsync(d_stat + (size_t)ST_COUNT * k, c->d_status,
                             sizeof(int32_t) * ST_COUNT, hipMemcpyDeviceToDevice, s));
    HIP_CHECK(hipEventRecord(done[k], s));
    HIP_CHECK(hipStreamWaitEvent(st, done[k], 0));
  }
  HIP_CHECK(hipMemcpyAsync(h_stat, d_stat, sizeof(int32_t) * ST_COUNT * n,
                           hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  for (int k = 0; k < n; k++) {
    moil_ctx *c = cs[k];
    const int32_t *status = h_stat + (size_t)ST_COUNT * k;
    const bool w = lists_grow(c, status), a = atoms_grow(c, status);
    if (w || a) {
      // buffers grew: the full build with its own retries (squeezed already)
      CHECK_RC(moil_gpu_build_lists_dev(c, d_coor[k], st));
      continue;
    }
    CHECK_RC(lists_finish(c, status));
    CHECK_RC(atoms_finish(c, status));
    CHECK_RC(ensure_parts(c, c->nwchunk + c->nachunk));
    c->have_lists = true;
    c->list_gen++;
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_build_lists(moil_ctx *c, const double *coor) {
  if (!c || !coor) return fail(MOIL_GPU_EINVAL, "null argument");
  HIP_CHECK(hipMemcpyAsync(c->d_coor, coor, sizeof(double) * 3 * c->npt, hipMemcpyHostToDevice, c->stream));
  CHECK_RC(moil_gpu_build_lists_dev(c, c->d_coor, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return MOIL_GPU_OK;
}

// ---------------------------------------------------------------------------

// dpot = (fixed-point accumulator + the atom's bonded gradient slots) x
// scale + reciprocal gradient; the accumulator is cleared for the next
// evaluation on the way
__device__ __forceinline__ void finalize_impl(const FinArgs &A, int bid) {
  size_t g = (size_t)bid * 256 + threadIdx.x;
  if (g >= A.n) return;
  long long acc = A.facc[g];
  const int a = (int)(g / 3), k = (int)(g - 3 * (size_t)a);
  for (int r = A.brstart[a]; r < A.brstart[a + 1]; r++) acc += A.bslot[3 * (size_t)A.brslot[r] + k];
  double v = (double)acc * A.inv;
  if (A.has_pme) v += A.fpme[g];
  A.dpot[g] = v;
  A.facc[g] = 0;
}
__device__ __forceinline__ void finalize_body(const FinArgs &A, int bid, int nblk) {
  finalize_impl(A, bid);
}
DEFINE_KERNEL(k_finalize_forces, FinArgs, 256, finalize_body)

// deterministic reduction of the per-workgroup partials: one workgroup
// per energy term (fixed strided order + fixed shuffle tree), then the
// last workgroup to finish assembles the energy record in eforce's order
// (pot/eforce.f:100-205) and resets the ticket for the next call
__device__ __forceinline__ void energies_impl(const double *__restrict__ ep, const EnCfg &C,
                                              double *__restrict__ tot,
                                              unsigned int *__restrict__ ticket,
                                              double *__restrict__ out, int bid, int nblk) {
  __shared__ double red[4];
  __shared__ bool last;
  const int term = bid, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double s = 0.0;
  const double *row = ep + (size_t)term * C.max_parts;
  for (int i = t; i < C.nparts[term]; i += 256) s += row[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (t == 0) {
    tot[term] = (red[0] + red[1]) + (red[2] + red[3]);
    __threadfence();
    last = atomicAdd(ticket, 1u) == (unsigned)(nblk - 1);
  }
  __syncthreads();
  if (!last || t != 0) return;
  __threadfence();
  volatile const double *T = tot;
  moil_energies e;
  e.e_bond = T[EN_BOND];
  e.e_theta = T[EN_THETA];
  e.e_tors = T[EN_TORS];
  e.e_imp = T[EN_IMP];
  e.e_vdw = T[EN_VDW];
  e.e_el = T[EN_EL];
  e.e_vsym = T[EN_VSYM];
  e.e_lsym = T[EN_LSYM];
  e.e_el14 = T[EN_EL14];
  e.e_vdw14 = T[EN_VDW14];
  e.e_corr = T[EN_CORR];
  e.e_self = C.e_self;
  e.e_ew_receip = T[EN_RECIP];
  double et = 0.0;
  if (C.has_bond) et += e.e_bond;
  if (C.has_angle) et += e.e_theta;
  if (C.has_tors) et += e.e_tors;
  if (C.has_imp) et += e.e_imp;
  if (C.has_nb) {
    et += e.e_ew_receip + e.e_corr + e.e_self;
    et += e.e_vdw + e.e_el;
    if (C.has_14) et += e.e_el14 + e.e_vdw14;
    et += e.e_vsym + e.e_lsym;
  }
  e.e_total = et;
  e.e_dir = e.e_el + e.e_lsym;
  const double *src = &e.e_total;
  // a force term beyond the fixed-point range made this gradient wrong:
  // the record says so (every energy NaN), and the word is cleared
  volatile int *ovf = (volatile int *)(ticket + 1);
  const bool bad = *ovf != 0;
  for (int k = 0; k < MOIL_N_ENERGIES; k++) out[k] = bad ? __builtin_nan("") : src[k];
  if (bad) *ovf = 0;
  *ticket = 0u;
}
__device__ __forceinline__ void energies_body(const EnArgs &A, int bid, int nblk) {
  energies_impl(A.ep, A.C, A.tot, A.ticket, A.out, bid, nblk);
}
DEFINE_KERNEL(k_energies, EnArgs, 256, energies_body)

// force assembly and the energy record in one launch: workgroups
// [0, n_fin) finalize the gradient, the next EN_COUNT reduce one energy
// term each (the two halves touch disjoint memory)
__device__ __forceinline__ void assemble_body(const AsmArgs &A, int bid, int nblk) {
  if (bid < A.n_fin) finalize_impl(A.fin, bid);
  else energies_impl(A.en.ep, A.en.C, A.en.tot, A.en.ticket, A.en.out, bid - A.n_fin, EN_COUNT);
}
DEFINE_KERNEL(k_assemble, AsmArgs, 256, assemble_body)

// stage timing: event pairs recorded on the launch stream, summed lazily
static hipEvent_t ev_get(moil_ctx *c) {
  if (c->ev_next == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_next++];
}
struct StageTimer {
  moil_ctx *c;
  int stage;
  hipStream_t st;
  hipEvent_t a = nullptr;
  // stage events only at timing level 2 (an event record between kernels
  // costs a few microseconds of dispatch gap)
  StageTimer(moil_ctx *c_, int s_, hipStream_t st_) : c(c_), stage(s_), st(st_) {
    if (c->timing >= 2 && (a = ev_get(c))) hipEventRecord(a, st);
  }
  ~StageTimer() {
    if (!a) return;
    hipEvent_t b = ev_get(c);
    if (!b) return;
    hipEventRecord(b, st);
    c->ev_rec.push_back({stage, {a, b}});
  }
};

// Every launch of one evaluation with its arguments: reciprocal space (if
// the nonbonded terms are on), bonded terms, water and atom real space,
// force assembly (fixed-point accumulator + reciprocal gradient -> dpot)
// and the energy record.
void eval_plan(moil_ctx *c, const double *d_coor, double *d_dpot, double *d_energies,
               EvalPlan *p) {
  const uint32_t T = c->par.terms;
  const bool nb = (T & (MOIL_TERM_VDW | MOIL_TERM_EL)) != 0;
  *p = EvalPlan();
  p->f64 = c->par.precision != MOIL_PREC_MIXED;
  if (nb) pme_plan(c, d_coor, p);
  bonded_plan(c, d_coor, p);
  int nwpart = 0;
  if (nb) {
    water_plan(c, d_coor, p);
    nwpart = p->n_water;
    atoms_plan(c, d_coor, nwpart, p);
  }
  const size_t n3 = 3 * (size_t)c->npt;
  p->fin = {c->d_facc, c->d_fpme, nb ? 1 : 0, n3, c->fix_inv, d_dpot,
            c->d_brstart, c->d_brslot, c->d_bslot};
  p->n_fin = (int)((n3 + 255) / 256);
  EnCfg C;
  C.max_parts = c->max_parts;
  const int nbond = bonded_blocks(c), nwi = water_intra_blocks(c);
  for (int k = 0; k < EN_COUNT; k++) C.nparts[k] = 0;
  C.nparts[EN_TORS] = C.nparts[EN_IMP] = C.nparts[EN_VDW14] = C.nparts[EN_EL14] = nbond;
  C.nparts[EN_BOND] = C.nparts[EN_THETA] = C.nparts[EN_CORR] = nbond + nwi;
  const int nrs = nwpart + (nb ? p->n_atoms : 0);
  C.nparts[EN_VDW] = C.nparts[EN_EL] = C.nparts[EN_VSYM] = C.nparts[EN_LSYM] = nrs;
  C.nparts[EN_RECIP] = nb ? pme_conv_blocks(c) : 0;
  C.e_self = c->e_self;
  C.has_bond = (T & MOIL_TERM_BOND) && c->nb > 0;
  C.has_angle = (T & MOIL_TERM_ANGLE) && c->nangl > 0;
  C.has_tors = (T & MOIL_TERM_TORS) && c->ntors > 0;
  C.has_imp = (T & MOIL_TERM_IMP) && c->nimp > 0;
  C.has_nb = nb;
  C.has_14 = c->totspe > 0 && (T & (MOIL_TERM_EL14 | MOIL_TERM_VDW14));
  p->en = {c->d_epart, C, c->d_etot, c->d_ticket, d_energies ? d_energies : c->d_energy};
  p->asmb = {p->fin, p->n_fin, p->en};
}

// One evaluation in order on stream st (eager path; timing level 2 puts
// events between the stages, level 1 has the real-space kernels record
// their own dispatch events).  The packed water positions and the long
// real-space pair kernels are submitted first; the intra-water terms and
// then the reciprocal sum go to the second stream after them (forked at
// the evaluation's start, so they wait only for the coordinates) and run
// beside them; the bonded terms and the partial-row reductions follow on
// st.  The intra-water terms store (not add) the water atoms' accumulator
// rows, so both reductions wait for them.  (A third stream per context
// for the intra-water terms was measured: no faster here, and with four
// hardware queues per process it put the main streams of all contexts on
// one queue, which serialised the side-by-side list builds of many
// contexts: chain call 6 -> 17 ms.)
int32_t eforce_eager(moil_ctx *c, const EvalPlan &p, hipStream_t st) {
  // level 1 times every timing_stride-th evaluation (dispatch events cost
  // ~2 % of the evaluation rate when recorded on every one)
  const bool sample =
      c->timing >= 2 || (c->timing == 1 && (c->timing_count++ % c->timing_stride) == 0);
  if (sample) c->n_timed++;
  StageTimer whole(c, 4, st);
  const bool fork = p.pme && c->pme_concurrent;
  if (fork) HIP_CHECK(hipEventRecord(c->ev_fork, st));
  if (p.f64) launch_k_pack_water_f64(p.pack, p.n_pack, st);
  else launch_k_pack_water_f32(p.pack, p.n_pack, st);
  hipEvent_t ea = nullptr, eb = nullptr;
  if (sample && p.n_water > 0) {
    ea = ev_get(c);
    eb = ev_get(c);
  }
  launch_k_water_ev(p.water, p.f64, p.n_water, st, ea, eb, c->water_wgs);
  if (ea && eb) c->ev_rec.push_back({0, {ea, eb}});
  ea = eb = nullptr;
  if (sample && p.n_atoms > 0) {
    ea = ev_get(c);
    eb = ev_get(c);
  }
  launch_k_atoms_ev(p.atoms, p.f64, p.n_atoms, st, ea, eb);
  if (ea && eb) c->ev_rec.push_back({1, {ea, eb}});
  if (fork) {
    // the intra-water terms first on the second stream (they touch no
    // buffer of the pair kernels), then the reciprocal sum
    HIP_CHECK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    launch_k_water_intra(p.intra, p.n_intra, c->stream2);
    HIP_CHECK(hipEventRecord(c->ev_intra, c->stream2));
    {
      StageTimer tp(c, 2, c->stream2);
      CHECK_RC(pme_run(c, p, c->stream2));
    }
    HIP_CHECK(hipEventRecord(c->ev_join, c->stream2));
  } else if (p.pme) {
    StageTimer tp(c, 2, st);
    CHECK_RC(pme_run(c, p, st));
  }
  {
    StageTimer tb(c, 3, st);
    if (!fork) launch_k_water_intra(p.intra, p.n_intra, st);
    launch_k_bonded(p.bonded, p.n_bonded, st);
    HIP_CHECK(hipGetLastError());
  }
  // the intra-water terms store the water atoms' accumulator rows, which
  // the reductions then add to
  if (fork) HIP_CHECK(hipStreamWaitEvent(st, c->ev_intra, 0));
  launch_k_water_reduce(p.wred, p.n_wred, st);
  launch_k_atoms_reduce(p.ared, p.n_ared, st);
  if (fork) HIP_CHECK(hipStreamWaitEvent(st, c->ev_join, 0));
  launch_k_assemble(p.asmb, p.n_fin + EN_COUNT, st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

// The same sequence into a graph being captured on the context's graph
// stream (no events: event nodes of a graph carry no timestamps here).
static int32_t eforce_capture_body(moil_ctx *c, const EvalPlan &p, hipStream_t st) {
  const bool fork = p.pme && c->pme_concurrent;
  if (fork) HIP_CHECK(hipEventRecord(c->ev_fork, st));
  if (p.f64) launch_k_pack_water_f64(p.pack, p.n_pack, st);
  else launch_k_pack_water_f32(p.pack, p.n_pack, st);
  launch_k_water_ev(p.water, p.f64, p.n_water, st, nullptr, nullptr, c->water_wgs);
  if (p.f64) launch_k_atoms_f64(p.atoms, p.n_atoms, st);
  else launch_k_atoms_f32(p.atoms, p.n_atoms, st);
  if (fork) {
    HIP_CHECK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    launch_k_water_intra(p.intra, p.n_intra, c->stream2);
    HIP_CHECK(hipEventRecord(c->ev_intra, c->stream2));
    CHECK_RC(pme_run(c, p, c->stream2));
    HIP_CHECK(hipEventRecord(c->ev_join, c->stream2));
  } else if (p.pme) {
    CHECK_RC(pme_run(c, p, st));
  }
  if (!fork) launch_k_water_intra(p.intra, p.n_intra, st);
  launch_k_bonded(p.bonded, p.n_bonded, st);
  if (fork) HIP_CHECK(hipStreamWaitEvent(st, c->ev_intra, 0));
  launch_k_water_reduce(p.wred, p.n_wred, st);
  launch_k_atoms_reduce(p.ared, p.n_ared, st);
  if (fork) HIP_CHECK(hipStreamWaitEvent(st, c->ev_join, 0));
  launch_k_assemble(p.asmb, p.n_fin + EN_COUNT, st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

// a replay of the old graph may still be in flight on the graph stream: it
// ends before the exec is destroyed
void egraph_drop(moil_ctx *c) {
  if (c->eg.exec && c->gstream) hipStreamSynchronize(c->gstream);
  if (c->eg.exec) hipGraphExecDestroy(c->eg.exec);
  c->eg.exec = nullptr;
}

extern "C" int32_t moil_gpu_eforce_dev(moil_ctx *c, const double *d_coor, double *d_dpot,
                                       double *d_energies, void *stream) {
  if (!c || !d_coor || !d_dpot) return fail(MOIL_GPU_EINVAL, "null argument");
  if (!c->have_lists) return fail(MOIL_GPU_ESTATE, "eforce called before build_lists");
  hipStream_t st = pick_stream(c, stream);
  if (c->facc_dirty) {
    HIP_CHECK(hipMemsetAsync(c->d_facc, 0, sizeof(long long) * 3 * (size_t)c->npt, st));
    c->facc_dirty = false;
  }
  EvalPlan p;
  eval_plan(c, d_coor, d_dpot, d_energies, &p);
  // the accumulator is dirty until an evaluation's finalize is enqueued
  c->facc_dirty = true;
  if (c->timing >= 2 || !c->use_graph) {
    CHECK_RC(eforce_eager(c, p, st));
    c->facc_dirty = false;
    return MOIL_GPU_OK;
  }
  // Replay of a captured evaluation: one submission instead of ~20 kernel
  // launches and two transforms.  The graph is recaptured when the lists,
  // the parameters or the pointers changed.  At timing level 1 the first
  // evaluation after such a change runs in order with the real-space
  // kernels' dispatch events (the timed sample); the replays are untimed.
  EGraph &G = c->eg;
  const void *key[3] = {d_coor, d_dpot, d_energies};
  bool fresh = G.exec && G.gen == c->list_gen && G.pgen == c->param_gen;
  for (int k = 0; fresh && k < 3; k++) fresh = G.key[k] == key[k];
  if (!fresh && c->timing >= 1 && c->timed_gen != c->list_gen) {
    c->timed_gen = c->list_gen;
    CHECK_RC(eforce_eager(c, p, st));
    c->facc_dirty = false;
    return MOIL_GPU_OK;
  }
  if (!fresh) {
    egraph_drop(c);
    HIP_CHECK(hipEventRecord(c->gev_in, st));
    HIP_CHECK(hipStreamWaitEvent(c->gstream, c->gev_in, 0));
    HIP_CHECK(hipStreamBeginCapture(c->gstream, hipStreamCaptureModeThreadLocal));
    const int32_t rc = eforce_capture_body(c, p, c->gstream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->gstream, &g);
    if (rc != MOIL_GPU_OK || e != hipSuccess || !g) {
      if (g) hipGraphDestroy(g);
      // capture not available: in order from now on
      c->use_graph = false;
      hipGetLastError();
      CHECK_RC(rc);
      CHECK_RC(eforce_eager(c, p, st));
      c->facc_dirty = false;
      return MOIL_GPU_OK;
    }
    const hipError_t ei = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    HIP_CHECK(ei);
    G.gen = c->list_gen;
    G.pgen = c->param_gen;
    for (int k = 0; k < 3; k++) G.key[k] = key[k];
  }
  HIP_CHECK(hipEventRecord(c->gev_in, st));
  HIP_CHECK(hipStreamWaitEvent(c->gstream, c->gev_in, 0));
  HIP_CHECK(hipGraphLaunch(G.exec, c->gstream));
  HIP_CHECK(hipEventRecord(c->gev_out, c->gstream));
  HIP_CHECK(hipStreamWaitEvent(st, c->gev_out, 0));
  c->facc_dirty = false;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_eforce(moil_ctx *c, const double *coor, double *dpot, moil_energies *e) {
  if (!c || !coor || !dpot) return fail(MOIL_GPU_EINVAL, "null argument");
  size_t n3 = 3 * (size_t)c->npt;
  HIP_CHECK(hipMemcpyAsync(c->d_coor, coor, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  CHECK_RC(moil_gpu_eforce_dev(c, c->d_coor, c->d_dpot, c->d_energy, c->stream));
  HIP_CHECK(hipMemcpyAsync(dpot, c->d_dpot, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  double en[MOIL_N_ENERGIES];
  HIP_CHECK(hipMemcpyAsync(en, c->d_energy, sizeof(en), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (e) std::memcpy(e, en, sizeof(en));
  if (std::isnan(en[0]))
    return fail(MOIL_GPU_ECAPACITY,
                "eforce: a force term beyond the fixed-point range (" +
                    std::to_string(FIX_TERM_UNITS / c->fix_scale) +
                    " kcal/mol/A); energies are NaN and the gradient is not valid");
  return MOIL_GPU_OK;
}

// ---------------------------------------------------------------------------
extern "C" int32_t moil_gpu_list_counts(moil_ctx *c, int64_t counts[8]) {
  if (!c || !counts) return fail(MOIL_GPU_EINVAL, "null argument");
  for (int k = 0; k < 8; k++) counts[k] = 0;
  if (!c->have_lists) return fail(MOIL_GPU_ESTATE, "no lists built");
  if (c->nwitems) {
    std::vector<WItem> it(c->nwitems);
    HIP_CHECK(hipMemcpy(it.data(), c->d_witems, sizeof(WItem) * it.size(), hipMemcpyDeviceToHost));
    for (const WItem &w : it) {
      int img = w.s > 0 ? 2 : 0;
      counts[img + 0] += __builtin_popcountll(w.m1);
      counts[img + 1] += __builtin_popcountll(w.m2);
    }
  }
  if (c->naitems) {
    std::vector<AItem> it(c->naitems);
    HIP_CHECK(hipMemcpy(it.data(), c->d_aitems, sizeof(AItem) * it.size(), hipMemcpyDeviceToHost));
    for (const AItem &a : it) {
      const int64_t n1 = __builtin_popcountll(a.m1), n2 = __builtin_popcountll(a.m2),
                    n3 = __builtin_popcountll(a.m3);
      counts[4] += n1;
      counts[5] += n2;
      counts[6] += n3;
      if (a.s > 0) counts[7] += n1 + n2 + n3;
    }
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_dump_water_pairs(moil_ctx *c, int32_t *rows, int64_t cap, int64_t *n) {
  if (!c || !n) return fail(MOIL_GPU_EINVAL, "null argument");
  if (!c->have_lists) return fail(MOIL_GPU_ESTATE, "no lists built");
  *n = 0;
  if (c->nwat == 0 || c->nwitems == 0) return MOIL_GPU_OK;
  std::vector<WItem> it(c->nwitems);
  std::vector<int32_t> s2w(c->nwat), boff(c->nwblk * WNBUCKET + 1);
  HIP_CHECK(hipMemcpy(it.data(), c->d_witems, sizeof(WItem) * it.size(), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(s2w.data(), c->d_slot2wat, sizeof(int32_t) * s2w.size(), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(boff.data(), c->d_boff, sizeof(int32_t) * boff.size(), hipMemcpyDeviceToHost));
  int64_t k = 0;
  for (int b = 0; b < c->nwblk; b++)
    for (int q = boff[b * WNBUCKET]; q < boff[(b + 1) * WNBUCKET]; q++) {
      const WItem &w = it[q];
      int jm = c->h_wat_mono[s2w[w.j]];
      for (int kind = 1; kind <= 2; kind++) {
        uint64_t m = kind == 1 ? w.m1 : w.m2;
        while (m) {
          int i = __builtin_ctzll(m);
          m &= m - 1;
          int im = c->h_wat_mono[s2w[b * WBLOCK + i]];
          if (rows && k < cap) {
            rows[4 * k] = im;
            rows[4 * k + 1] = jm;
            rows[4 * k + 2] = w.s;
            rows[4 * k + 3] = kind;
          }
          k++;
        }
      }
    }
  *n = k;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_dump_atom_pairs(moil_ctx *c, int32_t *rows, int64_t cap, int64_t *n) {
  if (!c || !n) return fail(MOIL_GPU_EINVAL, "null argument");
  if (!c->have_lists) return fail(MOIL_GPU_ESTATE, "no lists built");
  *n = 0;
  if (c->naitems == 0) return MOIL_GPU_OK;
  std::vector<AItem> it(c->naitems);
  std::vector<int32_t> s2a(c->nsslot), aboff((size_t)c->nablk * NBUCKET + 1);
  HIP_CHECK(hipMemcpy(aboff.data(), c->d_aboff, sizeof(int32_t) * aboff.size(), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(it.data(), c->d_aitems, sizeof(AItem) * it.size(), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(s2a.data(), c->d_sslot2atom, sizeof(int32_t) * s2a.size(), hipMemcpyDeviceToHost));
  int64_t k = 0;
  for (int b = 0; b < c->nablk; b++)
    for (int q = aboff[b * NBUCKET]; q < aboff[(b + 1) * NBUCKET]; q++) {
      const AItem &a = it[q];
      const bool rev = a.s > 13;
      const int sym = rev ? a.s - 13 : a.s;
      for (int kind = 1; kind <= 3; kind++) {
        uint64_t m = kind == 1 ? a.m1 : (kind == 2 ? a.m2 : a.m3);
        while (m) {
          const int i = __builtin_ctzll(m);
          m &= m - 1;
          const int ai = s2a[b * WBLOCK + i] + 1, aj = a.j + 1;
          int l, mm;
          if (sym == 0) {  // real: list owner is the lower atom
            l = std::min(ai, aj);
            mm = std::max(ai, aj);
          } else if (!rev) {
            l = ai;
            mm = aj;
          } else {
            l = aj;
            mm = ai;
          }
          if (rows && k < cap) {
            rows[4 * k] = l;
            rows[4 * k + 1] = mm;
            rows[4 * k + 2] = sym;
            rows[4 * k + 3] = kind;
          }
          k++;
        }
      }
    }
  *n = k;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_list_generation(moil_ctx *c, int64_t *gen) {
  if (!c || !gen) return fail(MOIL_GPU_EINVAL, "null argument");
  *gen = (int64_t)c->list_gen;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_set_timing(moil_ctx *c, int32_t level) {
  if (!c) return fail(MOIL_GPU_EINVAL, "null context");
  const int lv = level & 0xff, stride = (level >> 8) & 0xff;
  c->timing = lv < 0 ? 0 : (lv > 2 ? 2 : lv);
  c->timing_stride = stride < 1 ? 1 : stride;
  c->timing_count = 0;
  c->timed_gen = -1;
  c->ev_rec.clear();
  c->ev_next = 0;
  c->n_timed = 0;
  return MOIL_GPU_OK;
}

// t[stage] = summed milliseconds since the last call: t[0] water real
// space, t[1] atom real space (level >= 1); t[2] PME, t[3] bonded, t[4]
// whole eforce (level 2); t[5] number of eforce calls timed
extern "C" int32_t moil_gpu_stage_times(moil_ctx *c, double t[8]) {
  if (!c || !t) return fail(MOIL_GPU_EINVAL, "null argument");
  for (int k = 0; k < 8; k++) t[k] = 0.0;
  for (auto &r : c->ev_rec) {
    HIP_CHECK(hipEventSynchronize(r.second.second));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, r.second.first, r.second.second));
    if (r.first >= 0 && r.first < 5) t[r.first] += ms;
  }
  t[5] = (double)c->n_timed;
  c->ev_rec.clear();
  c->ev_next = 0;
  c->n_timed = 0;
  return MOIL_GPU_OK;
}

// ---------------------------------------------------------------------------
// Force-field parameters and their gradients (no counterpart in MOIL)

extern "C" int32_t moil_gpu_set_params(moil_ctx *c, const double *ptchg, const double *epsgm6,
                                       const double *epsgm12, double v14, double el14) {
  if (!c || !ptchg || !epsgm6 || !epsgm12) return fail(MOIL_GPU_EINVAL, "null argument");
  if (c->par.terms & MOIL_TERM_ARITH)
    return fail(MOIL_GPU_EINVAL, "set_params derives the 1-4 pairs by the geometric rule: "
                                 "not for an arith context");
  const size_t n = (size_t)c->npt;
  // Water-water pairs take the last water's charges and O LJ factors
  // (pot/watwat_ewald.f:40-72, the H charge being the last H's); the
  // parameter gradient spreads each such pair over the two waters' own
  // atoms, which is the gradient of the energy eforce evaluates only when
  // every water carries those same values.
  if (c->nwat >= 2) {
    const int lo = c->h_wat_o[c->nwat - 1];
    const double qo = ptchg[lo], qh = ptchg[lo + 2], e6o = epsgm6[lo], e12o = epsgm12[lo];
    for (int w = 0; w < c->nwat; w++) {
      const int o = c->h_wat_o[w];
      if (ptchg[o] != qo || ptchg[o + 1] != qh || ptchg[o + 2] != qh || epsgm6[o] != e6o ||
          epsgm12[o] != e12o)
        return fail(MOIL_GPU_EINVAL,
                    "set_params: water " + std::to_string(w + 1) +
                        " differs from the last water in its charges or O LJ factors "
                        "(water-water pairs use the last water's)");
    }
  }
  {
    std::string why;
    if (!pme_charges_fit(ptchg, n, c->par.eps, &why)) return fail(MOIL_GPU_EINVAL, "set_params: " + why);
  }
  c->h_q.assign(ptchg, ptchg + n);
  c->h_e6.assign(epsgm6, epsgm6 + n);
  c->h_e12.assign(epsgm12, epsgm12 + n);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipDeviceSynchronize());
  HIP_CHECK(hipMemcpy(c->d_ptchg, ptchg, sizeof(double) * n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_eps6, epsgm6, sizeof(double) * n, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(c->d_eps12, epsgm12, sizeof(double) * n, hipMemcpyHostToDevice));
  // self energy (symm/ewald_init.f:131-150)
  {
    double s = 0;
    for (size_t i = 0; i < n; i++) s += ptchg[i] * ptchg[i];
    c->e_self = -(c->beta * std::sqrt(1.0 / (4.0 * std::atan(1.0))) * (KOFDIE / c->par.eps) * s);
  }
  // last-water parameters (pot/watwat_ewald.f:40-72)
  if (c->nwat >= 2) {
    const int io = c->h_wat_o[c->nwat - 1], ih = io + 2;
    const double epst = (c->par.terms & MOIL_TERM_EL) ? 1.0 : 0.0;
    const double pick = (c->par.terms & MOIL_TERM_VDW) ? 1.0 : 0.0;
    c->oo6 = pick * epsgm6[io] * epsgm6[io];
    c->oo12 = pick * epsgm12[io] * epsgm12[io];
    c->qq_oo = epst * KOFDIE / c->par.eps * ptchg[io] * ptchg[io];
    c->qq_oh = epst * KOFDIE / c->par.eps * ptchg[io] * ptchg[ih];
    c->qq_hh = epst * KOFDIE / c->par.eps * ptchg[ih] * ptchg[ih];
  }
  // 1-4 pair parameters (generic/wconn.f:296-325, geometric rule, hvdw0)
  if (c->totspe > 0) {
    std::vector<double> p14(3 * (size_t)c->totspe);
    for (int m = 0; m < c->totspe; m++) {
      const int j = c->h_sp[2 * m] - 1, k = c->h_sp[2 * m + 1] - 1;
      if (epsgm12[j] < 1e-3 || epsgm12[k] < 1e-3) {
        p14[3 * m] = p14[3 * m + 1] = 0.0;
      } else {
        p14[3 * m] = epsgm12[j] * epsgm12[k] * v14;
        p14[3 * m + 1] = epsgm6[j] * epsgm6[k] * v14;
      }
      p14[3 * m + 2] = ptchg[j] * ptchg[k] * c->par.eps * el14 * KOFDIE;
    }
    HIP_CHECK(hipMemcpy(c->d_p14, p14.data(), sizeof(double) * p14.size(), hipMemcpyHostToDevice));
  }
  c->v14 = v14;
  c->el14 = el14;
  c->have_scale14 = true;
  c->param_gen++;  // captured evaluations hold the old derived parameters
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_param_grad_dev(moil_ctx *c, const double *d_coor, double *d_dq,
                                           double *d_de12, double *d_de6, void *stream) {
  if (!c || !d_coor || !d_dq || !d_de12 || !d_de6) return fail(MOIL_GPU_EINVAL, "null argument");
  if (c->par.terms & MOIL_TERM_ARITH)
    return fail(MOIL_GPU_EINVAL, "param_grad differentiates the geometric rule's factors: "
                                 "not for an arith context");
  if (!c->have_lists) return fail(MOIL_GPU_ESTATE, "param_grad called before build_lists");
  if (c->totspe > 0 && !c->have_scale14)
    return fail(MOIL_GPU_ESTATE,
                "param_grad needs the 1-4 scalings: call moil_gpu_set_params first");
  return adjoint_run(c, d_coor, d_dq, d_de12, d_de6, pick_stream(c, stream));
}

extern "C" int32_t moil_gpu_param_grad(moil_ctx *c, const double *coor, double *dq, double *de12,
                                       double *de6) {
  if (!c || !coor || !dq || !de12 || !de6) return fail(MOIL_GPU_EINVAL, "null argument");
  const size_t n = (size_t)c->npt;
  if (!c->d_adj) HIP_CHECK(hipMalloc(&c->d_adj, sizeof(double) * 3 * n));
  HIP_CHECK(hipMemcpyAsync(c->d_coor, coor, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                           c->stream));
  CHECK_RC(moil_gpu_param_grad_dev(c, c->d_coor, c->d_adj, c->d_adj + n, c->d_adj + 2 * n,
                                   c->stream));
  HIP_CHECK(hipMemcpyAsync(dq, c->d_adj, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(de12, c->d_adj + n, sizeof(double) * n, hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipMemcpyAsync(de6, c->d_adj + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return MOIL_GPU_OK;
}
