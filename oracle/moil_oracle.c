/*
 * moil_oracle.c -- CPU restatement of MOIL's energy/force path in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is
 * measured against; it is never linked into, loaded by or called from the
 * product.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Each routine restates one MOIL routine (file:line given at the routine)
 * in plain C, in MOIL's loop structure and FP64 operation order wherever
 * that order decides a discrete outcome (pair-list predicates, erfc table
 * index).  It is pinned against MOIL itself (oracle/_ref, golden fixtures
 * under tests/golden) and against closed forms in tests/test_oracle*.py.
 *
 * Restrictions (same as the GPU library): orthorhombic box, monomer
 * division off (mdivyes=.false.), no LES copies, no metal walls, no
 * virtual particles, PME order 4, one process (the serial comm_dummy
 * path).  Both LJ combination rules: geometric (arith=.false.) and
 * MOIL's arith rule with its AMBER torsion and improper forms
 * (MOIL_TERM_ARITH).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/moil_gpu.h"

#define KOFDIE 332.0716 /* COMMON/CONVERT.BLOCK:13 */

/* ------------------------------------------------------------------ */
/* small growable int array                                            */
typedef struct { int32_t *v; int64_t n, cap; } ivec;
static void iv_push(ivec *a, int32_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 1024;
    a->v = (int32_t *)realloc(a->v, (size_t)a->cap * sizeof(int32_t));
  }
  a->v[a->n++] = x;
}
static void iv_free(ivec *a) { free(a->v); a->v = NULL; a->n = a->cap = 0; }

typedef struct orc_ctx {
  /* topology (deep copies, 1-based values) */
  int npt, totmon, nb, nangl, ntors, nimp, totspe, nwaters;
  int32_t *poipt, *monoflag, *flagchr, *exc1, *exc2;
  double *ptchg, *epsgm6, *epsgm12;
  int32_t *ib1, *ib2, *ia1, *ia2, *ia3, *it1, *it2, *it3, *it4, *period;
  int32_t *ii1, *ii2, *ii3, *ii4, *sp1, *sp2, *idxtip3;
  double *kbond, *req, *kangl, *angleq, *kt1, *kt2, *kt3, *ph1, *ph2, *ph3;
  double *kimp, *impeq, *p14;
  moil_params p;
  /* ewald_init products */
  double beta, e_self, erftbdns, del;
  int ntab;
  double *erf_arr; /* [ntab][4] */
  double *bsp[3];
  /* lists in MOIL layout */
  double *mono_cent;        /* [totmon][3] */
  int nwat;                 /* waters in monomer order */
  int32_t *indxwat;         /* [nwat+1] 1-based monomer of water iwat */
  int32_t *poinwat1, *poinwat2; /* [nwat+2] */
  ivec listwt1, listwt2;    /* real then image entries */
  int nreal_wt1, nreal_wt2;
  ivec symmwat1, psymwt1, symmwat2, psymwt2;
  int iblckwt1[14], iblckwt2[14];
  int32_t *point1, *point2, *point3; /* [npt+1] */
  ivec list1, list2, list3;
  int nreal1, nreal2, nreal3;
  ivec symreal1, psym1, symreal2, psym2, symreal3, psym3;
  int iblock1[14], iblock2[14], iblock3[14];
  int symop[14][3];
  int have_lists;
  /* molecular virial (MOIL's pressON, COMMON/PRESS.BLOCK): molecules as
   * init_pressure defines them, centres of mass before every eforce */
  int nmol;
  int32_t *patom, *ordlist, *pmol; /* [nmol], [npt], [npt] (1-based) */
  double *ptms, *massmol, *xmol, *xrel; /* [npt], [nmol], [3*nmol], [3*npt] */
  double vir[13]; /* virial, virXX..ZZ, V_PIdir, V_PIrec, V_PIrel (XX, YY, ZZ) */
} orc_ctx;

static int32_t *dup_i(const int32_t *s, int64_t n) {
  int32_t *d = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  if (n > 0 && s) memcpy(d, s, (size_t)n * sizeof(int32_t));
  return d;
}
static double *dup_d(const double *s, int64_t n) {
  double *d = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  if (n > 0 && s) memcpy(d, s, (size_t)n * sizeof(double));
  return d;
}

/* ------------------------------------------------------------------ */
/* ewald_init restatement                                              */

/* Cody's erfc, the NETLIB SPECFUN CALERF with JINT = 1 that MOIL ships as
 * symm/erfcfun.f:14-288 (W. J. Cody, Math. Comp. 23 (1969) 631-638):
 * three rational approximations, on |x| <= 0.46875 (erf), 0.46875..4 and
 * above 4 (in 1/x^2), the Gaussian factor split at a multiple of 1/16.
 * Coefficients as published; MOIL's operation order. */
static double cody_erfc(double x) {
  static const double a[5] = {3.16112374387056560e00, 1.13864154151050156e02,
                              3.77485237685302021e02, 3.20937758913846947e03,
                              1.85777706184603153e-1};
  static const double b[4] = {2.36012909523441209e01, 2.44024637934444173e02,
                              1.28261652607737228e03, 2.84423683343917062e03};
  static const double cc[9] = {5.64188496988670089e-1, 8.88314979438837594e0,
                               6.61191906371416295e01, 2.98635138197400131e02,
                               8.81952221241769090e02, 1.71204761263407058e03,
                               2.05107837782607147e03, 1.23033935479799725e03,
                               2.15311535474403846e-8};
  static const double d[8] = {1.57449261107098347e01, 1.17693950891312499e02,
                              5.37181101862009858e02, 1.62138957456669019e03,
                              3.29079923573345963e03, 4.36261909014324716e03,
                              3.43936767414372164e03, 1.23033935480374942e03};
  static const double p[6] = {3.05326634961232344e-1, 3.60344899949804439e-1,
                              1.25781726111229246e-1, 1.60837851487422766e-2,
                              6.58749161529837803e-4, 1.63153871373020978e-2};
  static const double q[5] = {2.56852019228982242e00, 1.87295284992346047e00,
                              5.27905102951428412e-1, 6.05183413124413191e-2,
                              2.33520497626869185e-3};
  double y = fabs(x), res, num, den, ysq, dl;
  int i;
  if (y <= 0.46875) {
    ysq = y > 1.11e-16 ? y * y : 0.0;
    num = a[4] * ysq;
    den = ysq;
    for (i = 0; i < 3; i++) {
      num = (num + a[i]) * ysq;
      den = (den + b[i]) * ysq;
    }
    res = x * (num + a[3]) / (den + b[3]);
    return 1.0 - res;
  }
  if (y <= 4.0) {
    num = cc[8] * y;
    den = y;
    for (i = 0; i < 7; i++) {
      num = (num + cc[i]) * y;
      den = (den + d[i]) * y;
    }
    res = (num + cc[7]) / (den + d[7]);
  } else {
    if (y >= 26.543) return x < 0.0 ? 2.0 : 0.0;
    ysq = 1.0 / (y * y);
    num = p[5] * ysq;
    den = ysq;
    for (i = 0; i < 4; i++) {
      num = (num + p[i]) * ysq;
      den = (den + q[i]) * ysq;
    }
    res = ysq * (num + p[4]) / (den + q[4]);
    res = (5.6418958354775628695e-1 - res) / y;
  }
  ysq = trunc(y * 16.0) / 16.0;
  dl = (y - ysq) * (y + ysq);
  res = exp(-ysq * ysq) * exp(-dl) * res;
  return x < 0.0 ? 2.0 - res : res;
}

/* find_ewaldcof, symm/ewald_init.f:298-335: doubling then 2^-60 bisection
 * of erfc(x*rc)/rc >= dtol with Cody's erfc; dtol first clamped as
 * symm/ewald_init.f:78-82 does. */
static double find_ewaldcof(double cutoff, double dtol) {
  double x = 0.5, xlo, xhi, term;
  int i = 0, n;
  if (dtol < 1e-10) dtol = 1e-10;
  else if (dtol > 0.1) dtol = 5.2e-05;
  do {
    x = 2.0 * x;
    i = i + 1;
    term = cody_erfc(x * cutoff) / cutoff;
  } while (term >= dtol);
  n = i + 60;
  xlo = 0.0;
  xhi = x;
  for (i = 1; i <= n; i++) {
    x = (xlo + xhi) / 2;
    term = cody_erfc(x * cutoff) / cutoff;
    if (term >= dtol) xlo = x; else xhi = x;
  }
  return x;
}

/* de Boor's cubspl (symm/ewald_init.f:453-574) for MOIL's table: abscissae
 * tau_i = i*del, values in tab[4i], end slopes in tab[1] and tab[4(n-1)+1].
 * Gaussian elimination of the slope system written with a step h_m =
 * tau_m - tau_{m-1} and divided difference dd_m per node (MOIL keeps
 * them in rows 3 and 4 of erf_arr, which is why its last row ends up
 * holding h and dd), then each piece's derivatives at its left end. */
static void clamped_spline(double *tab, int n, double del) {
  double *h = (double *)malloc(sizeof(double) * n);   /* h[m], m >= 1 */
  double *dd = (double *)malloc(sizeof(double) * n);  /* dd[m], then the pivots */
  double *sl = (double *)malloc(sizeof(double) * n);  /* slopes (right sides first) */
  int m;
  for (m = 1; m < n; m++) {
    h[m] = del * m - del * (m - 1);
    dd[m] = (tab[4 * m] - tab[4 * (m - 1)]) / h[m];
  }
  h[0] = 0.0;   /* the first equation is s_0 = given slope */
  dd[0] = 1.0;
  sl[0] = tab[1];
  sl[n - 1] = tab[4 * (n - 1) + 1];
  for (m = 1; m <= n - 2; m++) {
    double g = -h[m + 1] / dd[m - 1];
    sl[m] = g * sl[m - 1] + 3.0 * (h[m] * dd[m + 1] + h[m + 1] * dd[m]);
    dd[m] = g * h[m - 1] + 2.0 * (h[m] + h[m + 1]);
  }
  for (m = n - 2; m >= 0; m--) sl[m] = (sl[m] - h[m] * sl[m + 1]) / dd[m];
  for (m = 0; m < n; m++) tab[4 * m + 1] = sl[m];
  for (m = 1; m < n; m++) {
    double d1 = (tab[4 * m] - tab[4 * (m - 1)]) / h[m];
    double d3 = sl[m - 1] + sl[m] - 2.0 * d1;
    tab[4 * (m - 1) + 2] = 2.0 * (d1 - sl[m - 1] - d3) / h[m];
    tab[4 * (m - 1) + 3] = (d3 / h[m]) * (6.0 / h[m]);
  }
  tab[4 * (n - 1) + 2] = h[n - 1];
  tab[4 * (n - 1) + 3] = (tab[4 * (n - 1)] - tab[4 * (n - 2)]) / h[n - 1];
  free(h); free(dd); free(sl);
}

/* fill_bspline/init/one_pass/diff, pot/ewald_receip.f:1013-1080 */
static void one_pass(double *c, double x, int k) {
  double div = 1.0 / (k - 1);
  int j;
  c[k - 1] = div * x * c[k - 2];
  for (j = 1; j <= k - 2; j++)
    c[k - j - 1] = div * ((x + j) * c[k - j - 2] + ((k - j) - x) * c[k - j - 1]);
  c[0] = div * (1 - x) * c[0];
}
static void fill_bspline(double w, int order, double *arr, double *darr) {
  int k, j;
  arr[order - 1] = 0.0;
  arr[1] = w;
  arr[0] = 1.0 - w;
  for (k = 3; k <= order - 1; k++) one_pass(arr, w, k);
  darr[0] = -arr[0];
  for (j = 2; j <= order; j++) darr[j - 1] = arr[j - 2] - arr[j - 1];
  one_pass(arr, w, order);
}

/* load_bsp_moduli + DFTMOD, pot/ewald_receip.f:413-481 */
static void bsp_moduli(double *mod, int nfft, int order) {
  double arr[32], darr[32];
  double *bsp_arr = (double *)calloc((size_t)nfft + order + 2, sizeof(double));
  double twopi = 2.0 * 3.14159265358979323846;
  int j, k;
  fill_bspline(0.0, order, arr, darr);
  for (k = 2; k <= order + 1; k++) bsp_arr[k - 1] = arr[k - 2];
  for (k = 1; k <= nfft; k++) {
    double s1 = 0, s2 = 0;
    for (j = 1; j <= nfft; j++) {
      double arg = twopi * (k - 1) * (j - 1) / nfft;
      s1 += bsp_arr[j - 1] * cos(arg);
      s2 += bsp_arr[j - 1] * sin(arg);
    }
    mod[k - 1] = s1 * s1 + s2 * s2;
  }
  for (k = 1; k <= nfft; k++)
    if (mod[k - 1] < 1e-7) {
      double lo = mod[(k - 2 + nfft) % nfft], hi = mod[k % nfft];
      mod[k - 1] = 0.5 * (lo + hi);
    }
  free(bsp_arr);
}

orc_ctx *orc_create(const moil_topology *t, const moil_params *p, char *err, int errlen) {
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  int i;
  double totchg = 0, cutoff, fac, spi, epstmp;
  c->p = *p;
  c->npt = t->npt; c->totmon = t->totmon; c->nb = t->nb; c->nangl = t->nangl;
  c->ntors = t->ntors; c->nimp = t->nimp; c->totspe = t->totspe; c->nwaters = t->nwaters;
  c->poipt = dup_i(t->poipt, t->totmon + 1);
  c->monoflag = dup_i(t->monoflag, t->totmon);
  c->flagchr = dup_i(t->flagchr, t->npt);
  c->exc1 = dup_i(t->exc1, t->npt + 1);
  c->exc2 = dup_i(t->exc2, t->exc1[t->npt]);
  c->ptchg = dup_d(t->ptchg, t->npt);
  c->epsgm6 = dup_d(t->epsgm6, t->npt);
  c->epsgm12 = dup_d(t->epsgm12, t->npt);
  c->ib1 = dup_i(t->ib1, t->nb); c->ib2 = dup_i(t->ib2, t->nb);
  c->kbond = dup_d(t->kbond, t->nb); c->req = dup_d(t->req, t->nb);
  c->ia1 = dup_i(t->iangl1, t->nangl); c->ia2 = dup_i(t->iangl2, t->nangl);
  c->ia3 = dup_i(t->iangl3, t->nangl);
  c->kangl = dup_d(t->kangl, t->nangl); c->angleq = dup_d(t->angleq, t->nangl);
  c->it1 = dup_i(t->itor1, t->ntors); c->it2 = dup_i(t->itor2, t->ntors);
  c->it3 = dup_i(t->itor3, t->ntors); c->it4 = dup_i(t->itor4, t->ntors);
  c->period = dup_i(t->period, t->ntors);
  c->kt1 = dup_d(t->ktors1, t->ntors); c->kt2 = dup_d(t->ktors2, t->ntors);
  c->kt3 = dup_d(t->ktors3, t->ntors);
  c->ph1 = dup_d(t->phase1, t->ntors); c->ph2 = dup_d(t->phase2, t->ntors);
  c->ph3 = dup_d(t->phase3, t->ntors);
  c->ii1 = dup_i(t->iimp1, t->nimp); c->ii2 = dup_i(t->iimp2, t->nimp);
  c->ii3 = dup_i(t->iimp3, t->nimp); c->ii4 = dup_i(t->iimp4, t->nimp);
  c->kimp = dup_d(t->kimp, t->nimp); c->impeq = dup_d(t->impeq, t->nimp);
  c->sp1 = dup_i(t->spec1, t->totspe); c->sp2 = dup_i(t->spec2, t->totspe);
  c->p14 = dup_d(t->p14, 3 * (int64_t)t->totspe);
  c->idxtip3 = dup_i(t->idxtip3, t->nwaters);

  /* symm/ewald_init.f:62-72 net charge check */
  for (i = 0; i < t->npt; i++) totchg += t->ptchg[i];
  if (fabs(totchg) > 1e-7) {
    if (err) snprintf(err, errlen, "net charge %.10g: add counterions", totchg);
    free(c);
    return NULL;
  }
  if (p->order != 4) {
    if (err) snprintf(err, errlen, "PME order %d unsupported (4 only)", p->order);
    free(c);
    return NULL;
  }
  if (p->terms & MOIL_TERM_ARITH)
    for (i = 0; i < t->ntors; i++)
      if (t->period[i] < 1 || t->period[i] > 4) {
        if (err) snprintf(err, errlen, "arith torsion %d with period %d", i + 1, t->period[i]);
        free(c);
        return NULL;
      }
  cutoff = sqrt(p->cutele2);
  c->beta = find_ewaldcof(cutoff, p->dtol);
  /* erftbdns = 0.5*perft with perft=10000 (symm/ewald_init.f:103-104) */
  c->erftbdns = 5000.0;
  c->del = 1.0 / c->erftbdns;
  c->ntab = (int)(c->beta * c->erftbdns * cutoff * 1.5);
  c->erf_arr = (double *)calloc((size_t)4 * c->ntab, sizeof(double));
  fac = 2.0 / sqrt(3.14159265358979323846);
  for (i = 1; i <= c->ntab; i++) {
    double x = c->del * (i - 1);
    c->erf_arr[4 * (i - 1)] = cody_erfc(x);
  }
  c->erf_arr[1] = -fac;
  {
    double x = (c->ntab - 1) * c->del;
    c->erf_arr[4 * (c->ntab - 1) + 1] = -fac * exp(-x * x);
  }
  clamped_spline(c->erf_arr, c->ntab, c->del);
  /* e_self, symm/ewald_init.f:131-150 */
  {
    double pi = 4.0 * atan(1.0), s = 0;
    spi = 1.0 / pi;
    spi = sqrt(spi);
    epstmp = KOFDIE / p->eps;
    for (i = 0; i < t->npt; i++) s += t->ptchg[i] * t->ptchg[i];
    c->e_self = -(c->beta * spi * epstmp * s);
  }
  c->bsp[0] = (double *)malloc(sizeof(double) * p->nfft1);
  c->bsp[1] = (double *)malloc(sizeof(double) * p->nfft2);
  c->bsp[2] = (double *)malloc(sizeof(double) * p->nfft3);
  bsp_moduli(c->bsp[0], p->nfft1, p->order);
  bsp_moduli(c->bsp[1], p->nfft2, p->order);
  bsp_moduli(c->bsp[2], p->nfft3, p->order);
  /* symm/syminit.f:80-164 */
  {
    static const int ops[13][3] = {{1, 0, 1}, {1, 0, 0}, {1, 0, -1}, {0, 0, 1},
                                   {1, 1, 1}, {1, 1, 0}, {1, 1, -1}, {0, 1, 1},
                                   {0, 1, 0}, {0, 1, -1}, {1, -1, 1}, {1, -1, 0},
                                   {1, -1, -1}};
    int s, k;
    for (s = 0; s < 13; s++)
      for (k = 0; k < 3; k++) c->symop[s + 1][k] = ops[s][k];
  }
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  free(c->patom); free(c->ordlist); free(c->pmol); free(c->ptms); free(c->massmol);
  free(c->xmol); free(c->xrel);
  free(c->poipt); free(c->monoflag); free(c->flagchr); free(c->exc1); free(c->exc2);
  free(c->ptchg); free(c->epsgm6); free(c->epsgm12);
  free(c->ib1); free(c->ib2); free(c->kbond); free(c->req);
  free(c->ia1); free(c->ia2); free(c->ia3); free(c->kangl); free(c->angleq);
  free(c->it1); free(c->it2); free(c->it3); free(c->it4); free(c->period);
  free(c->kt1); free(c->kt2); free(c->kt3); free(c->ph1); free(c->ph2); free(c->ph3);
  free(c->ii1); free(c->ii2); free(c->ii3); free(c->ii4); free(c->kimp); free(c->impeq);
  free(c->sp1); free(c->sp2); free(c->p14); free(c->idxtip3);
  free(c->erf_arr); free(c->bsp[0]); free(c->bsp[1]); free(c->bsp[2]);
  free(c->mono_cent); free(c->indxwat); free(c->poinwat1); free(c->poinwat2);
  iv_free(&c->listwt1); iv_free(&c->listwt2);
  iv_free(&c->symmwat1); iv_free(&c->psymwt1); iv_free(&c->symmwat2); iv_free(&c->psymwt2);
  free(c->point1); free(c->point2); free(c->point3);
  iv_free(&c->list1); iv_free(&c->list2); iv_free(&c->list3);
  iv_free(&c->symreal1); iv_free(&c->psym1); iv_free(&c->symreal2); iv_free(&c->psym2);
  iv_free(&c->symreal3); iv_free(&c->psym3);
  free(c);
}

int orc_ewald_info(orc_ctx *c, double *beta, double *e_self, int *ntab, double *tab,
                   int cap_rows, double *bsp1, double *bsp2, double *bsp3) {
  if (beta) *beta = c->beta;
  if (e_self) *e_self = c->e_self;
  if (ntab) *ntab = c->ntab;
  if (tab) memcpy(tab, c->erf_arr, sizeof(double) * 4 * (cap_rows < c->ntab ? cap_rows : c->ntab));
  if (bsp1) memcpy(bsp1, c->bsp[0], sizeof(double) * c->p.nfft1);
  if (bsp2) memcpy(bsp2, c->bsp[1], sizeof(double) * c->p.nfft2);
  if (bsp3) memcpy(bsp3, c->bsp[2], sizeof(double) * c->p.nfft3);
  return 0;
}

/* ------------------------------------------------------------------ */
/* squeeze, symm/squeeze.f:1-147 (box [-L/2, L/2], one wrap per axis)  */
void orc_squeeze(orc_ctx *c, double *coor) {
  double a = c->p.a, b = c->p.b, cc = c->p.c;
  double xlow = -a * 0.5, xhigh = a * 0.5, ylow = -b * 0.5, yhigh = b * 0.5;
  double zlow = -cc * 0.5, zhigh = cc * 0.5;
  int i, k;
  for (i = 1; i <= c->totmon; i++) {
    int f = c->monoflag[i - 1];
    int cm = (f & MOIL_MONO_SQZ_CENT) != 0;
    int istart = c->poipt[i - 1] + 1, iend = c->poipt[i];
    double xcm, ycm, zcm;
    if (!(cm || (f & MOIL_MONO_SQZ_FIRST))) continue;
    if (cm) {
      double xn;
      xcm = ycm = zcm = 0;
      for (k = istart; k <= iend; k++) {
        xcm += coor[3 * (k - 1)];
        ycm += coor[3 * (k - 1) + 1];
        zcm += coor[3 * (k - 1) + 2];
      }
      xn = 1.0 / (double)(iend - istart + 1);
      xcm *= xn; ycm *= xn; zcm *= xn;
    } else {
      xcm = coor[3 * (istart - 1)];
      ycm = coor[3 * (istart - 1) + 1];
      zcm = coor[3 * (istart - 1) + 2];
    }
    if (ycm < ylow) for (k = istart; k <= iend; k++) coor[3 * (k - 1) + 1] += b;
    if (ycm > yhigh) for (k = istart; k <= iend; k++) coor[3 * (k - 1) + 1] -= b;
    if (xcm < xlow) for (k = istart; k <= iend; k++) coor[3 * (k - 1)] += a;
    if (zcm < zlow) for (k = istart; k <= iend; k++) coor[3 * (k - 1) + 2] += cc;
    if (xcm > xhigh) for (k = istart; k <= iend; k++) coor[3 * (k - 1)] -= a;
    if (zcm > zhigh) for (k = istart; k <= iend; k++) coor[3 * (k - 1) + 2] -= cc;
  }
}

/* ------------------------------------------------------------------ */
/* nbondm + nbond (pot/nbondm.f:1-247, pot/nbond.f:1-262)              */

#define X(i, k) coor[3 * ((i) - 1) + (k)]

static int is_water(const orc_ctx *c, int mono) { return (c->monoflag[mono - 1] & MOIL_MONO_WATER) != 0; }

static void nbond(orc_ctx *c, const double *coor, const int32_t *pointm, const int32_t *nblistm) {
  int totdmon = c->totmon, i, l, j, m;
  double cutvbig2 = c->p.cutvbig2, cutebig2 = c->p.cutebig2;
  /* flag2 = (.not.arith) .and. ... (pot/nbond.f:115-117) */
  int hvdw0 = (c->p.terms & MOIL_TERM_HVDW0) != 0 && !(c->p.terms & MOIL_TERM_ARITH);
  const int32_t *dpoipt = c->poipt;
  c->point1[0] = c->point2[0] = c->point3[0] = 0;
  for (i = 1; i <= totdmon; i++) {
    int ki = dpoipt[i - 1] + 1;
    int jbeg = pointm[i], jend = pointm[i + 1] - 1;
    for (l = ki; l <= dpoipt[i]; l++) {
      for (j = jbeg; j <= jend; j++) {
        int k = nblistm[j];
        int kj = (i != k) ? dpoipt[k - 1] + 1 : l + 1;
        for (m = kj; m <= dpoipt[k]; m++) {
          int jx, excluded = 0;
          double rx, ry, rz, r2;
          for (jx = c->exc1[l - 1] + 1; jx <= c->exc1[l]; jx++)
            if (c->exc2[jx - 1] == m) { excluded = 1; break; }
          if (excluded) continue;
          rx = X(l, 0) - X(m, 0);
          ry = X(l, 1) - X(m, 1);
          rz = X(l, 2) - X(m, 2);
          r2 = rx * rx + ry * ry + rz * rz;
          if (hvdw0) {
            int flag2 = (r2 < cutebig2) && (c->epsgm12[l - 1] < 1e-3 || c->epsgm12[m - 1] < 1e-3);
            if (flag2 && c->flagchr[l - 1] && c->flagchr[m - 1]) { iv_push(&c->list3, m); continue; }
            if (flag2) continue;
          }
          if (r2 < cutvbig2) {
            if (c->flagchr[l - 1] && c->flagchr[m - 1]) iv_push(&c->list1, m);
            else iv_push(&c->list2, m);
          } else if (r2 < cutebig2) {
            if (c->flagchr[l - 1] && c->flagchr[m - 1]) iv_push(&c->list3, m);
          }
        }
      }
      c->point1[l] = (int32_t)c->list1.n;
      c->point2[l] = (int32_t)c->list2.n;
      c->point3[l] = (int32_t)c->list3.n;
    }
  }
}

static void nbondm(orc_ctx *c, const double *coor) {
  int totdmon = c->totmon, i, j, k, nlist = 0, iwat = 0, nflag = 0;
  int32_t *pointm = (int32_t *)calloc((size_t)totdmon + 3, sizeof(int32_t));
  ivec nblistm = {0, 0, 0};
  double cutmono2 = c->p.cutmono2, cutvbig2 = c->p.cutvbig2, cutebig2 = c->p.cutebig2;
  iv_push(&nblistm, 0); /* 1-based */
  /* monomer centres, pot/nbondm.f:79-124 */
  for (i = 1; i <= totdmon; i++) {
    if (is_water(c, i)) {
      j = c->poipt[i] - 2;
      for (k = 0; k < 3; k++) c->mono_cent[3 * (i - 1) + k] = X(j, k);
      continue;
    }
    {
      int ls = c->poipt[i - 1] + 1, le = c->poipt[i];
      double r2;
      for (k = 0; k < 3; k++) c->mono_cent[3 * (i - 1) + k] = 0.0;
      for (j = ls; j <= le; j++)
        for (k = 0; k < 3; k++) c->mono_cent[3 * (i - 1) + k] += X(j, k);
      r2 = 1.0 / (double)(le - ls + 1);
      for (k = 0; k < 3; k++) c->mono_cent[3 * (i - 1) + k] *= r2;
    }
  }
  c->poinwat1[0] = c->poinwat2[0] = 0;
  for (i = 1; i <= totdmon; i++) {
    int watflag = is_water(c, i);
    const double *ci = &c->mono_cent[3 * (i - 1)];
    nlist++;
    pointm[i] = nlist;
    if (nlist >= nblistm.n) iv_push(&nblistm, i); else nblistm.v[nlist] = i;
    if (watflag) {
      if (nflag) {
        nlist--;
        pointm[i] = nlist;
        nblistm.v[nlist] = i;
      }
      iwat++;
      c->indxwat[iwat] = i;
      c->poinwat1[iwat] = (int32_t)c->listwt1.n;
      c->poinwat2[iwat] = (int32_t)c->listwt2.n;
      nflag = 1;
    } else {
      nflag = 0;
    }
    if (i == totdmon) continue;
    for (j = i + 1; j <= totdmon; j++) {
      const double *cj = &c->mono_cent[3 * (j - 1)];
      double rxyz, r2;
      rxyz = ci[0] - cj[0];
      r2 = rxyz * rxyz;
      if (r2 > cutmono2) continue;
      rxyz = ci[1] - cj[1];
      r2 = r2 + rxyz * rxyz;
      if (r2 > cutmono2) continue;
      rxyz = ci[2] - cj[2];
      r2 = r2 + rxyz * rxyz;
      if (watflag && is_water(c, j)) {
        if (r2 <= cutvbig2) iv_push(&c->listwt1, j);
        else if (r2 <= cutebig2) iv_push(&c->listwt2, j);
      } else if (r2 <= cutmono2) {
        nlist++;
        if (nlist >= nblistm.n) iv_push(&nblistm, j); else nblistm.v[nlist] = j;
        nflag = 0;
      }
    }
  }
  pointm[totdmon + 1] = nflag ? nlist : nlist + 1;
  c->nwat = iwat;
  c->poinwat1[iwat + 1] = (int32_t)c->listwt1.n;
  c->poinwat2[iwat + 1] = (int32_t)c->listwt2.n;
  c->nreal_wt1 = (int)c->listwt1.n;
  c->nreal_wt2 = (int)c->listwt2.n;
  nbond(c, coor, pointm, nblistm.v);
  c->nreal1 = (int)c->list1.n;
  c->nreal2 = (int)c->list2.n;
  c->nreal3 = (int)c->list3.n;
  free(pointm);
  iv_free(&nblistm);
}

/* nbsym, symm/nbsym.f:1-192 */
static void nbsym(orc_ctx *c, const double *coor, const int32_t *pointm, const int32_t *nblistm,
                  int lsym, int ia, int jb, int kc) {
  double da = ia * c->p.a, db = jb * c->p.b, dc = kc * c->p.c;
  double cutvbig2 = c->p.cutvbig2, cutebig2 = c->p.cutebig2;
  int i, l, j, m, ki = 1, jbeg = 1;
  for (i = 1; i <= c->totmon; i++) {
    int jend = pointm[i];
    for (l = ki; l <= c->poipt[i]; l++) {
      int o1 = 0, o2 = 0, o3 = 0;
      double xl = X(l, 0) + da, yl = X(l, 1) + db, zl = X(l, 2) + dc;
      for (j = jbeg; j <= jend; j++) {
        int k = nblistm[j];
        int kj = (k == 1) ? 1 : c->poipt[k - 1] + 1;
        for (m = kj; m <= c->poipt[k]; m++) {
          double rx = xl - X(m, 0), ry = yl - X(m, 1), rz = zl - X(m, 2);
          double r2 = rx * rx + ry * ry + rz * rz;
          int flag2 = !(c->p.terms & MOIL_TERM_ARITH) &&
                      (c->epsgm12[l - 1] < 1e-3 || c->epsgm12[m - 1] < 1e-3) && (r2 < cutebig2);
          int chg = c->flagchr[l - 1] && c->flagchr[m - 1];
          if (flag2 && chg) { o3 = 1; iv_push(&c->list3, m); continue; }
          if (flag2) continue;
          if (r2 < cutvbig2) {
            if (chg) { o1 = 1; iv_push(&c->list1, m); }
            else { o2 = 1; iv_push(&c->list2, m); }
          } else if (r2 < cutebig2) {
            if (chg) { o3 = 1; iv_push(&c->list3, m); }
          }
        }
      }
      if (o1) { iv_push(&c->symreal1, l); iv_push(&c->psym1, 0); }
      if (o2) { iv_push(&c->symreal2, l); iv_push(&c->psym2, 0); }
      if (o3) { iv_push(&c->symreal3, l); iv_push(&c->psym3, 0); }
      if (c->symreal1.n) c->psym1.v[c->psym1.n - 1] = (int32_t)c->list1.n;
      if (c->symreal2.n) c->psym2.v[c->psym2.n - 1] = (int32_t)c->list2.n;
      if (c->symreal3.n) c->psym3.v[c->psym3.n - 1] = (int32_t)c->list3.n;
    }
    ki = c->poipt[i] + 1;
    jbeg = pointm[i] + 1;
  }
  c->iblock1[lsym] = (int)c->symreal1.n;
  c->iblock2[lsym] = (int)c->symreal2.n;
  c->iblock3[lsym] = (int)c->symreal3.n;
}

/* nbmsym, symm/nbmsym.f:1-351 (no metal walls, one process) */
static void nbmsym(orc_ctx *c, const double *coor, int ia, int jb, int kc, int lsym) {
  int totdmon = c->totmon, i, k, j, total_need_only = 0, nlist = 0;
  double cutmono2 = c->p.cutmono2, cutvbig2 = c->p.cutvbig2, cutebig2 = c->p.cutebig2;
  double large_cutmono = cutmono2;
  double ha = 0.5 * c->p.a, hb = 0.5 * c->p.b, hc = 0.5 * c->p.c;
  double da = ia * c->p.a, db = jb * c->p.b, dc = kc * c->p.c;
  double barx = 0.5 * da, bary = 0.5 * db, barz = 0.5 * dc;
  int lbarx = fabs(barx) > 0, lbary = fabs(bary) > 0, lbarz = fabs(barz) > 0;
  int32_t *need_only = (int32_t *)malloc(sizeof(int32_t) * (totdmon + 1));
  int32_t *pointm = (int32_t *)calloc((size_t)totdmon + 2, sizeof(int32_t));
  ivec nblistm = {0, 0, 0};
  double *csym = (double *)malloc(sizeof(double) * 3 * (totdmon + 1));
  const double *mc = c->mono_cent;
  iv_push(&nblistm, 0);
  for (i = 1; i <= totdmon; i++) {
    if (fabs(mc[3 * (i - 1)]) > ha || fabs(mc[3 * (i - 1) + 1]) > hb ||
        fabs(mc[3 * (i - 1) + 2]) > hc) {
      large_cutmono = cutmono2 * 4.0;
      break;
    }
  }
  for (i = 1; i <= totdmon; i++) {
    if (lbarx) { double cx = mc[3 * (i - 1)] - barx; if (cx * cx > cutmono2) continue; }
    if (lbary) { double cy = mc[3 * (i - 1) + 1] - bary; if (cy * cy > cutmono2) continue; }
    if (lbarz) { double cz = mc[3 * (i - 1) + 2] - barz; if (cz * cz > cutmono2) continue; }
    need_only[++total_need_only] = i;
  }
  for (i = 1; i <= totdmon; i++) {
    csym[3 * i] = mc[3 * (i - 1)] + da;
    csym[3 * i + 1] = mc[3 * (i - 1) + 1] + db;
    csym[3 * i + 2] = mc[3 * (i - 1) + 2] + dc;
  }
  pointm[0] = 0;
  for (i = 1; i <= totdmon; i++) {
    double xch = fabs(csym[3 * i]) - ha, ych = fabs(csym[3 * i + 1]) - hb;
    double zch = fabs(csym[3 * i + 2]) - hc, rch2;
    int watflag, o1 = 0, o2 = 0;
    if (xch < 0.0) xch = 0.0;
    if (ych < 0.0) ych = 0.0;
    if (zch < 0.0) zch = 0.0;
    rch2 = xch * xch + ych * ych + zch * zch;
    if (rch2 > large_cutmono) { pointm[i] = nlist; continue; }
    watflag = is_water(c, i);
    for (k = 1; k <= total_need_only; k++) {
      double rx, r2;
      j = need_only[k];
      rx = csym[3 * i] - mc[3 * (j - 1)];
      r2 = rx * rx;
      if (r2 > cutmono2) continue;
      rx = csym[3 * i + 1] - mc[3 * (j - 1) + 1];
      r2 = r2 + rx * rx;
      if (r2 > cutmono2) continue;
      rx = csym[3 * i + 2] - mc[3 * (j - 1) + 2];
      r2 = r2 + rx * rx;
      if (watflag && is_water(c, j)) {
        if (r2 < cutvbig2) { o1 = 1; iv_push(&c->listwt1, j); }
        else if (r2 < cutebig2) { o2 = 1; iv_push(&c->listwt2, j); }
      } else if (r2 < cutmono2) {
        nlist++;
        iv_push(&nblistm, j);
      }
    }
    pointm[i] = nlist;
    if (watflag) {
      if (o1) { iv_push(&c->symmwat1, i); iv_push(&c->psymwt1, 0); }
      if (o2) { iv_push(&c->symmwat2, i); iv_push(&c->psymwt2, 0); }
      if (c->symmwat1.n) c->psymwt1.v[c->psymwt1.n - 1] = (int32_t)c->listwt1.n;
      if (c->symmwat2.n) c->psymwt2.v[c->psymwt2.n - 1] = (int32_t)c->listwt2.n;
    }
  }
  c->iblckwt1[lsym] = (int)c->symmwat1.n;
  c->iblckwt2[lsym] = (int)c->symmwat2.n;
  nbsym(c, coor, pointm, nblistm.v, lsym, ia, jb, kc);
  free(need_only); free(pointm); free(csym);
  iv_free(&nblistm);
}

int orc_build_lists(orc_ctx *c, const double *coor) {
  int s;
  free(c->mono_cent); free(c->indxwat); free(c->poinwat1); free(c->poinwat2);
  free(c->point1); free(c->point2); free(c->point3);
  iv_free(&c->listwt1); iv_free(&c->listwt2);
  iv_free(&c->symmwat1); iv_free(&c->psymwt1); iv_free(&c->symmwat2); iv_free(&c->psymwt2);
  iv_free(&c->list1); iv_free(&c->list2); iv_free(&c->list3);
  iv_free(&c->symreal1); iv_free(&c->psym1); iv_free(&c->symreal2); iv_free(&c->psym2);
  iv_free(&c->symreal3); iv_free(&c->psym3);
  c->mono_cent = (double *)calloc((size_t)3 * c->totmon + 3, sizeof(double));
  c->indxwat = (int32_t *)calloc((size_t)c->totmon + 2, sizeof(int32_t));
  c->poinwat1 = (int32_t *)calloc((size_t)c->totmon + 3, sizeof(int32_t));
  c->poinwat2 = (int32_t *)calloc((size_t)c->totmon + 3, sizeof(int32_t));
  c->point1 = (int32_t *)calloc((size_t)c->npt + 1, sizeof(int32_t));
  c->point2 = (int32_t *)calloc((size_t)c->npt + 1, sizeof(int32_t));
  c->point3 = (int32_t *)calloc((size_t)c->npt + 1, sizeof(int32_t));
  nbondm(c, coor);
  /* symm/syminit.f:170-174 */
  c->iblckwt1[0] = c->iblckwt2[0] = 0;
  c->iblock1[0] = c->iblock2[0] = c->iblock3[0] = 0;
  for (s = 1; s <= 13; s++) nbmsym(c, coor, c->symop[s][0], c->symop[s][1], c->symop[s][2], s);
  c->have_lists = 1;
  return 0;
}

/* Canonical pair rows (see moil_gpu.h): water (i, j, symop, list),
 * atoms (l, m, symop, list). */
int64_t orc_dump_water_pairs(orc_ctx *c, int32_t *rows, int64_t cap) {
  int64_t n = 0;
  int iw, s, w, kind;
  for (kind = 1; kind <= 2; kind++) {
    const int32_t *poin = kind == 1 ? c->poinwat1 : c->poinwat2;
    const ivec *lw = kind == 1 ? &c->listwt1 : &c->listwt2;
    const ivec *sw = kind == 1 ? &c->symmwat1 : &c->symmwat2;
    const ivec *pw = kind == 1 ? &c->psymwt1 : &c->psymwt2;
    const int *ib = kind == 1 ? c->iblckwt1 : c->iblckwt2;
    int nreal = kind == 1 ? c->nreal_wt1 : c->nreal_wt2;
    for (iw = 1; iw <= c->nwat; iw++)
      for (int64_t k = poin[iw]; k < poin[iw + 1]; k++) {
        if (rows && n < cap) {
          rows[4 * n] = c->indxwat[iw]; rows[4 * n + 1] = lw->v[k];
          rows[4 * n + 2] = 0; rows[4 * n + 3] = kind;
        }
        n++;
      }
    {
      int64_t jb = nreal;
      for (s = 1; s <= 13; s++)
        for (w = ib[s - 1]; w < ib[s]; w++) {
          for (int64_t k = jb; k < pw->v[w]; k++) {
            if (rows && n < cap) {
              rows[4 * n] = sw->v[w]; rows[4 * n + 1] = lw->v[k];
              rows[4 * n + 2] = s; rows[4 * n + 3] = kind;
            }
            n++;
          }
          jb = pw->v[w];
        }
    }
  }
  return n;
}

int64_t orc_dump_atom_pairs(orc_ctx *c, int32_t *rows, int64_t cap) {
  int64_t n = 0;
  int kind, l, s, w;
  for (kind = 1; kind <= 3; kind++) {
    const int32_t *pt = kind == 1 ? c->point1 : kind == 2 ? c->point2 : c->point3;
    const ivec *ls = kind == 1 ? &c->list1 : kind == 2 ? &c->list2 : &c->list3;
    const ivec *sr = kind == 1 ? &c->symreal1 : kind == 2 ? &c->symreal2 : &c->symreal3;
    const ivec *ps = kind == 1 ? &c->psym1 : kind == 2 ? &c->psym2 : &c->psym3;
    const int *ib = kind == 1 ? c->iblock1 : kind == 2 ? c->iblock2 : c->iblock3;
    int nreal = kind == 1 ? c->nreal1 : kind == 2 ? c->nreal2 : c->nreal3;
    for (l = 1; l <= c->npt; l++)
      for (int64_t k = pt[l - 1]; k < pt[l]; k++) {
        if (rows && n < cap) {
          rows[4 * n] = l; rows[4 * n + 1] = ls->v[k]; rows[4 * n + 2] = 0; rows[4 * n + 3] = kind;
        }
        n++;
      }
    {
      int64_t jb = nreal;
      for (s = 1; s <= 13; s++)
        for (w = ib[s - 1]; w < ib[s]; w++) {
          for (int64_t k = jb; k < ps->v[w]; k++) {
            if (rows && n < cap) {
              rows[4 * n] = sr->v[w]; rows[4 * n + 1] = ls->v[k];
              rows[4 * n + 2] = s; rows[4 * n + 3] = kind;
            }
            n++;
          }
          jb = ps->v[w];
        }
    }
  }
  return n;
}

/* ------------------------------------------------------------------ */
/* energy terms                                                        */

typedef struct { double e_bond, e_theta, e_tors, e_imp, e_vdw, e_el, e_vsym, e_lsym,
                 e_el14, e_vdw14, e_corr, e_ew_receip; } eterms;

/* cubic-spline erfc: pot/cdie_ewald.f:152-160.  MOIL reads row
 * int(erftbdns x) without a bound; past the table's last row (a distance
 * MOIL's lists reach only with stale lists or split molecules, where MOIL
 * reads beyond its array) the last row is read at its right end, the
 * library's definition of that case. */
static inline void spline_erfc(const orc_ctx *c, double xerfc, double *erfcc, double *derfc) {
  int idx;
  double dx;
  if (c->erftbdns * xerfc < (double)c->ntab) {
    idx = (int)(c->erftbdns * xerfc);
    dx = xerfc - idx * c->del;
  } else {
    idx = c->ntab - 1;
    dx = c->del;
  }
  const double *t = &c->erf_arr[4 * idx];
  *derfc = -t[1] - dx * (t[2] + 0.5 * dx * t[3]);
  *erfcc = dx * (t[1] + 0.5 * dx * (t[2] + dx * t[3] * (1.0 / 3.0))) + t[0];
}

/* ebond, pot/ebond.f:33-72 */
static void ebond(const orc_ctx *c, const double *coor, double *dpot, eterms *E) {
  int mm;
  for (mm = 0; mm < c->nb; mm++) {
    int i = c->ib1[mm], j = c->ib2[mm];
    double rx = X(i, 0) - X(j, 0), ry = X(i, 1) - X(j, 1), rz = X(i, 2) - X(j, 2);
    double r2 = rx * rx + ry * ry + rz * rz, s = sqrt(r2), r = 2.0 / s;
    double db = s - c->req[mm], df = c->kbond[mm] * db, e = df * db;
    E->e_bond += e;
    df = df * r;
    dpot[3 * (i - 1)] += rx * df; dpot[3 * (i - 1) + 1] += ry * df; dpot[3 * (i - 1) + 2] += rz * df;
    dpot[3 * (j - 1)] -= rx * df; dpot[3 * (j - 1) + 1] -= ry * df; dpot[3 * (j - 1) + 2] -= rz * df;
  }
}

/* etheta, pot/etheta.f:25-113 */
static void etheta(const orc_ctx *c, const double *coor, double *dpot, eterms *E) {
  int n, k;
  for (n = 0; n < c->nangl; n++) {
    int i = c->ia1[n], j = c->ia2[n], l = c->ia3[n];
    double di[3], dj[3], ri, rj, cst, at, da, df, s;
    for (k = 0; k < 3; k++) { di[k] = X(i, k) - X(j, k); dj[k] = X(l, k) - X(j, k); }
    ri = sqrt(di[0] * di[0] + di[1] * di[1] + di[2] * di[2]);
    rj = sqrt(dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2]);
    for (k = 0; k < 3; k++) { di[k] /= ri; dj[k] /= rj; }
    cst = di[0] * dj[0] + di[1] * dj[1] + di[2] * dj[2];
    if (cst > 1.0) cst = 1.0;
    if (cst < -1.0) cst = -1.0;
    at = acos(cst);
    da = at - c->angleq[n];
    df = c->kangl[n] * da;
    E->e_theta += df * da;
    df = df + df;
    s = sin(at);
    if (s < 1e-6 && s > 0.0) s = 1e-6;
    else if (s < 0.0 && s > -1e-6) s = -1e-6;
    df = -df / s;
    for (k = 0; k < 3; k++) {
      double t1 = (dj[k] - cst * di[k]) / ri, t2 = (di[k] - cst * dj[k]) / rj;
      dpot[3 * (i - 1) + k] += df * t1;
      dpot[3 * (j - 1) + k] -= df * (t1 + t2);
      dpot[3 * (l - 1) + k] += df * t2;
    }
  }
}

/* gradient of cos(phi) for the dihedral i-j-k-l, phi defined through
 * a=x_j-x_i, b=x_k-x_j, c=x_l-x_k and cos(phi)=((a.b)(b.c)-(a.c)(b.b))/
 * sqrt(((a.a)(b.b)-(a.b)^2)((b.b)(c.c)-(b.c)^2)) (pot/ephi.f:88-109).
 * Derived here by the chain rule on u=(a.a)(b.b)-(a.b)^2,
 * v=(b.b)(c.c)-(b.c)^2, w=(a.b)(b.c)-(a.c)(b.b). */
static double dihedral_cos_grad(const double *xi, const double *xj, const double *xk,
                                const double *xl, double g[4][3], double *sgn) {
  double a[3], b[3], cc[3], aa, bb, ccc, ab, bc, ac, u, v, w, co, inv;
  double dw_da[3], dw_db[3], dw_dc[3], du_da[3], du_db[3], dv_db[3], dv_dc[3];
  int k;
  for (k = 0; k < 3; k++) { a[k] = xj[k] - xi[k]; b[k] = xk[k] - xj[k]; cc[k] = xl[k] - xk[k]; }
  aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  ccc = cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2];
  ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  bc = b[0] * cc[0] + b[1] * cc[1] + b[2] * cc[2];
  ac = a[0] * cc[0] + a[1] * cc[1] + a[2] * cc[2];
  u = aa * bb - ab * ab;
  v = bb * ccc - bc * bc;
  w = ab * bc - ac * bb;
  inv = 1.0 / sqrt(u * v);
  co = w * inv;
  for (k = 0; k < 3; k++) {
    dw_da[k] = bc * b[k] - bb * cc[k];
    dw_db[k] = ab * cc[k] + bc * a[k] - 2.0 * ac * b[k];
    dw_dc[k] = ab * b[k] - bb * a[k];
    du_da[k] = 2.0 * (bb * a[k] - ab * b[k]);
    du_db[k] = 2.0 * (aa * b[k] - ab * a[k]);
    dv_db[k] = 2.0 * (ccc * b[k] - bc * cc[k]);
    dv_dc[k] = 2.0 * (bb * cc[k] - bc * b[k]);
  }
  /* d co = dw*inv - 0.5*co*(du/u + dv/v) */
  for (k = 0; k < 3; k++) {
    double gA = dw_da[k] * inv - 0.5 * co * (du_da[k] / u);
    double gB = dw_db[k] * inv - 0.5 * co * (du_db[k] / u + dv_db[k] / v);
    double gC = dw_dc[k] * inv - 0.5 * co * (dv_dc[k] / v);
    /* a = xj - xi, b = xk - xj, c = xl - xk */
    g[0][k] = -gA;
    g[1][k] = gA - gB;
    g[2][k] = gB - gC;
    g[3][k] = gC;
  }
  if (sgn) {
    /* sign of phi as pot/eimphi.f:130-143: ((a x b) x (b x c)) . b */
    double uu[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double vv[3] = {b[1] * cc[2] - b[2] * cc[1], b[2] * cc[0] - b[0] * cc[2], b[0] * cc[1] - b[1] * cc[0]};
    double dx[3] = {uu[1] * vv[2] - uu[2] * vv[1], uu[2] * vv[0] - uu[0] * vv[2], uu[0] * vv[1] - uu[1] * vv[0]};
    *sgn = dx[0] * b[0] + dx[1] * b[1] + dx[2] * b[2];
  }
  return co;
}

/* etors, pot/ephi.f:66-277 */
static void etors(const orc_ctx *c, const double *coor, double *dpot, eterms *E) {
  int n, k, q;
  for (n = 0; n < c->ntors; n++) {
    int id[4] = {c->it1[n], c->it2[n], c->it3[n], c->it4[n]};
    double g[4][3], co, e, df;
    co = dihedral_cos_grad(&X(id[0], 0), &X(id[1], 0), &X(id[2], 0), &X(id[3], 0), g, NULL);
    if (c->p.terms & MOIL_TERM_ARITH) {
      /* arith: one cosine of period 1..4 (pot/ephi.f:113-138, 238-259) */
      double k2 = c->kt2[n], h2 = c->ph2[n];
      switch (c->period[n]) {
        case 1: e = k2 + h2 * k2 * co; df = k2 * h2; break;
        case 2: e = k2 + h2 * k2 * (2.0 * co * co - 1.0); df = 4.0 * k2 * h2 * co; break;
        case 3:
          e = k2 + h2 * k2 * co * (4.0 * co * co - 3.0);
          df = k2 * h2 * (12.0 * co * co - 3.0);
          break;
        case 4:
          e = k2 + h2 * k2 * (2.0 * (2.0 * co * co - 1.0) * (2.0 * co * co - 1.0) - 1.0);
          df = 16.0 * k2 * h2 * co * (2.0 * co * co - 1.0);
          break;
        default: /* MOIL leaves e and df unset; the library refuses these */
          fprintf(stderr, "oracle: arith torsion %d with period %d\n", n + 1, c->period[n]);
          abort();
      }
    } else {
      e = (c->kt1[n] + c->kt2[n] + c->kt3[n]) + c->ph1[n] * (c->kt1[n] * co) +
          c->ph2[n] * (c->kt2[n] * (2.0 * co * co - 1.0)) +
          c->ph3[n] * (c->kt3[n] * co * (4.0 * co * co - 3.0));
      df = c->ph1[n] * c->kt1[n] + c->ph2[n] * c->kt2[n] * 4.0 * co +
           c->ph3[n] * c->kt3[n] * (12.0 * co * co - 3.0);
    }
    E->e_tors += e;
    for (q = 0; q < 4; q++)
      for (k = 0; k < 3; k++) dpot[3 * (id[q] - 1) + k] += df * g[q][k];
  }
}

/* eimphi, pot/eimphi.f:86-293 */
static void eimphi(const orc_ctx *c, const double *coor, double *dpot, eterms *E) {
  int n, k, q;
  double pi = 4.0 * atan(1.0);
  for (n = 0; n < c->nimp; n++) {
    const int arith = (c->p.terms & MOIL_TERM_ARITH) != 0;
    /* arith: i = iimp4, j = iimp2, k = iimp1, l = iimp3 (pot/eimphi.f:98-115) */
    int id[4] = {arith ? c->ii4[n] : c->ii1[n], c->ii2[n], arith ? c->ii1[n] : c->ii3[n],
                 arith ? c->ii3[n] : c->ii4[n]};
    double g[4][3], co, e, df, sgn;
    co = dihedral_cos_grad(&X(id[0], 0), &X(id[1], 0), &X(id[2], 0), &X(id[3], 0), g, &sgn);
    if (co > 1.0) co = 1.0;
    if (co < -1.0) co = -1.0;
    if (arith) { /* k - k cos(2 phi), pot/eimphi.f:251-262 */
      e = c->kimp[n] - c->kimp[n] * (2.0 * co * co - 1.0);
      df = -4.0 * c->kimp[n] * co;
    } else if (fabs(c->impeq[n]) < 1e-6) {
      df = -2.0 * c->kimp[n];
      e = -df * (1.0 - co);
    } else if (fabs(fabs(c->impeq[n]) - pi) < 1e-6) {
      df = 2.0 * c->kimp[n];
      e = df * (1.0 + co);
    } else {
      double phi = acos(co), delta, yy;
      if (sgn < 0) phi = -phi;
      delta = phi - c->impeq[n];
      if (delta > pi) delta -= 2.0 * pi;
      if (delta < -pi) delta += 2.0 * pi;
      df = c->kimp[n] * delta;
      e = df * delta;
      yy = sin(phi);
      if (fabs(yy) > 1e-6) df = -2.0 * df / yy;
      else if ((phi > 0.0 && phi < pi / 2.0) || (phi < 0.0 && phi > -pi / 2.0)) df = df * 1e6;
      else df = -df * 1e6;
    }
    E->e_imp += e;
    for (q = 0; q < 4; q++)
      for (k = 0; k < 3; k++) dpot[3 * (id[q] - 1) + k] += df * g[q][k];
  }
}

/* ener14, pot/ener14.f:38-95 */
static void ener14(const orc_ctx *c, const double *coor, double *dpot, eterms *E) {
  int n;
  double epstmp = (c->p.terms & MOIL_TERM_EL) ? 1.0 : 0.0;
  double pick = (c->p.terms & MOIL_TERM_VDW) ? 1.0 : 0.0;
  int doel = (c->p.terms & MOIL_TERM_EL14) != 0, dov = (c->p.terms & MOIL_TERM_VDW14) != 0;
  if (!doel && !dov) return;
  for (n = 0; n < c->totspe; n++) {
    int j = c->sp1[n], k = c->sp2[n];
    double rx = X(j, 0) - X(k, 0), ry = X(j, 1) - X(k, 1), rz = X(j, 2) - X(k, 2);
    double r2 = rx * rx + ry * ry + rz * rz, s2 = 1.0 / r2, e1 = 0, e2 = 0, df1 = 0, df2 = 0, df;
    const double *p = &c->p14[3 * n];
    if (p[0] > 1e-12) {
      double s6 = s2 * s2 * s2, aa = p[0] * s6 * s6 * pick, bb = p[1] * s6 * pick;
      e1 = aa - bb;
      df1 = -6.0 * s2 * (aa + e1);
    }
    if (fabs(p[2]) > 1e-12) {
      double s = sqrt(s2);
      e2 = p[2] * s * epstmp;
      df2 = -e2 * s2;
    }
    df = df1 + df2;
    dpot[3 * (j - 1)] += df * rx; dpot[3 * (j - 1) + 1] += df * ry; dpot[3 * (j - 1) + 2] += df * rz;
    dpot[3 * (k - 1)] -= df * rx; dpot[3 * (k - 1) + 1] -= df * ry; dpot[3 * (k - 1) + 2] -= df * rz;
    E->e_vdw14 += e1;
    E->e_el14 += e2;
  }
}

/* One real-space pair (atom lists).  kind: 1 = list1 (LJ if r2<=cutvdw2,
 * Coulomb always), 2 = list2 (LJ if r2<=cutvdw2), 3 = list3 (Coulomb if
 * r2<=cutele2).  Returns the energies via ev, ee and the scalar df; with
 * vir, also the LJ part of df (vir[0]) and MOIL's V_PIdir of the pair
 * (vir[1], pot/cdie_ewald.f:351-355: its own exp, not the table's
 * derivative; 0 without a Coulomb term). */
static inline int pair_atoms(const orc_ctx *c, int kind, double r2, double qq, double A, double B,
                             double *ev, double *ee, double *df, double *vir) {
  double rij, s, s2 = 1.0 / r2, erfcc, derfc;
  *ev = *ee = *df = 0;
  if (vir) vir[0] = vir[1] = 0;
  if (kind == 2) {
    double s6;
    if (r2 > c->p.cutvdw2) return 0;
    s6 = s2 * s2 * s2;
    A = A * s6 * s6;
    B = B * s6;
    *ev = A - B;
    *df = -6.0 * s2 * (A + *ev);
    if (vir) vir[0] = *df;
    return 1;
  }
  if (kind == 3 && r2 > c->p.cutele2) return 0;
  rij = sqrt(r2);
  s = 1.0 / rij;
  spline_erfc(c, c->beta * rij, &erfcc, &derfc);
  *ee = qq * s * erfcc;
  *df = -qq * s * (s2 * erfcc + s * c->beta * derfc);
  if (vir) {
    const double xerfc = c->beta * rij;
    vir[1] = qq * (erfcc * s * s * s +
                   2.0 * c->beta / sqrt(M_PI) * exp(-(xerfc * xerfc)) * s * s);
  }
  if (kind == 1 && r2 <= c->p.cutvdw2) {
    double s6 = s2 * s2 * s2, dfv;
    A = A * s6 * s6;
    B = B * s6;
    *ev = A - B;
    dfv = -6.0 * s2 * (A + *ev);
    if (vir) vir[0] = dfv;
    *df += dfv;
  }
  return 1;
}

/* The virial terms of one pair (pot/cdie_ewald.f:351-368,440-460 for a
 * real pair, symm/symcdie_ewald.f:150-200 for an image): r = x_i + T -
 * x_j, molecular separation rm = M_i + T - M_j written as MOIL writes it
 * (M_i - M_j for a real pair; M_j - M_i - T, with the signs folded in,
 * for an image), dfv the LJ part of df, vpd MOIL's V_PIdir. */
static inline void pair_virial(const orc_ctx *c, double *V, int img, int i, int j, const double *T,
                               const double *r, double dfv, double vpd) {
  const double *mi = c->xmol + 3 * (c->pmol[i - 1] - 1), *mj = c->xmol + 3 * (c->pmol[j - 1] - 1);
  double rm[3];
  int k;
  if (!img) {
    for (k = 0; k < 3; k++) rm[k] = mi[k] - mj[k];
    if (dfv != 0.0) {
      V[0] = V[0] - rm[0] * dfv * r[0] - rm[1] * dfv * r[1] - rm[2] * dfv * r[2];
      for (k = 0; k < 3; k++) V[1 + k] = V[1 + k] - dfv * r[k] * rm[k];
    }
    for (k = 0; k < 3; k++) V[4 + k] = V[4 + k] + vpd * r[k] * rm[k];
  } else {
    for (k = 0; k < 3; k++) rm[k] = mj[k] - mi[k] - T[k];
    if (dfv != 0.0) {
      V[0] = V[0] + rm[0] * dfv * r[0] + rm[1] * dfv * r[1] + rm[2] * dfv * r[2];
      for (k = 0; k < 3; k++) V[1 + k] = V[1 + k] + dfv * r[k] * rm[k];
    }
    for (k = 0; k < 3; k++) V[4 + k] = V[4 + k] + vpd * (-r[k]) * rm[k];
  }
}

/* The LJ coefficients of a pair as pair_atoms takes them (A of s6*s6, B
 * of s6).  Geometric: A = ai*aj, B = bi*bj.  Arith (epsgm6 = sigma):
 * a = ai*aj, b = 0.5*(bi+bj); b1 = b*b*b; b1 = b1*b1; B = a*b1;
 * b1 = b1*b1; A = a*b1, MOIL's products (pot/cdie_ewald.f:377-383,
 * 461-470, 536-548; symm/symcdie_ewald.f:127-140, 301-314). */
static inline void lj_pair(int arith, double ai, double bi, double aj, double bj, double *A,
                           double *B) {
  double a, b, b1;
  if (!arith) {
    *A = ai * aj;
    *B = bi * bj;
    return;
  }
  a = ai * aj;
  b = 0.5 * (bi + bj);
  b1 = b * b * b;
  b1 = b1 * b1;
  *B = a * b1;
  b1 = b1 * b1;
  *A = a * b1;
}

/* cdie_ewald exclusion correction + list loops, pot/cdie_ewald.f:67-673,
 * and symcdie_ewald, symm/symcdie_ewald.f:58-549. */
static void real_atoms(const orc_ctx *c, const double *coor, double *dpot, eterms *E, double *V) {
  double epstmp = (c->p.terms & MOIL_TERM_EL) ? KOFDIE / c->p.eps : 0.0;
  double pick = (c->p.terms & MOIL_TERM_VDW) ? 1.0 : 0.0;
  const int arith = (c->p.terms & MOIL_TERM_ARITH) != 0;
  int i, kind, s;
  /* exclusion correction (e_corr), atoms 1..npt-1 (pot/cdie_ewald.f:83-199) */
  for (i = 1; i <= c->npt - 1; i++) {
    double qi = c->ptchg[i - 1] * epstmp;
    for (int k = c->exc1[i - 1] + 1; k <= c->exc1[i]; k++) {
      int j = c->exc2[k - 1];
      double rx = X(i, 0) - X(j, 0), ry = X(i, 1) - X(j, 1), rz = X(i, 2) - X(j, 2);
      double r2 = rx * rx + ry * ry + rz * rz, s2 = 1.0 / r2, rij = sqrt(r2);
      double q = qi * c->ptchg[j - 1], sv = 1.0 / rij, erfcc, derfc, erf, df;
      spline_erfc(c, c->beta * rij, &erfcc, &derfc);
      erf = 1.0 - erfcc;
      E->e_corr += -q * sv * erf;
      df = q * sv * (s2 * erf - sv * c->beta * derfc);
      dpot[3 * (i - 1)] += df * rx; dpot[3 * (i - 1) + 1] += df * ry; dpot[3 * (i - 1) + 2] += df * rz;
      dpot[3 * (j - 1)] -= df * rx; dpot[3 * (j - 1) + 1] -= df * ry; dpot[3 * (j - 1) + 2] -= df * rz;
    }
  }
  for (kind = 1; kind <= 3; kind++) {
    const int32_t *pt = kind == 1 ? c->point1 : kind == 2 ? c->point2 : c->point3;
    const ivec *ls = kind == 1 ? &c->list1 : kind == 2 ? &c->list2 : &c->list3;
    const ivec *sr = kind == 1 ? &c->symreal1 : kind == 2 ? &c->symreal2 : &c->symreal3;
    const ivec *ps = kind == 1 ? &c->psym1 : kind == 2 ? &c->psym2 : &c->psym3;
    const int *ib = kind == 1 ? c->iblock1 : kind == 2 ? c->iblock2 : c->iblock3;
    int nreal = kind == 1 ? c->nreal1 : kind == 2 ? c->nreal2 : c->nreal3;
    /* real pairs */
    for (i = 1; i <= c->npt; i++) {
      /* arith: bi = epsgm6(i) without pick (pot/cdie_ewald.f:98-104) */
      double qi = c->ptchg[i - 1] * epstmp, ai = c->epsgm12[i - 1] * pick,
             bi = arith ? c->epsgm6[i - 1] : c->epsgm6[i - 1] * pick;
      double d[3] = {0, 0, 0};
      for (int64_t k = pt[i - 1]; k < pt[i]; k++) {
        int j = ls->v[k];
        double rx = X(i, 0) - X(j, 0), ry = X(i, 1) - X(j, 1), rz = X(i, 2) - X(j, 2);
        double r2 = rx * rx + ry * ry + rz * rz, ev, ee, df, vr[2], A, B;
        lj_pair(arith, ai, bi, c->epsgm12[j - 1], c->epsgm6[j - 1], &A, &B);
        if (!pair_atoms(c, kind, r2, qi * c->ptchg[j - 1], A, B, &ev, &ee, &df, V ? vr : NULL))
          continue;
        if (V) {
          const double r[3] = {rx, ry, rz};
          /* arith list1: the LJ virial is only in the geometric branch
           * (pot/cdie_ewald.f:432-473) */
          if (arith && kind == 1) vr[0] = 0.0;
          pair_virial(c, V, 0, i, j, NULL, r, vr[0], vr[1]);
        }
        E->e_vdw += ev;
        E->e_el += ee;
        d[0] += df * rx; d[1] += df * ry; d[2] += df * rz;
        dpot[3 * (j - 1)] -= df * rx; dpot[3 * (j - 1) + 1] -= df * ry; dpot[3 * (j - 1) + 2] -= df * rz;
      }
      dpot[3 * (i - 1)] += d[0]; dpot[3 * (i - 1) + 1] += d[1]; dpot[3 * (i - 1) + 2] += d[2];
    }
    /* image pairs */
    {
      int64_t jb = nreal;
      for (s = 1; s <= 13; s++) {
        double T[3] = {c->symop[s][0] * c->p.a, c->symop[s][1] * c->p.b, c->symop[s][2] * c->p.c};
        for (int w = ib[s - 1]; w < ib[s]; w++) {
          int ir = sr->v[w];
          double xi = X(ir, 0) + T[0], yi = X(ir, 1) + T[1], zi = X(ir, 2) + T[2];
          /* arith: list1 keeps pick on bi, list2 drops it
           * (symm/symcdie_ewald.f:101-102,275-284) */
          double qi = c->ptchg[ir - 1] * epstmp, ai = c->epsgm12[ir - 1] * pick,
                 bi = (arith && kind == 2) ? c->epsgm6[ir - 1] : c->epsgm6[ir - 1] * pick;
          for (int64_t k = jb; k < ps->v[w]; k++) {
            int j = ls->v[k];
            double rx = xi - X(j, 0), ry = yi - X(j, 1), rz = zi - X(j, 2);
            double r2 = rx * rx + ry * ry + rz * rz, ev, ee, df, vr[2], A, B;
            lj_pair(arith, ai, bi, c->epsgm12[j - 1], c->epsgm6[j - 1], &A, &B);
            if (!pair_atoms(c, kind, r2, qi * c->ptchg[j - 1], A, B, &ev, &ee, &df, V ? vr : NULL))
              continue;
            if (V) {
              const double r[3] = {rx, ry, rz};
              pair_virial(c, V, 1, ir, j, T, r, vr[0], vr[1]);
            }
            E->e_vsym += ev;
            E->e_lsym += ee;
            dpot[3 * (ir - 1)] += df * rx; dpot[3 * (ir - 1) + 1] += df * ry; dpot[3 * (ir - 1) + 2] += df * rz;
            dpot[3 * (j - 1)] -= df * rx; dpot[3 * (j - 1) + 1] -= df * ry; dpot[3 * (j - 1) + 2] -= df * rz;
          }
          jb = ps->v[w];
        }
      }
    }
  }
}

/* One water-water group entry: O-O LJ (if r2 <= cutvdw2) and the 9
 * site-pair erfc Coulomb terms; kind 2 entries are skipped whole when
 * the O-O distance exceeds relx (pot/watwat_ewald.f:131-612, 624-1085). */
static void water_pair(const orc_ctx *c, int io, const double *T, int jo, int kind,
                       double qq[3][3], double oo12, double oo6, double *dpot,
                       const double *coor, double *ev, double *ee, double *V, int img) {
  double xi[3][3], r[3], r2;
  int a, b, k;
  for (a = 0; a < 3; a++)
    for (k = 0; k < 3; k++) xi[a][k] = X(io + a, k) + T[k];
  for (k = 0; k < 3; k++) r[k] = xi[0][k] - X(jo, k);
  r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  if (kind == 2 && r2 > c->p.cutele2) return;
  for (a = 0; a < 3; a++)
    for (b = 0; b < 3; b++) {
      double rr[3], q2, s2, rij, s, erfcc, derfc, df = 0, dfv = 0;
      for (k = 0; k < 3; k++) rr[k] = xi[a][k] - X(jo + b, k);
      q2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
      s2 = 1.0 / q2;
      if (a == 0 && b == 0 && kind == 1 && q2 <= c->p.cutvdw2) {
        double s6 = s2 * s2 * s2, A = oo12 * s6 * s6, B = oo6 * s6, e1 = A - B;
        *ev += e1;
        df = -6.0 * s2 * (A + e1);
        dfv = df;
      }
      rij = sqrt(q2);
      s = 1.0 / rij;
      spline_erfc(c, c->beta * rij, &erfcc, &derfc);
      *ee += qq[a][b] * s * erfcc;
      if (V) {
        /* pot/watwat_ewald.f:162-206, symm/symwat_ewald.f:160-203: every
         * site pair with the two waters' molecular separation */
        const double xerfc = c->beta * rij;
        const double vpd = qq[a][b] * (erfcc * s * s * s +
                                       2.0 * c->beta / sqrt(M_PI) * exp(-(xerfc * xerfc)) * s * s);
        pair_virial(c, V, img, io + a, jo + b, T, rr, dfv, vpd);
      }
      df += -qq[a][b] * s * (s2 * erfcc + s * c->beta * derfc);
      for (k = 0; k < 3; k++) {
        dpot[3 * (io + a - 1) + k] += df * rr[k];
        dpot[3 * (jo + b - 1) + k] -= df * rr[k];
      }
    }
}

static void real_waters(const orc_ctx *c, const double *coor, double *dpot, eterms *E, double *V) {
  double epstmp = (c->p.terms & MOIL_TERM_EL) ? 1.0 : 0.0;
  double pick = (c->p.terms & MOIL_TERM_VDW) ? 1.0 : 0.0;
  double qq[3][3], oo12, oo6, zero[3] = {0, 0, 0};
  int io, ih, iw, kind, s;
  if (c->nwaters < 2) return;
  /* parameters of the last water (pot/watwat_ewald.f:40-72) */
  {
    int m = c->idxtip3[c->nwaters - 1];
    io = c->poipt[m] - 2;
    ih = c->poipt[m];
    if (!(c->p.terms & MOIL_TERM_ARITH)) {
      oo6 = pick * c->epsgm6[io - 1] * c->epsgm6[io - 1];
      oo12 = pick * c->epsgm12[io - 1] * c->epsgm12[io - 1];
    } else { /* pick*epsgm12(io)**2*epsgm6(io)**6, **12 (pot/watwat_ewald.f:65-66) */
      double e12 = c->epsgm12[io - 1], sg = c->epsgm6[io - 1];
      oo6 = pick * pow(e12, 2) * pow(sg, 6);
      oo12 = pick * pow(e12, 2) * pow(sg, 12);
    }
    qq[0][0] = epstmp * KOFDIE / c->p.eps * c->ptchg[io - 1] * c->ptchg[io - 1];
    qq[0][1] = qq[0][2] = qq[1][0] = qq[2][0] =
        epstmp * KOFDIE / c->p.eps * c->ptchg[io - 1] * c->ptchg[ih - 1];
    qq[1][1] = qq[1][2] = qq[2][1] = qq[2][2] =
        epstmp * KOFDIE / c->p.eps * c->ptchg[ih - 1] * c->ptchg[ih - 1];
  }
  for (kind = 1; kind <= 2; kind++) {
    const int32_t *poin = kind == 1 ? c->poinwat1 : c->poinwat2;
    const ivec *lw = kind == 1 ? &c->listwt1 : &c->listwt2;
    const ivec *sw = kind == 1 ? &c->symmwat1 : &c->symmwat2;
    const ivec *pw = kind == 1 ? &c->psymwt1 : &c->psymwt2;
    const int *ib = kind == 1 ? c->iblckwt1 : c->iblckwt2;
    int nreal = kind == 1 ? c->nreal_wt1 : c->nreal_wt2;
    /* watwat_ewald loops i=1..my_nwat = nwaters-1 (pot/nbondm.f:224) */
    for (iw = 1; iw <= c->nwat - 1; iw++) {
      int i0 = c->poipt[c->indxwat[iw]] - 2;
      for (int64_t k = poin[iw]; k < poin[iw + 1]; k++) {
        int j0 = c->poipt[lw->v[k]] - 2;
        water_pair(c, i0, zero, j0, kind, qq, oo12, oo6, dpot, coor, &E->e_vdw, &E->e_el, V, 0);
      }
    }
    {
      int64_t jb = nreal;
      for (s = 1; s <= 13; s++) {
        double T[3] = {c->symop[s][0] * c->p.a, c->symop[s][1] * c->p.b, c->symop[s][2] * c->p.c};
        for (int w = ib[s - 1]; w < ib[s]; w++) {
          int i0 = c->poipt[sw->v[w]] - 2;
          for (int64_t k = jb; k < pw->v[w]; k++) {
            int j0 = c->poipt[lw->v[k]] - 2;
            water_pair(c, i0, T, j0, kind, qq, oo12, oo6, dpot, coor, &E->e_vsym, &E->e_lsym, V, 1);
          }
          jb = pw->v[w];
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* PME (pot/ewald_receip.f:8-213, 600-970) with a self-contained FFT  */

typedef struct { double re, im; } cplx;

/* in-place complex DFT of length n with stride, sign -1 forward
 * (FFTPACK cfftf) or +1 backward (cfftb), unnormalised.  Mixed radix
 * recursion on the smallest prime factor, O(n^2) for a prime n. */
static void dft_rec(cplx *x, int n, int sign, cplx *tmp) {
  int p = 2, m, q, r, k;
  if (n == 1) return;
  while (p * p <= n && n % p) p++;
  if (n % p) p = n; /* prime */
  m = n / p;
  /* split into p subsequences x[q + p*r] */
  for (q = 0; q < p; q++)
    for (r = 0; r < m; r++) tmp[q * m + r] = x[q + p * r];
  for (q = 0; q < p; q++) dft_rec(tmp + q * m, m, sign, x);
  for (k = 0; k < n; k++) {
    double sr = 0, si = 0;
    for (q = 0; q < p; q++) {
      double ang = sign * 2.0 * M_PI * (double)((long)q * k % n) / n;
      cplx v = tmp[q * m + (k % m)];
      double cr = cos(ang), ci = sin(ang);
      sr += v.re * cr - v.im * ci;
      si += v.re * ci + v.im * cr;
    }
    x[k].re = sr;
    x[k].im = si;
  }
}

static void fft1(cplx *line, int n, int sign) {
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * n);
  cplx *buf = (cplx *)malloc(sizeof(cplx) * n);
  memcpy(buf, line, sizeof(cplx) * n);
  dft_rec(buf, n, sign, tmp);
  memcpy(line, buf, sizeof(cplx) * n);
  free(tmp);
  free(buf);
}

static void fft3(cplx *Q, int n1, int n2, int n3, int sign) {
  int i, j, k, nm = n1 > n2 ? n1 : n2;
  cplx *line;
  nm = nm > n3 ? nm : n3;
  line = (cplx *)malloc(sizeof(cplx) * nm);
  for (k = 0; k < n3; k++)
    for (j = 0; j < n2; j++) fft1(&Q[((size_t)k * n2 + j) * n1], n1, sign);
  for (k = 0; k < n3; k++)
    for (i = 0; i < n1; i++) {
      for (j = 0; j < n2; j++) line[j] = Q[((size_t)k * n2 + j) * n1 + i];
      fft1(line, n2, sign);
      for (j = 0; j < n2; j++) Q[((size_t)k * n2 + j) * n1 + i] = line[j];
    }
  for (j = 0; j < n2; j++)
    for (i = 0; i < n1; i++) {
      for (k = 0; k < n3; k++) line[k] = Q[((size_t)k * n2 + j) * n1 + i];
      fft1(line, n3, sign);
      for (k = 0; k < n3; k++) Q[((size_t)k * n2 + j) * n1 + i] = line[k];
    }
  free(line);
}

static void pme(const orc_ctx *c, const double *coor, double *dpot, eterms *E, double *V) {
  int n1 = c->p.nfft1, n2 = c->p.nfft2, n3 = c->p.nfft3, order = c->p.order;
  int npt = c->npt, n, a, b, g;
  size_t G = (size_t)n1 * n2 * n3;
  double recip[3] = {1.0 / c->p.a, 1.0 / c->p.b, 1.0 / c->p.c};
  double fc = sqrt(KOFDIE / c->p.eps), energy = 0;
  double *fr = (double *)malloc(sizeof(double) * 3 * npt);
  double *th = (double *)malloc(sizeof(double) * 3 * 4 * npt);
  double *dth = (double *)malloc(sizeof(double) * 3 * 4 * npt);
  cplx *Q = (cplx *)calloc(G, sizeof(cplx));
  int nn[3] = {n1, n2, n3};
  if (!(c->p.terms & MOIL_TERM_EL)) fc = 0.0;
  /* get_scaled_fractionals + get_bspline_coeffs (ewald_receip.f:250-260, 973-1011) */
  for (n = 0; n < npt; n++)
    for (a = 0; a < 3; a++) {
      double w = coor[3 * n + a] * recip[a];
      double f = nn[a] * (w - round(w) + 0.5);
      fr[3 * n + a] = f;
      fill_bspline(f - (int)f, order, &th[(3 * n + a) * 4], &dth[(3 * n + a) * 4]);
    }
  /* fill_charge_grid (ewald_receip.f:637-660) */
  for (n = 0; n < npt; n++) {
    int i0 = (int)fr[3 * n] - order, j0 = (int)fr[3 * n + 1] - order, k0 = (int)fr[3 * n + 2] - order;
    double fc1 = c->ptchg[n] * fc;
    for (int i3 = 0; i3 < order; i3++) {
      int k = k0 + 1 + i3;
      double prod;
      if (k < 0) k += n3;
      prod = th[(3 * n + 2) * 4 + i3] * fc1;
      for (int i2 = 0; i2 < order; i2++) {
        int j = j0 + 1 + i2;
        double prod1;
        if (j < 0) j += n2;
        prod1 = th[(3 * n + 1) * 4 + i2] * prod;
        for (int i1 = 0; i1 < order; i1++) {
          int i = i0 + 1 + i1;
          if (i < 0) i += n1;
          Q[((size_t)k * n2 + j) * n1 + i].re += th[(3 * n) * 4 + i1] * prod1;
        }
      }
    }
  }
  fft3(Q, n1, n2, n3, -1);
  /* scalar_sum (ewald_receip.f:723-812) */
  {
    double pi = 3.14159265358979323846, fac = pi * pi / (c->beta * c->beta);
    double pivol = pi * (c->p.a * c->p.b * c->p.c);
    int nf[3];
    for (a = 0; a < 3; a++) { nf[a] = nn[a] / 2; if (2 * nf[a] < nn[a]) nf[a]++; }
    for (size_t ind = 1; ind < G; ind++) {
      int k3 = (int)(ind / ((size_t)n1 * n2)), rem = (int)(ind % ((size_t)n1 * n2));
      int k2 = rem / n1, k1 = rem % n1;
      int m1 = k1 + 1 > nf[0] ? k1 - n1 : k1;
      int m2 = k2 + 1 > nf[1] ? k2 - n2 : k2;
      int m3 = k3 + 1 > nf[2] ? k3 - n3 : k3;
      double mh1 = recip[0] * m1, mh2 = recip[1] * m2, mh3 = recip[2] * m3;
      double msq = mh1 * mh1 + mh2 * mh2 + mh3 * mh3;
      double denom = pivol * c->bsp[0][k1] * c->bsp[1][k2] * c->bsp[2][k3] * msq;
      double eterm = exp(-fac * msq) / denom;
      cplx *q = &Q[((size_t)k3 * n2 + k2) * n1 + k1];
      energy += eterm * (q->re * q->re + q->im * q->im);
      if (V) {
        /* ewald_receip.f:799-807 */
        const double vr = 0.5 * eterm * (q->re * q->re + q->im * q->im);
        const double mh[3] = {mh1, mh2, mh3};
        for (a = 0; a < 3; a++)
          V[7 + a] = V[7 + a] + vr * (1.0 - 2.0 * (1.0 + fac * msq) / msq * mh[a] * mh[a]);
      }
      q->re *= eterm;
      q->im *= eterm;
    }
    /* the m = 0 term is left untouched, as MOIL's loop starts at ind=1 */
  }
  E->e_ew_receip = 0.5 * energy;
  fft3(Q, n1, n2, n3, +1);
  /* grad_sum (ewald_receip.f:886-927) */
  for (n = 0; n < npt; n++) {
    double f1 = 0, f2 = 0, f3 = 0, term1 = fc * c->ptchg[n];
    int i0 = (int)fr[3 * n] - order, j0 = (int)fr[3 * n + 1] - order, k0 = (int)fr[3 * n + 2] - order;
    const double *t1 = &th[(3 * n) * 4], *t2 = &th[(3 * n + 1) * 4], *t3 = &th[(3 * n + 2) * 4];
    const double *d1 = &dth[(3 * n) * 4], *d2 = &dth[(3 * n + 1) * 4], *d3 = &dth[(3 * n + 2) * 4];
    for (int i3 = 0; i3 < order; i3++) {
      int k = k0 + 1 + i3;
      if (k < 0) k += n3;
      for (int i2 = 0; i2 < order; i2++) {
        int j = j0 + 1 + i2;
        if (j < 0) j += n2;
        for (int i1 = 0; i1 < order; i1++) {
          int i = i0 + 1 + i1;
          double term;
          if (i < 0) i += n1;
          term = term1 * Q[((size_t)k * n2 + j) * n1 + i].re;
          f1 -= term * d1[i1] * t2[i2] * t3[i3] * n1;
          f2 -= term * t1[i1] * d2[i2] * t3[i3] * n2;
          f3 -= term * t1[i1] * t2[i2] * d3[i3] * n3;
        }
      }
    }
    dpot[3 * n] -= recip[0] * f1;
    dpot[3 * n + 1] -= recip[1] * f2;
    dpot[3 * n + 2] -= recip[2] * f3;
    if (V) {
      /* ewald_receip.f:929-933: the reciprocal force on n times its
       * position relative to its molecule's centre */
      V[10] = V[10] + recip[0] * f1 * c->xrel[3 * n];
      V[11] = V[11] + recip[1] * f2 * c->xrel[3 * n + 1];
      V[12] = V[12] + recip[2] * f3 * c->xrel[3 * n + 2];
    }
  }
  (void)b; (void)g;
  free(fr); free(th); free(dth); free(Q);
}

/* Molecules for the virial (COMMON/PRESS.BLOCK as analysis/init_pressure.f
 * fills it: patom[m] the last ordlist position of molecule m, ordlist the
 * atoms molecule by molecule, pmol the molecule of each atom) and the
 * masses; the molecular masses summed as compute_massmol does
 * (analysis/init_pressure.f:411-428).  nmol = 0 switches the virial off. */
int orc_set_molecules(orc_ctx *c, int nmol, const int32_t *patom, const int32_t *ordlist,
                      const int32_t *pmol, const double *ptms) {
  int m, k = 1;
  free(c->patom); free(c->ordlist); free(c->pmol); free(c->ptms); free(c->massmol);
  free(c->xmol); free(c->xrel);
  c->patom = c->ordlist = c->pmol = NULL;
  c->ptms = c->massmol = c->xmol = c->xrel = NULL;
  c->nmol = 0;
  if (nmol <= 0) return 0;
  c->patom = dup_i(patom, nmol);
  c->ordlist = dup_i(ordlist, c->npt);
  c->pmol = dup_i(pmol, c->npt);
  c->ptms = dup_d(ptms, c->npt);
  c->massmol = (double *)calloc(nmol, sizeof(double));
  c->xmol = (double *)calloc(3 * (size_t)nmol, sizeof(double));
  c->xrel = (double *)calloc(3 * (size_t)c->npt, sizeof(double));
  for (m = 1; m <= nmol; m++) {
    double sm = 0.0;
    for (int a = k; a <= patom[m - 1]; a++) sm = sm + ptms[ordlist[a - 1] - 1];
    c->massmol[m - 1] = sm;
    k = patom[m - 1] + 1;
  }
  c->nmol = nmol;
  return 0;
}

/* centres of mass and relative positions, as the coordinate part of
 * atom2molecule (analysis/pressure.f:419-505): per molecule and
 * component the mass-weighted sum in ordlist order over the molecular
 * mass */
static void mol_coords(orc_ctx *c, const double *coor) {
  int m, k0 = 0, d, i;
  for (m = 1; m <= c->nmol; m++) {
    for (d = 0; d < 3; d++) {
      double sm = 0.0;
      for (int k = k0 + 1; k <= c->patom[m - 1]; k++) {
        const int a = c->ordlist[k - 1];
        sm = sm + coor[3 * (a - 1) + d] * c->ptms[a - 1];
      }
      c->xmol[3 * (m - 1) + d] = sm / c->massmol[m - 1];
    }
    k0 = c->patom[m - 1];
  }
  for (i = 1; i <= c->npt; i++)
    for (d = 0; d < 3; d++)
      c->xrel[3 * (i - 1) + d] = coor[3 * (i - 1) + d] - c->xmol[3 * (c->pmol[i - 1] - 1) + d];
}

/* the virial terms of the last orc_eforce with molecules set: virial,
 * virXX/YY/ZZ, V_PIdirXX/YY/ZZ, V_PIrecXX/YY/ZZ, V_PIrelXX/YY/ZZ */
int orc_virial(const orc_ctx *c, double *out) {
  memcpy(out, c->vir, sizeof(c->vir));
  return c->nmol > 0 ? 0 : MOIL_GPU_ESTATE;
}

/* eforce, pot/eforce.f:37-311 (energy path subset) */
int orc_eforce(orc_ctx *c, const double *coor, double *dpot, double *en /* moil_energies */) {
  eterms E;
  moil_energies *out = (moil_energies *)en;
  uint32_t t = c->p.terms;
  double *V = c->nmol > 0 ? c->vir : NULL;
  if (!c->have_lists) return MOIL_GPU_ESTATE;
  memset(&E, 0, sizeof(E));
  memset(dpot, 0, sizeof(double) * 3 * c->npt);
  memset(c->vir, 0, sizeof(c->vir));
  if (V) mol_coords(c, coor);
  if ((t & MOIL_TERM_BOND) && c->nb > 0) ebond(c, coor, dpot, &E);
  if ((t & MOIL_TERM_ANGLE) && c->nangl > 0) etheta(c, coor, dpot, &E);
  if ((t & MOIL_TERM_TORS) && c->ntors > 0) etors(c, coor, dpot, &E);
  if ((t & MOIL_TERM_IMP) && c->nimp > 0) eimphi(c, coor, dpot, &E);
  if (t & (MOIL_TERM_VDW | MOIL_TERM_EL)) {
    real_atoms(c, coor, dpot, &E, V);
    pme(c, coor, dpot, &E, V);
    real_waters(c, coor, dpot, &E, V);
    if (c->totspe > 0) ener14(c, coor, dpot, &E);
  }
  memset(out, 0, sizeof(*out));
  out->e_bond = E.e_bond; out->e_theta = E.e_theta; out->e_tors = E.e_tors; out->e_imp = E.e_imp;
  out->e_vdw = E.e_vdw; out->e_el = E.e_el; out->e_vsym = E.e_vsym; out->e_lsym = E.e_lsym;
  out->e_el14 = E.e_el14; out->e_vdw14 = E.e_vdw14; out->e_corr = E.e_corr;
  out->e_self = c->e_self; out->e_ew_receip = E.e_ew_receip;
  out->e_total = 0;
  if ((t & MOIL_TERM_BOND) && c->nb > 0) out->e_total += out->e_bond;
  if ((t & MOIL_TERM_ANGLE) && c->nangl > 0) out->e_total += out->e_theta;
  if ((t & MOIL_TERM_TORS) && c->ntors > 0) out->e_total += out->e_tors;
  if ((t & MOIL_TERM_IMP) && c->nimp > 0) out->e_total += out->e_imp;
  if (t & (MOIL_TERM_VDW | MOIL_TERM_EL)) {
    out->e_total += out->e_ew_receip + out->e_corr + out->e_self;
    out->e_total += out->e_vdw + out->e_el;
    out->e_total += out->e_el14 + out->e_vdw14;
    out->e_total += out->e_vsym + out->e_lsym;
  }
  out->e_dir = out->e_el + out->e_lsym;
  return 0;
}

/* Convenience for the CPU baseline: build lists every nlist calls. */
int orc_num_waters_real(orc_ctx *c) { return c->nwat; }
