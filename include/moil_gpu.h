/*
 * moil_gpu.h -- C ABI of the MI355X energy/force library for MOIL.
 *
 * The library replaces MOIL's energy-path subroutines, which in MOIL are
 * argument-less Fortran routines communicating through COMMON blocks:
 *
 *   ewald_init   symm/ewald_init.f:8-242     -> moil_gpu_init
 *   squeeze      symm/squeeze.f:1-147        -> moil_gpu_squeeze
 *   nbondm+nbond pot/nbondm.f:1-247,
 *                pot/nbond.f:1-262           -> moil_gpu_build_lists
 *   syminit+nbmsym+nbsym symm/syminit.f:1-178,
 *                symm/nbmsym.f:1-351,
 *                symm/nbsym.f:1-192          -> moil_gpu_build_lists
 *   eforce       pot/eforce.f:1-325          -> moil_gpu_eforce
 *
 * Every pointer argument is a plain host (or, for the *_dev variants,
 * device) pointer; arrays follow MOIL's COMMON layouts and 1-based index
 * conventions so a Fortran ISO_C_BINDING shim can pass COMMON arrays
 * straight through (coor/dpot are MOIL's coor(3,npt)/dpot(3,npt)).
 *
 * Error convention: every entry point returns 0 on success and a
 * MOIL_GPU_E* code otherwise; moil_gpu_last_error() returns the message.
 * MOIL itself prints and STOPs through alert() (intrprtr/alert.f:10-24);
 * the Fortran shim maps a non-zero status to alert(...,1).
 *
 * Threading: one host thread per context; calls on one context are not
 * re-entrant (MOIL's global state has the same property).
 */
#ifndef MOIL_GPU_H
#define MOIL_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOIL_GPU_ABI_VERSION 1

/* status codes */
#define MOIL_GPU_OK        0
#define MOIL_GPU_EINVAL    1  /* bad argument / unsupported input      */
#define MOIL_GPU_ECHARGE   2  /* |sum q| > 1e-7 (symm/ewald_init.f:66) */
#define MOIL_GPU_ECAPACITY 3  /* a list or grid exceeds its capacity   */
#define MOIL_GPU_EHIP      4  /* HIP / rocFFT runtime failure          */
#define MOIL_GPU_ESTATE    5  /* call order (e.g. eforce before lists) */

/* monomer flags, derived by the caller from MOIL's monomer names */
#define MOIL_MONO_WATER     1 /* 'TIP3'/'SPC*' water group (pot/nbondm.f:81-82) */
#define MOIL_MONO_SQZ_FIRST 2 /* squeezed by first atom (symm/squeeze.f:75-79)  */
#define MOIL_MONO_SQZ_CENT  4 /* squeezed by centroid (symm/squeeze.f:70-72)    */

/* precision modes */
#define MOIL_PREC_FP64  0  /* FP64 pair math: MOIL parity mode              */
#define MOIL_PREC_MIXED 1  /* FP32 pair math, 64-bit fixed-point forces,
                              FP64 energies / Ewald self+reciprocal sums    */

/* energy-term switches (MOIL /LENER/ flags, pot/init_ef.f:24-47) */
#define MOIL_TERM_BOND   (1u << 0)   /* ebyes  */
#define MOIL_TERM_ANGLE  (1u << 1)   /* ethyes */
#define MOIL_TERM_TORS   (1u << 2)   /* etoyes */
#define MOIL_TERM_IMP    (1u << 3)   /* eimyes */
#define MOIL_TERM_VDW    (1u << 4)   /* evdyes */
#define MOIL_TERM_EL     (1u << 5)   /* eelyes */
#define MOIL_TERM_EL14   (1u << 6)   /* e14el_yes */
#define MOIL_TERM_VDW14  (1u << 7)   /* e14v_yes  */
#define MOIL_TERM_HVDW0  (1u << 8)   /* hvdw0 hydrogen rule (pot/nbond.f:115) */
/* arith: MOIL's 'arit' connectivity keyword (connect/connect.f:124), the
 * combination and term forms of its AMBER-style parameter files.  The
 * topology's per-atom arrays then hold epsgm12 = sqrt(4 eps) and epsgm6 =
 * sigma (connect/fillpt.f:136-179), and
 *   pair LJ   A = epsgm12_i epsgm12_j, sigma_ij = (epsgm6_i + epsgm6_j)/2,
 *             E = A sigma_ij^12 / r^12 - A sigma_ij^6 / r^6
 *             (pot/cdie_ewald.f:377-383,432-473,515-548,
 *             symm/symcdie_ewald.f:127-140,275-314); the water O-O
 *             constants follow (pot/watwat_ewald.f:62-68)
 *   torsion   E = k2 + phase2 k2 cos(n phi), n = period 1..4 (ktors1,
 *             ktors3, phase1, phase3 unused; pot/ephi.f:113-138,238-259)
 *   improper  atoms taken as (iimp4, iimp2, iimp1, iimp3),
 *             E = k - k cos(2 phi) (pot/eimphi.f:98-115,251-262)
 *   lists     no hydrogen rule (pot/nbond.f:115-117, symm/nbsym.f:104-106)
 * The 1-4 pair coefficients p14 are inputs either way.  Torsion periods
 * outside 1..4 are refused at init (MOIL leaves such a term's energy
 * unset).  The parameter-gradient entry points and moil_gpu_set_params
 * assume the geometric rule and refuse an arith context. */
#define MOIL_TERM_ARITH  (1u << 9)
#define MOIL_TERM_DEFAULT 0x1FFu

/* Topology: the energy-path part of MOIL's /CONNINT/ /CONNDBL/ /CONNLOG/
 * (COMMON/CONNECT.BLOCK).  All indices are 1-based as in MOIL.          */
typedef struct moil_topology {
  int32_t npt;
  int32_t totmon;            /* monomers (monomer division unsupported)   */
  const int32_t *poipt;      /* [totmon+1], poipt[0]=0, last atom of mono */
  const int32_t *monoflag;   /* [totmon] MOIL_MONO_* bits                 */
  const double *ptchg;       /* [npt] charges (e)                         */
  const double *epsgm6;      /* [npt] 2*sqrt(eps)*sigma^3                 */
  const double *epsgm12;     /* [npt] 2*sqrt(eps)*sigma^6                 */
  const int32_t *flagchr;    /* [npt] nonzero = charged (Fortran LOGICAL) */
  const int32_t *exc1;       /* [npt+1] exclusion CSR ends, exc1[0]=0     */
  const int32_t *exc2;       /* [exc1[npt]] excluded partners (> atom)   */
  int32_t nb;                /* bonds  E = k (r-r0)^2                     */
  const int32_t *ib1, *ib2;
  const double *kbond, *req;
  int32_t nangl;             /* angles E = k (theta-theta0)^2, radians    */
  const int32_t *iangl1, *iangl2, *iangl3;
  const double *kangl, *angleq;
  int32_t ntors;             /* proper torsions (cosine polynomial)       */
  const int32_t *itor1, *itor2, *itor3, *itor4, *period;
  const double *ktors1, *ktors2, *ktors3, *phase1, *phase2, *phase3;
  int32_t nimp;              /* improper torsions                          */
  const int32_t *iimp1, *iimp2, *iimp3, *iimp4;
  const double *kimp, *impeq;
  int32_t totspe;            /* scaled 1-4 pairs                           */
  const int32_t *spec1, *spec2;
  const double *p14;         /* [3*totspe]: A*v14, B*v14, qq*eps*el14*kofdie */
  int32_t nwaters;
  const int32_t *idxtip3;    /* [nwaters] monomer numbers of the waters    */
} moil_topology;

/* Run parameters (MOIL /NB_D/, /SYMMD/, /EWALD2/, /EWALD3/, eps). */
typedef struct moil_params {
  double a, b, c;            /* orthorhombic box edges, box centred at 0  */
  double cutvdw2;            /* rvmx^2  */
  double cutele2;            /* relx^2  */
  double cutvbig2;           /* rvbg^2  */
  double cutebig2;           /* rebg^2  */
  double cutmono2;           /* cutm^2  */
  double eps;                /* dielectric constant                       */
  double dtol;               /* Ewald direct-sum tolerance -> beta        */
  int32_t nfft1, nfft2, nfft3;
  int32_t order;             /* B-spline order (4 supported)              */
  int32_t precision;         /* MOIL_PREC_*                               */
  uint32_t terms;            /* MOIL_TERM_* bits                          */
} moil_params;

/* Energies, same names as MOIL's /ENERGY/ and /EWALD3/ scalars. */
typedef struct moil_energies {
  double e_total, e_bond, e_theta, e_tors, e_imp;
  double e_vdw, e_el, e_vsym, e_lsym, e_el14, e_vdw14;
  double e_corr, e_self, e_ew_receip, e_dir;
} moil_energies;
#define MOIL_N_ENERGIES 15

typedef struct moil_ctx moil_ctx;

/* Library / device info. */
int32_t moil_gpu_abi_version(void);
const char *moil_gpu_last_error(void);

/* ewald_init (symm/ewald_init.f:8-242): copies the topology to the GPU,
 * derives beta, the erfc spline table, e_self and the B-spline moduli,
 * and plans the FFTs.  Returns MOIL_GPU_ECHARGE for a non-neutral system. */
int32_t moil_gpu_init(const moil_topology *topo, const moil_params *par,
                      moil_ctx **out);
int32_t moil_gpu_finalize(moil_ctx *ctx);

/* beta (ewaldcof), e_self, table rows (mxerftab) and density (erftbdns). */
int32_t moil_gpu_ewald_info(moil_ctx *ctx, double *ewaldcof, double *e_self,
                            int32_t *ntab, double *erftbdns);
/* erfc spline table as MOIL's erf_arr(4,ntab) (symm/ewald_init.f:418-449),
 * read back from the device copy the FP64 kernels use */
int32_t moil_gpu_erf_table(moil_ctx *ctx, double *erf_arr, int32_t cap_rows);

/* ewald_init's host arithmetic without a context (symm/ewald_init.f:53-126,
 * 298-335, 418-574): beta from relx^2 and dtol by MOIL's bisection over
 * Cody's erfc (symm/erfcfun.f), the table size mxerftab and density
 * erftbdns, and (if erf_arr is not NULL) the first cap_rows rows of the
 * erfc spline table erf_arr(4,mxerftab), all bit-identical to MOIL's.
 * moil_gpu_init derives the same numbers.  Host only. */
int32_t moil_ewald_setup(double cutele2, double dtol, double *ewaldcof, int32_t *ntab,
                         double *erftbdns, double *erf_arr, int32_t cap_rows);

/* squeeze (symm/squeeze.f:1-147): wraps molecules back into the box;
 * coor(3,npt) is updated in place. */
int32_t moil_gpu_squeeze(moil_ctx *ctx, double *coor);

/* nbondm + nbond + syminit + nbmsym + nbsym: builds the buffered pair
 * lists (real and 13 half-shell image translations) from coor(3,npt). */
int32_t moil_gpu_build_lists(moil_ctx *ctx, const double *coor);

/* eforce (pot/eforce.f:1-325): energies and the gradient dpot(3,npt). */
int32_t moil_gpu_eforce(moil_ctx *ctx, const double *coor, double *dpot,
                        moil_energies *e);

/* Device-resident variants: d_coor/d_dpot are device pointers to
 * (3,npt) float64 arrays, d_energies a device pointer to
 * MOIL_N_ENERGIES float64 in moil_energies order; `stream` is the
 * caller's hipStream_t (NULL = the default null stream); the work is
 * ordered after the caller's earlier work on that stream.  Asynchronous
 * (build_lists_dev synchronizes the stream internally). */
int32_t moil_gpu_squeeze_dev(moil_ctx *ctx, double *d_coor, void *stream);
int32_t moil_gpu_build_lists_dev(moil_ctx *ctx, const double *d_coor, void *stream);
int32_t moil_gpu_eforce_dev(moil_ctx *ctx, const double *d_coor, double *d_dpot,
                            double *d_energies, void *stream);

/* Launch shapes of a context (no MOIL counterpart; the gradient is
 * bit-identical for every shape, the energies agree to FP64 summation
 * order, since a chunk budget regroups their sums): the pair budget of one
 * water real-space workgroup and of one atom-list workgroup, the
 * mixed-precision water kernel's workgroups per CU (0: one per chunk) and
 * the largest edge of a PME spreading tile (grid points, 4..8).
 * moil_gpu_autotune times evaluations of d_coor (device pointer, the
 * context's stream; reps timed evaluations per candidate, <= 0 for 20)
 * over a fixed candidate set, one shape at a time, keeps the fastest
 * shapes and rebuilds the lists with them; us_default / us_tuned are the
 * mean evaluation times before and after.  Results are cached per device
 * and system signature for the process (and in the file named by
 * MOIL_GPU_TUNE_CACHE, if set); a cached system is not timed again.
 * moil_gpu_set_tuning applies shapes directly (chunk budgets and tiles
 * take effect at the next moil_gpu_build_lists*, which must precede the
 * next evaluation when the tile shape changed). */
typedef struct moil_tuning {
  int32_t wchunk_pairs;     /* 256..65536, default 2048 */
  int32_t achunk_pairs;     /* 256..65536, default 8192 */
  int32_t water_wg_per_cu;  /* 0..16, default 4 */
  int32_t pme_tile;         /* 4..8, default 8 */
  double us_default, us_tuned;
} moil_tuning;
int32_t moil_gpu_get_tuning(moil_ctx *ctx, moil_tuning *t);
int32_t moil_gpu_set_tuning(moil_ctx *ctx, const moil_tuning *t);
int32_t moil_gpu_autotune(moil_ctx *ctx, const double *d_coor, int32_t reps, moil_tuning *out);

/* No counterpart in MOIL: moil_gpu_eforce_dev for callers that cannot
 * allocate during the call (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The same evaluation with the same
 * results, ordered on `stream` as moil_gpu_eforce_dev's (the reciprocal
 * sum forks to the context's second stream and joins back through the
 * context's own events), never replayed as a graph, with its reduction
 * scratch (energy partial sums, per-term totals, the reduction's ticket
 * and the fixed-point overflow word) in a static 8 MiB device array of
 * the library, one per GPU, held until the process ends: no allocation
 * or free during the call.  The lists must have been built
 * (moil_gpu_build_lists*).  MOIL_GPU_ECAPACITY when the evaluation needs
 * more than the array's 87,376 partial sums per energy term (real-space
 * chunks, reciprocal convolution blocks or bonded blocks).  Every call on
 * one GPU shares that GPU's array: calls that overlap in time on one GPU
 * race. */
int32_t moil_gpu_eforce_noalloc_dev(moil_ctx *ctx, const double *d_coor, double *d_dpot,
                                    double *d_energies, void *stream);

/* Batched evaluation (no counterpart in MOIL, which evaluates one system
 * per process; the chain/path drivers loop over slices, chain/sds.f:71-140).
 * A batch groups existing contexts of one precision (independent systems
 * or path slices of any sizes); moil_gpu_batch_eforce_dev evaluates all of
 * them, d_coor[i], d_dpot[i], d_energies[i] being system i's device arrays
 * as for moil_gpu_eforce_dev.  Every kernel runs once over all systems
 * (per-system argument tables and workgroup offsets), the transforms as
 * one batched plan per grid size; the step is captured once into two
 * hipGraphs (real space; reciprocal sums with the transform groups on up
 * to four streams) replayed side by side, and recaptured when a system's
 * lists were rebuilt or the pointers change (moil_gpu_batch_set_graph
 * switches to direct launches).  `nstreams` is accepted for
 * compatibility and unused.  Ordered after the caller's earlier work on
 * `stream` (NULL = null stream) and before its later work.  Results are
 * identical to one moil_gpu_eforce_dev per system. */
typedef struct moil_batch moil_batch;
int32_t moil_gpu_batch_create(moil_ctx **ctxs, int32_t n, int32_t nstreams, moil_batch **out);
int32_t moil_gpu_batch_eforce_dev(moil_batch *batch, const double *const *d_coor,
                                  double *const *d_dpot, double *const *d_energies,
                                  void *stream);
int32_t moil_gpu_batch_destroy(moil_batch *batch);
/* Lists of every system of the batch (as moil_gpu_build_lists_dev per
 * system, preceded by moil_gpu_squeeze_dev when `squeeze` is nonzero;
 * d_coor[i] is updated by the squeeze): the builds are queued side by side
 * on the systems' own streams and the call waits on the host once for all
 * of them (again for a system whose lists outgrew their buffers).
 * Ordered after the caller's earlier work on `stream`. */
int32_t moil_gpu_batch_build_lists_dev(moil_batch *batch, double *const *d_coor,
                                       int32_t squeeze, void *stream);
/* Graph replay of the batch step on (the default) or off.  Off, every
 * evaluation launches the same kernels directly on the batch's two
 * streams: for callers that rebuild lists before every evaluation, where
 * each call would recapture.  Results are identical either way. */
int32_t moil_gpu_batch_set_graph(moil_batch *batch, int32_t on);

/* MOIL connectivity (.con) files: moil_con_read parses one the way MOIL's
 * rconn does (generic/rconn.f:46-497, version 12.1, fixed-column records)
 * into a topology for moil_gpu_init; the arrays belong to the returned
 * handle and stay valid until moil_con_free.  Monomer flags come from the
 * monomer names as MOIL's list and squeeze code tests them.  Files with LES
 * copies, Morse bonds, the beta list or monomer division are refused
 * (MOIL_GPU_EINVAL).  Host only: no GPU is touched. */
typedef struct moil_con moil_con;
int32_t moil_con_read(const char *path, moil_con **out, moil_topology *topo);
/* the file's arith flag (rconn.f:60-61): *arith = 1 when the context needs
 * MOIL_TERM_ARITH in its parameters' term bits */
int32_t moil_con_arith(const moil_con *con, int32_t *arith);
int32_t moil_con_free(moil_con *con);
/* CHARMM-style coordinate file as MOIL's getcrd reads it
 * (generic/getcrd.f:48-97, format(i7,i7,1x,a4,1x,a4,3(f10.5),...)) into
 * coor(3,npt).  Only complete files: a particle record out of sequence is
 * refused (getcrd would fill placeholders and build hydrogens). */
int32_t moil_crd_read(const char *path, int32_t npt, double *coor);
/* Binary trajectory as MOIL's wdyncrd writes it (dynamics/wdyncrd.f:24-88,
 * binary, no frozen particles): moil_traj_open writes the header for nfrm
 * frames of npt particles, each moil_traj_write one frame from coor(3,npt)
 * (host, stored in single precision).  Byte-identical to MOIL's file. */
typedef struct moil_traj moil_traj;
int32_t moil_traj_open(const char *path, int32_t npt, int32_t nfrm, moil_traj **out);
int32_t moil_traj_write(moil_traj *traj, const double *coor);
int32_t moil_traj_close(moil_traj *traj);
/* Frame `frame` (0-based) of such a trajectory into coor(3,npt), as MOIL's
 * rdyncrd reads it (dynamics/rdyncrd.f:40-70); *nframes (if not NULL)
 * gets the header's frame count. */
int32_t moil_traj_read(const char *path, int32_t npt, int32_t frame, double *coor,
                       int32_t *nframes);

/* force_norm (pot/force_norm.f:1-44): the RMS gradient component
 * sqrt(sum |dpot_i|^2 / (3 npt)) of dpot(3,npt), summed in MOIL's order.
 * Host only. */
int32_t moil_force_norm(const double *dpot, int32_t npt, double *gradf);

/* Chain of replicas: MOIL's sds (chain/sds.f:1-257), the energy and
 * gradient of a chain of igrid replicas of one system that the chain
 * minimiser, annealer and chain dynamics evaluate (chain/chnneal.f:85,
 * chain/chndyn.f:112, chain/pwl_chn.f:200):
 *   S = sum_{j=2}^{igrid-1} e_total(r_j) + gamma sum_j (d_j - <d>)^2
 *       + rho/lambda sum_j exp(-lambda D_j / <d>^2),
 * d_j = |r_{j+1}-r_j|, D_j = |r_{j+2}-r_j|^2 over atoms with ipick > 0.
 * moil_gpu_chain_create takes igrid-2 contexts, one per interior replica
 * (same topology and parameters, same precision; igrid >= 3).  A call
 * squeezes a copy of each interior replica and rebuilds its lists, as
 * sds does on every call (chain/sds.f:84-86), evaluates the interior
 * replicas as one batch, and adds the chain terms.  Arrays as sds's
 * arguments: r and dv (3, igrid*npt) replica-major; ipick (npt);
 * d0 and e1 (2*igrid-3); e0 (igrid), of which entries 2..igrid-1
 * (1-based) are written.  dv must not alias r.  The _dev variant takes
 * device pointers (sener a device scalar) and is ordered on `stream`;
 * the replicas' list builds run side by side on their contexts' own
 * streams, and the call waits on the host once for all of them (again
 * for a replica whose lists outgrew their buffers). */
typedef struct moil_chain moil_chain;
int32_t moil_gpu_chain_create(moil_ctx **interior, int32_t igrid, moil_chain **out);
int32_t moil_gpu_chain_sds(moil_chain *chain, const double *r, const int32_t *ipick,
                           double gamma, double rho, double lambda, double *sener,
                           double *dv, double *d0, double *e0, double *e1);
int32_t moil_gpu_chain_sds_dev(moil_chain *chain, const double *d_r, const int32_t *d_ipick,
                               double gamma, double rho, double lambda, double *d_dv,
                               double *d_d0, double *d_e0, double *d_e1, double *d_sener,
                               void *stream);
int32_t moil_gpu_chain_destroy(moil_chain *chain);

/* Resident molecular dynamics: MOIL's dynamics step loop
 * (dynamics/dyna.f:105-301) on coordinates and velocities that stay on the
 * GPU, with the matrix SHAKE of rigid waters (dynamics/mshakpt.f,
 * dynamics/mshakvl.f).  Every nlist steps: squeeze, list build and one
 * force evaluation on the new lists; then per step velo <- velo dt -
 * factor1 dpot, matrix SHAKE of that step, coor <- coor + velo,
 * velo <- velo / dt, eforce, velo <- velo - factor2 dpot, matrix SHAKE of
 * the velocities, with factor1 = dt^2/2 / m and factor2 = dt/2 / m
 * (dynamics/init_dyna.f:279-288).  Velocities in A per MOIL time unit,
 * dt in MOIL's time unit (dt[ps] / tconv, tconv = 0.0488882 ps,
 * COMMON/CONVERT.BLOCK).  The matrix SHAKE constants are MOIL's
 * MSHAKE.BLOCK after mshakinit (dynamics/mshakinit.f: the inverted 3x3
 * constraint matrix, column-major as in Fortran, tolcons, the squared O-H
 * and H-H distances and the inverse H and O masses).  No temperature
 * control, frozen particles or bond SHAKE (the Fortran shim refuses those
 * modes for the energy path).  A water whose matrix SHAKE does not
 * converge where MOIL would rebuild it or stop makes the run return
 * MOIL_GPU_ECAPACITY (checked at each list update and at the end). */
typedef struct moil_dyna_params {
  double dt;                 /* time step, MOIL time units                  */
  int32_t nlist;             /* steps between list updates                  */
  int32_t mshake;            /* nonzero: matrix SHAKE of every water (shakm) */
  double cmat[9];            /* MSHAKE.BLOCK cmat(3,3) (inverted), column-major */
  double tolcons, reqoh2, reqhh2, invmh, invmo;
} moil_dyna_params;
typedef struct moil_dyna moil_dyna;
/* ptms: [npt] masses (MOIL's ptms).  The dynamics object uses the
 * context's lists and evaluation. */
int32_t moil_gpu_dyna_create(moil_ctx *ctx, const double *ptms, const moil_dyna_params *par,
                             moil_dyna **out);
/* nstep steps from d_coor, d_velo (device (3,npt) float64, updated in
 * place); d_dpot gets the last gradient; d_energies (device, may be NULL)
 * the MOIL_N_ENERGIES record of every step's evaluation, step-major.
 * Ordered on `stream` (the list updates synchronize it). */
int32_t moil_gpu_dyna_run_dev(moil_dyna *dyna, double *d_coor, double *d_velo, double *d_dpot,
                              int32_t nstep, double *d_energies, void *stream);
/* The same from host arrays (energies: [nstep][MOIL_N_ENERGIES] or NULL). */
int32_t moil_gpu_dyna_run(moil_dyna *dyna, double *coor, double *velo, double *dpot,
                          int32_t nstep, double *energies);
int32_t moil_gpu_dyna_destroy(moil_dyna *dyna);

/* Force-field parameters and the energy's gradient with respect to them
 * (no counterpart in MOIL; for force-field fitting and the differentiable
 * binding).  moil_gpu_set_params replaces the per-atom charges and LJ
 * factors (MOIL's ptchg, epsgm6, epsgm12 of CONNECT.BLOCK) and rebuilds
 * what MOIL derives from them: e_self (symm/ewald_init.f:131-150), the
 * last-water parameters (pot/watwat_ewald.f:40-72) and the 1-4 pair
 * parameters p14 from the scalings v14 and el14 (generic/wconn.f:296-325).
 * The list classification (flagchr, the hydrogen rule) keeps the values of
 * moil_gpu_init; no neutrality check.  Refused (MOIL_GPU_EINVAL): water
 * charges or water O LJ factors that are not the same in every water
 * (MOIL's water-water pairs use the last water's, so per-water values
 * would make the energy's parameter gradient differ from what is
 * returned).
 * moil_gpu_param_grad[_dev] returns dE/dq, dE/d(epsgm12) and dE/d(epsgm6)
 * per atom, (npt) float64 each, of the energy moil_gpu_eforce evaluates on
 * the current lists (dE/dx is eforce's dpot).  Every pair distance is
 * recomputed; the reciprocal grid is rebuilt.  Water pairs follow MOIL's
 * use of the last water's parameters, each water atom receiving its own
 * share.  Requires moil_gpu_set_params when the system has 1-4 pairs. */
int32_t moil_gpu_set_params(moil_ctx *ctx, const double *ptchg, const double *epsgm6,
                            const double *epsgm12, double v14, double el14);
int32_t moil_gpu_param_grad(moil_ctx *ctx, const double *coor, double *dq, double *de12,
                            double *de6);
int32_t moil_gpu_param_grad_dev(moil_ctx *ctx, const double *d_coor, double *d_dq,
                                double *d_de12, double *d_de6, void *stream);

/* The molecular virial of MOIL's constant-pressure path (pressON,
 * COMMON/PRESS.BLOCK): the terms eforce adds up when pressON is set
 * (pot/eforce.f:70-85 zeroes them), in MOIL's names and order.  Real
 * space, for every listed pair with r = x_i - x_j and molecular
 * separation R = M_i - M_j of the two atoms' molecules' centres of mass
 * (each translated with its atom for an image pair):
 *   virial, virXX..ZZ  from the LJ part df of the pair's force factor:
 *                      virXX -= df rx Rx, virial -= df (r . R)
 *                      (pot/cdie_ewald.f:440-460, pot/watwat_ewald.f:162-174,
 *                      symm/symcdie_ewald.f:150-166, symm/symwat_ewald.f:160-173);
 *   V_PIdirXX..ZZ      V_PIdir rx Rx, V_PIdir = qq (erfc(beta r)/r^3 +
 *                      2 beta/sqrt(pi) exp(-beta^2 r^2)/r^2) for every
 *                      Coulomb pair and water site pair (pot/cdie_ewald.f:351-368);
 *   V_PIrecXX..ZZ      the reciprocal sum's sum over k of E(k) (1 - 2 (1 +
 *                      pi^2 m^2/beta^2) m_x^2/m^2) (pot/ewald_receip.f:799-807);
 *   V_PIrelXX..ZZ      the reciprocal force on every atom times its
 *                      position relative to its molecule's centre
 *                      (pot/ewald_receip.f:929-933).
 * The exclusion correction, 1-4 pairs and bonded terms add nothing
 * (MOIL's correction term is disabled, pot/cdie_ewald.f:162-171).
 * The real-space terms are evaluated in FP64 with MOIL's erfc table in
 * either precision mode; the reciprocal ones in the context's precision. */
#define MOIL_N_VIRIAL 13
typedef struct moil_virial {
  double virial, virXX, virYY, virZZ;
  double V_PIdirXX, V_PIdirYY, V_PIdirZZ;
  double V_PIrecXX, V_PIrecYY, V_PIrecZZ;
  double V_PIrelXX, V_PIrelYY, V_PIrelZZ;
} moil_virial;

/* Molecules for the virial (PRESS.BLOCK as MOIL's init_pressure fills it,
 * analysis/init_pressure.f:1-77): nmol molecules; patom[m] (1-based) the
 * position in ordlist of molecule m+1's last atom; ordlist the atoms
 * molecule by molecule; pmol[i] the molecule of atom i+1; ptms the atom
 * masses (CONNECT.BLOCK).  The molecular masses are summed in ordlist order
 * (compute_massmol, analysis/init_pressure.f:411-428).  nmol = 0 removes
 * the molecules.  Refused (MOIL_GPU_EINVAL): an atom in no molecule or in
 * one out of range, ordlist and pmol disagreeing, a molecule of mass <= 0. */
int32_t moil_gpu_set_molecules(moil_ctx *ctx, int32_t nmol, const int32_t *patom,
                               const int32_t *ordlist, const int32_t *pmol, const double *ptms);
/* MOIL's molecule division (create_mol_BFS, analysis/init_pressure.f:79-187)
 * for callers without MOIL: patom and pmol hold npt entries, ordlist npt.
 * Host only. */
int32_t moil_molecules(const moil_topology *topo, int32_t *nmol, int32_t *patom,
                       int32_t *ordlist, int32_t *pmol);
/* eforce with pressON (pot/eforce.f:1-325 with the virial terms above):
 * energies, the gradient and the virial.  The molecules' centres of mass
 * are computed from coor first (the coordinate part of atom2molecule,
 * analysis/pressure.f:419-505).  Requires moil_gpu_set_molecules.  The
 * _dev form writes MOIL_N_VIRIAL float64 (moil_virial order) to d_virial. */
int32_t moil_gpu_eforce_virial(moil_ctx *ctx, const double *coor, double *dpot,
                               moil_energies *e, moil_virial *v);
int32_t moil_gpu_eforce_virial_dev(moil_ctx *ctx, const double *d_coor, double *d_dpot,
                                   double *d_energies, double *d_virial, void *stream);

/* List inspection for parity tests.  counts[0..7]: water pairs real
 * listwt1, real listwt2, image listwt1, image listwt2; atom pairs list1,
 * list2, list3 (real+image), total image atom pairs. */
int32_t moil_gpu_list_counts(moil_ctx *ctx, int64_t counts[8]);
/* Canonical water pairs, rows of 4 int32: i monomer (the translated one
 * for images), j monomer, symmetry operation (0 real, 1..13 as
 * symm/syminit.f:80-164), list (1 = listwt1, 2 = listwt2). */
int32_t moil_gpu_dump_water_pairs(moil_ctx *ctx, int32_t *rows, int64_t cap,
                                  int64_t *n);
/* Canonical atom pairs, rows of 4 int32: l (translated for images), m,
 * symmetry operation, list (1, 2, 3 as pot/nbond.f:136-149). */
int32_t moil_gpu_dump_atom_pairs(moil_ctx *ctx, int32_t *rows, int64_t cap,
                                 int64_t *n);

/* The number of list builds of the context so far (moil_gpu_build_lists,
 * its _dev form and batched builds of the context each add one; no MOIL
 * counterpart): a caller that keeps results of an evaluation checks with
 * it that the lists they were computed on are still the context's. */
int32_t moil_gpu_list_generation(moil_ctx *ctx, int64_t *gen);

/* Stage timing with HIP events on the launch stream.  Evaluations run in
 * order on the launch stream.  With MOIL_GPU_GRAPH=1 in the environment
 * (experimental, off by default: a process running ~40 recaptures in a
 * parameter finite-difference loop aborted with the cause not identified)
 * they are replays of a captured hipGraph of the whole evaluation,
 * recaptured when the lists, the parameters or the pointers change.
 * moil_gpu_set_timing (which also clears) selects, in bits 0..7, level 0
 * (off), 1 (the real-space kernels are timed by events their own dispatch
 * records; under graph replay only the first evaluation after each such
 * change runs in order and is timed) or 2 (every evaluation in order, with
 * event records between its stages: PME, bonded, the whole eforce); bits
 * 8..15 = n >= 2 time only every n-th evaluation at level 1 (the dispatch
 * events cost ~2 % of the evaluation rate).  moil_gpu_stage_times returns
 * the summed milliseconds since the last call: t[0] real-space water
 * kernel, t[1] real-space atom kernel, t[2] PME, t[3] bonded + exclusion +
 * 1-4, t[4] whole eforce, t[5] the number of eforce calls timed. */
int32_t moil_gpu_set_timing(moil_ctx *ctx, int32_t level);
int32_t moil_gpu_stage_times(moil_ctx *ctx, double t[8]);

#ifdef __cplusplus
}
#endif

#endif /* MOIL_GPU_H */
