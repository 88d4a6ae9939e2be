// Argument blocks of the force-evaluation kernels.
//
// Every kernel of one evaluation takes its arguments as one of these
// blocks.  A single-system launch passes the block by value; a
// multi-system launch (batch.hip) passes a device table holding one block
// per system plus the per-system workgroup offsets: each workgroup finds
// its system by binary search over the offsets and runs the same kernel
// body with that system's block and its workgroup number inside the
// system.  Pointers to grid data whose element type depends on the
// precision (float or double) are carried untyped.
#pragma once

#include <string>

#include "ctx.h"

// ---- real space, water groups (realspace.hip) ----
struct WaterRS {
  double qq_oo, qq_oh, qq_hh, oo12, oo6;
  double beta, cutvdw2, cutele2, erftbdns;
  double a, b, c;
  double fix_scale;
  int ntab;
  int *ovf;  // fixed-point overflow word of the context
};
struct PackArgs {
  const double *coor;
  const int32_t *slot_o;
  int nwat;
  void *wpos;  // float4[3] (mixed) or double[9] (FP64) per water slot
};
struct WaterArgs {
  const int32_t *chunks;  // [nchunk][3] block, begin, end
  const int32_t *order;   // [nchunk] dispatch order
  const WItem *items;
  const int32_t *jpos;
  const void *wpos;
  int nwat;
  WaterRS P;
  const double *erf;  // FP64 only
  long long *jpart, *ipart;  // jpart rows: float[9] (mixed) or int64[9] (FP64)
  double *epart;
  int max_parts;
};
struct WReduceArgs {
  const int32_t *jstart;
  const long long *jpart;  // float[9] rows (mixed) or fixed-point int64[9] (FP64)
  int jf32;
  const int32_t *cfirst;
  const long long *ipart;
  const int32_t *slot_o;
  int nwat;
  long long *facc;
  int *ovf;
};

// ---- real space, atom lists (atoms.hip) ----
struct ARS {
  double qfac, pick;  // kofdie/eps (or 0), 1 (or 0) for LJ
  double beta, cutvdw2, cutele2, erftbdns;
  double a, b, c;
  double fix_scale;
  int ntab;  // rows of the erfc table
  int arith;  // LJ by the arith rule (epsgm6 holds sigma)
  int *ovf;
};
struct AtomArgs {
  const int32_t *chunks;
  const AItem *items;
  const double *coor;
  const int32_t *sslot2atom;
  int nsol;
  const double *q, *e12, *e6, *erf;
  ARS P;
  const int32_t *jpos;
  long long *jpart, *ipart;
  double *epart;
  int max_parts, part0;
};
struct AReduceArgs {
  const int32_t *jstart;
  const long long *jpart;
  const int32_t *slot_of, *cfirst;
  const long long *ipart;
  int npt;
  long long *facc;
};

// ---- bonded terms (bonded.hip) ----
struct BPar {
  double fix_scale;
  int *ovf;
  double epstmp14, pick;  // ener14 switches
  double qfac;            // kofdie/eps (exclusion correction)
  double beta, erftbdns;
  int ntab;  // rows of the erfc table
  int do_el14, do_v14;
  int arith;  // arith torsion and improper forms
  int max_parts;
};
struct BSeg {
  int nb, nangl, ntors, nimp, nexcl, n14;
  const int32_t *ib, *ia, *it, *ii, *ei, *ej, *sp;
  const double *bp, *ap, *tp, *ip, *p14, *q, *erf;
};
struct WIntra {
  int nwat;
  const int32_t *first;               // [nwat] first atom (0-based)
  const int32_t *boff, *aoff, *eoff;  // [nwat+1] CSR offsets
  const int32_t *bidx, *aidx, *eidx;  // local atom numbers 0..2
  const double *bpar, *apar;          // k, r0 / k, theta0
  const double *q, *erf;
};
struct BondedArgs {
  const double *coor;
  BSeg S;
  BPar P;
  long long *bslot;  // per-term gradient slots, fixed point [nbslot][3]
  double *ep;
  int mp;
};
struct IntraArgs {
  const double *coor;
  WIntra W;
  BPar P;
  long long *facc;
  double *ep;
  int mp, part0;
};

// ---- reciprocal space (pme.hip) ----
struct PmePar {
  int n1, n2, n3, nc1;
  double ra, rb, rc;  // 1/a, 1/b, 1/c
  double fc;          // sqrt(kofdie/eps)
  double qscale, qinv;
};
// grid tiles: dimension k cut into nt_k tiles; LDS regions r1 r2 r3
// (the largest tile + 3 stencil points + the margin on both sides)
struct TilePar {
  int n1, n2, n3;
  int nt1, nt2, nt3;
  int r1, r2, r3;
  int rsz;
};
struct TbinArgs {
  const double *coor;
  int npt;
  const int32_t *order;
  PmePar P;
  TilePar T;
  int32_t *tcnt, *atile, *arank;
  int32_t *tstart;   // scan output when the last workgroup scans (fuse_scan)
  int fuse_scan;     // 1: the last workgroup scans the counts (tcnt[ntile] is its ticket)
};
struct TscanArgs {
  int32_t *cnt;
  int n;
  int32_t *start;
};
struct TplaceArgs {
  int npt;
  const int32_t *atile, *arank, *tstart;
  int32_t *sidx;
};
struct TspreadArgs {
  const double *coor, *q;
  const int32_t *tstart, *sidx;
  PmePar P;
  TilePar T;
  long long *tbuf;
  unsigned long long *qover;  // grid points of atoms outside their tile's region
};
struct TsumArgs {
  const long long *tbuf;
  unsigned long long *qover;  // read and cleared
  TilePar T;
  double qinv;
  void *grid;
};
struct ConvArgs {
  void *S;
  const double *eterm;
  int nc1, n1;
  size_t H;
  double *epart;
};
struct GatherArgs {
  const double *coor, *q;
  const int32_t *tstart, *sidx;
  PmePar P;
  TilePar T;
  const void *grid;
  double *fpme;
};

// ---- assembly (api.hip) ----
struct FinArgs {
  long long *facc;
  const double *fpme;
  int has_pme;
  size_t n;
  double inv;
  double *dpot;
  const int32_t *brstart, *brslot;  // per-atom lists of bonded gradient slots
  const long long *bslot;
};
struct EnCfg {
  int max_parts;
  int nparts[EN_COUNT];
  double e_self;
  int has_bond, has_angle, has_tors, has_imp, has_nb, has_14;
};
struct EnArgs {
  const double *ep;
  EnCfg C;
  double *tot;
  unsigned int *ticket;  // [0] last-workgroup ticket, [1] fixed-point overflow word
  double *out;
};
struct AsmArgs {
  FinArgs fin;
  int n_fin;
  EnArgs en;
};

// Every launch of one evaluation of one system: reciprocal space,
// intra-water + generic bonded terms, water real space, atom real space,
// force and energy assembly.  A count of 0 skips the launch.
struct EvalPlan {
  bool f64 = false;
  bool pme = false;
  TspreadArgs tspread;
  int n_tile = 0;
  TsumArgs tsum;
  int n_tsum = 0;
  ConvArgs conv;
  int n_conv = 0;
  GatherArgs gather;
  IntraArgs intra;
  int n_intra = 0;
  BondedArgs bonded;
  int n_bonded = 0;
  PackArgs pack;
  int n_pack = 0;
  WaterArgs water;
  int n_water = 0;
  WReduceArgs wred;
  int n_wred = 0;
  AtomArgs atoms;
  int n_atoms = 0;
  AReduceArgs ared;
  int n_ared = 0;
  FinArgs fin;
  int n_fin = 0;
  EnArgs en;
  AsmArgs asmb;  // fin + en in one launch of n_fin + EN_COUNT workgroups
  // the reciprocal virial terms (moil_gpu_eforce_virial_dev): V_PIrec per
  // k after the forward transform, V_PIrel per atom after the gather
  bool vir = false;
  const double *acom = nullptr;
  double *vrec = nullptr, *vrel = nullptr;
};
// the reciprocal virial kernels (virial.hip), launched by pme_run
void launch_pme_virial(const moil_ctx *c, const EvalPlan &p, hipStream_t st);
void launch_rel_virial(const moil_ctx *c, const EvalPlan &p, hipStream_t st);

// plan builders (one per module; the caller fills d_coor / outputs)
void pme_plan(const moil_ctx *c, const double *d_coor, EvalPlan *p);
void bonded_plan(const moil_ctx *c, const double *d_coor, EvalPlan *p);
void water_plan(const moil_ctx *c, const double *d_coor, EvalPlan *p);
void atoms_plan(const moil_ctx *c, const double *d_coor, int part0, EvalPlan *p);

// Launchers.  launch_X runs kernel X for one system (nblk workgroups);
// launch_X_multi runs it over nsys systems from a device table of
// argument blocks and nsys+1 workgroup offsets (total = offsets[nsys]).
#define DECLARE_LAUNCH(NAME, ARGS)                                                        \
  void launch_##NAME(const ARGS &a, int nblk, hipStream_t st);                            \
  void launch_##NAME##_multi(const ARGS *tab, const int32_t *bstart, int nsys, int total, \
                             hipStream_t st);
DECLARE_LAUNCH(k_pack_water_f32, PackArgs)
DECLARE_LAUNCH(k_pack_water_f64, PackArgs)
DECLARE_LAUNCH(k_water_f32, WaterArgs)
DECLARE_LAUNCH(k_water_f64, WaterArgs)
DECLARE_LAUNCH(k_water_reduce, WReduceArgs)
DECLARE_LAUNCH(k_atoms_f32, AtomArgs)
DECLARE_LAUNCH(k_atoms_f64, AtomArgs)
DECLARE_LAUNCH(k_atoms_reduce, AReduceArgs)
DECLARE_LAUNCH(k_water_intra, IntraArgs)
DECLARE_LAUNCH(k_bonded, BondedArgs)
DECLARE_LAUNCH(k_pme_tbin, TbinArgs)
DECLARE_LAUNCH(k_pme_tscan, TscanArgs)
DECLARE_LAUNCH(k_pme_tplace, TplaceArgs)
DECLARE_LAUNCH(k_pme_tspread_f32, TspreadArgs)
DECLARE_LAUNCH(k_pme_tspread_f64, TspreadArgs)
DECLARE_LAUNCH(k_pme_tsum_f32, TsumArgs)
DECLARE_LAUNCH(k_pme_tsum_f64, TsumArgs)
DECLARE_LAUNCH(k_conv_f32, ConvArgs)
DECLARE_LAUNCH(k_conv_f64, ConvArgs)
DECLARE_LAUNCH(k_gather_f32, GatherArgs)
DECLARE_LAUNCH(k_gather_f64, GatherArgs)
DECLARE_LAUNCH(k_finalize_forces, FinArgs)
DECLARE_LAUNCH(k_energies, EnArgs)
DECLARE_LAUNCH(k_assemble, AsmArgs)
#undef DECLARE_LAUNCH

// the real-space kernels with dispatch-recorded start/stop events (either
// may be null); the water kernel optionally on a short grid of nwg
// workgroups
void launch_k_water_ev(const WaterArgs &a, bool f64, int nblk, hipStream_t st, hipEvent_t e0,
                       hipEvent_t e1, int nwg = 0);
void launch_k_atoms_ev(const AtomArgs &a, bool f64, int nblk, hipStream_t st, hipEvent_t e0,
                       hipEvent_t e1);

// PME of one system: spreading, transforms on the context's buffers,
// convolution, gather; pme_bin assigns the atoms to spreading tiles (at
// every list build)
int32_t pme_bin(moil_ctx *c, const double *d_coor, hipStream_t st);
// spreading tiles of at most `edge` points a side (4..8), buffers resized
int32_t pme_set_tile(moil_ctx *c, int edge);
void pme_launch_spread(const EvalPlan &p, hipStream_t st);
int32_t pme_run(moil_ctx *c, const EvalPlan &p, hipStream_t st);
int pme_conv_blocks(const moil_ctx *c);
// whether every charge's stencil weight q sqrt(kofdie/eps) fits the grid's
// fixed-point conversion (|q| sqrt(kofdie/eps) < 2^11); else the message
bool pme_charges_fit(const double *q, size_t n, double eps, std::string *why);
int bonded_blocks(const moil_ctx *c);
int water_intra_blocks(const moil_ctx *c);
int bonded_max_blocks(const moil_ctx *c);
// the whole evaluation of one system (api.hip); d_energies may be null
// (the context's own record)
void eval_plan(moil_ctx *c, const double *d_coor, double *d_dpot, double *d_energies,
               EvalPlan *p);
// the plan's launches in order on st (api.hip)
int32_t eforce_eager(moil_ctx *c, const EvalPlan &p, hipStream_t st);
// the caller's stream argument as a HIP stream
hipStream_t pick_stream(moil_ctx *c, void *stream);
// looks up the no-allocation evaluation's scratch address on device dev
// (noalloc.hip; once per device, with dev current)
int32_t scratch_attach(int dev);
