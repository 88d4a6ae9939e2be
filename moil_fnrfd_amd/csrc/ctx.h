// Library context: device copies of the topology, Ewald tables, pair
// lists and work buffers for one MOIL system.
#pragma once

#include <rocfft/rocfft.h>

#include <vector>

#include "common.h"

// a captured evaluation (api.hip): graph exec and what it was captured for
struct EGraph {
  hipGraphExec_t exec = nullptr;
  long gen = -1, pgen = -1;
  const void *key[3] = {nullptr, nullptr, nullptr};
};

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
};

struct moil_ctx {
  moil_params par;
  int npt = 0, totmon = 0, nwat = 0, nb = 0, nangl = 0, ntors = 0, nimp = 0, totspe = 0;
  int nexcl = 0;
  // generic (not intra-water) bonds, angles and exclusions; the
  // intra-water ones are listed per water in the d_wi_* arrays
  int gnb = 0, gnangl = 0, gnexcl = 0, wi_nterms = 0;
  // bonded gradient slots ([nbslot][3] fixed point, written whole by the
  // bonded kernel) and each atom's list of its slots (CSR, fixed topology)
  size_t nbslot = 0;
  long long *d_bslot = nullptr;
  int32_t *d_brstart = nullptr, *d_brslot = nullptr;
  int device = 0;                 // the HIP device current at moil_gpu_init
  int ncu = 0;                    // its compute units
  hipStream_t stream = nullptr;   // default stream of the context
  hipStream_t stream2 = nullptr;  // reciprocal space, overlapped with real space
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_intra = nullptr;  // the intra-water terms done on the second stream
  bool pme_concurrent = true;     // off with MOIL_GPU_PME_CONCURRENT=0 in the environment
  bool facc_dirty = false;
  int fix_shift = FIX_SHIFT_FP64;
  double fix_scale = 0, fix_inv = 0;

  // ---- topology (device) ----
  int32_t *d_poipt = nullptr, *d_monoflag = nullptr, *d_flagchr = nullptr;
  double *d_ptchg = nullptr, *d_eps6 = nullptr, *d_eps12 = nullptr;
  int32_t *d_exc_i = nullptr, *d_exc_j = nullptr;  // exclusion pairs (0-based)
  int32_t *d_ib = nullptr;     // [2*nb]
  double *d_bpar = nullptr;    // [2*nb] k, r0
  int32_t *d_ia = nullptr;     // [3*nangl]
  double *d_apar = nullptr;    // [2*nangl]
  int32_t *d_it = nullptr;     // [4*ntors]
  double *d_tpar = nullptr;    // [6*ntors] k1 k2 k3 ph1 ph2 ph3
  int32_t *d_ii = nullptr;     // [4*nimp]
  double *d_ipar = nullptr;    // [2*nimp]
  int32_t *d_sp = nullptr;     // [2*totspe]
  double *d_p14 = nullptr;     // [3*totspe]
  int32_t *d_wi_boff = nullptr, *d_wi_aoff = nullptr, *d_wi_eoff = nullptr;  // [nwat+1]
  int32_t *d_wi_zero = nullptr;                                             // [nwat+1] zeros
  int32_t *d_wi_bidx = nullptr, *d_wi_aidx = nullptr, *d_wi_eidx = nullptr;  // local 0..2
  double *d_wi_bpar = nullptr, *d_wi_apar = nullptr;
  int32_t *d_wat_o = nullptr;     // [nwat] O atom (0-based)
  int32_t *d_wat_mono = nullptr;  // [nwat] monomer number (1-based)
  std::vector<int32_t> h_wat_mono, h_wat_o;
  double qq_oo = 0, qq_oh = 0, qq_hh = 0, oo12 = 0, oo6 = 0;  // last-water parameters

  // ---- Ewald ----
  double beta = 0, e_self = 0, erftbdns = 5000.0;
  int ntab = 0;
  std::vector<double> h_erf;   // [ntab][4]
  double *d_erf = nullptr;
  double *d_eterm = nullptr;   // influence function on the half spectrum
  std::vector<double> bsp[3];
  int nf1 = 0, nf2 = 0, nf3 = 0, nc1 = 0;  // grid dims; nc1 = nf1/2+1
  int pt_nt[3] = {0, 0, 0}, pt_r[3] = {0, 0, 0}, pt_ntile = 0;  // spreading tiles
  int pt_edge = 0;                                   // their largest edge (points)
  int32_t *d_tcnt = nullptr, *d_tstart = nullptr;  // [ntile+1] atoms per tile, scan
  long long *d_tbuf = nullptr;                      // [ntile][r1 r2 r3] tile regions
  int32_t *d_atile = nullptr, *d_arank = nullptr;   // [npt] tile, rank in tile
  int32_t *d_sidx = nullptr;    // [npt] atom index, tile order (from the last list build)
  unsigned long long *d_qover = nullptr;  // [G] fixed-point charges of atoms outside their region
  void *d_qreal = nullptr;      // real grid (float or double)
  void *d_qcplx = nullptr;      // half spectrum
  rocfft_plan plan_fwd = nullptr, plan_bwd = nullptr;
  rocfft_execution_info exec_fwd = nullptr, exec_bwd = nullptr;
  void *d_fftwork = nullptr;
  size_t fftwork_bytes = 0;
  double *d_fpme = nullptr;     // [3*npt] reciprocal-space gradient

  // ---- monomers / water lists ----
  double *d_mono_cent = nullptr;  // [3*totmon]
  int32_t *d_flags = nullptr;     // [4]: large_cutmono flag, overflow flags
  int gx = 1, gy = 1, gz = 1, ncell = 1;  // water cell grid (brick-major order)
  double cs[3] = {1, 1, 1}, gmin[3] = {0, 0, 0};
  int32_t *d_cell_of = nullptr;    // [nwat] cell key
  int32_t *d_cell_cnt = nullptr;   // [ncell+1]
  int32_t *d_cell_start = nullptr; // [ncell+1]
  int32_t *d_cell_fill = nullptr;  // [ncell]
  int32_t *d_slot2wat = nullptr;   // [nwat]
  double *d_scent = nullptr;       // [3*nwat] slot centres (FP64)
  int32_t *d_smono = nullptr;      // [nwat] slot monomer numbers
  int nwblk = 0;                   // water blocks of WBLOCK slots
  int32_t *d_bcount = nullptr;     // [nwblk*NBUCKET+1]
  int32_t *d_boff = nullptr;       // [nwblk*NBUCKET+1]
  WRaw *d_wraw = nullptr;          // [nwblk][wraw_cap] candidate records
  int wraw_cap = 0;
  int32_t *d_rcount = nullptr;     // [nwblk] records per block
  WItem *d_witems = nullptr;
  size_t witem_cap = 0, nwitems = 0;
  int32_t *d_wchunk = nullptr;     // [cap][3] block, begin, end
  int32_t *d_wcpairs = nullptr;    // [cap] pairs per chunk
  int32_t *d_worder = nullptr;     // [cap] dispatch order (largest first)
  int32_t *d_wccount = nullptr;    // [nwblk+1] chunks per block
  int wnpart = 1;                  // workgroups per block in the predicate pass
  int32_t *d_bpart = nullptr;      // [nwblk][wnpart][NBUCKET] item counts per part
  int nwchunk = 0, wchunk_cap = 0;
  // gradient partials of the real-space water kernel (no global atomics):
  // jpart[row][9] (an item's j-water; the rows of one water contiguous),
  // ipart[chunk][64][9] (the chunk's block waters), both 64-bit fixed
  // point; jstart: row range per slot j, jpos: row of each item
  long long *d_jpart = nullptr, *d_ipart = nullptr;
  int ipart_cap = 0;
  int32_t *d_jcnt = nullptr, *d_jfill = nullptr, *d_jstart = nullptr;
  int32_t *d_jpos = nullptr;       // [item] row of the item in jpart
  int32_t *d_cfirst = nullptr;     // [nwblk+1] first chunk of each block
  int wchunk_pairs = 2048;         // pair budget of one water real-space workgroup
  int achunk_pairs = 8192;         // pair budget of one atom-list workgroup
  // single-system water kernel: workgroups per CU of the short-grid
  // launch (0: one workgroup per chunk) and the resulting grid
  int water_wg_per_cu = 4;
  int water_wgs = 0;
  int wchunk_max_items = 2048;
  // packed per-slot water coordinates for the real-space kernel
  void *d_wpos = nullptr;          // FP32: 3 float4 per slot; FP64: 9 double per slot
  int32_t *d_slot_o = nullptr;     // [nwat] O atom of the slot

  // ---- atom lists (pairs with an atom outside water monomers) ----
  int nsol = 0;                          // atoms of non-water monomers ("solute")
  int nsslot = 0;                        // solute slots: blocks of WBLOCK, padded per brick row
  int nrow = 0, rowcells = 0;            // brick rows of the cell grid, cell keys per row
  int32_t *d_sslot_u = nullptr;          // [nsol] solute atoms in cell order, unpadded
  int32_t *d_rowcnt = nullptr, *d_rowoff = nullptr;  // padded slots per row, their scan
  std::vector<int32_t> h_sol;            // their indices (0-based)
  int32_t *d_sol = nullptr;
  int32_t *d_mono_of = nullptr;          // [npt] monomer (0-based)
  int32_t *d_aflag = nullptr;            // [npt] charged / small LJ / water bits
  uint64_t *d_xmask = nullptr;           // [npt] exclusions l+1 .. l+64
  int32_t *d_xfar = nullptr;             // [npt] has exclusions beyond l+64
  int32_t *d_exc1 = nullptr, *d_exc2 = nullptr;  // MOIL exclusion CSR
  int ag[3] = {0, 0, 0}, ancell = 0;     // all-atom cell grid
  double acs[3] = {0, 0, 0};
  int32_t *d_acell_of = nullptr, *d_astart = nullptr, *d_aslot2atom = nullptr;
  int32_t *d_acnt = nullptr;  // [4][ancell+1] cell counters of the four counting passes
  int32_t *d_sstart = nullptr, *d_sslot2atom = nullptr, *d_sslot_of = nullptr;  // solute slots
  int nablk = 0;                         // solute blocks of WBLOCK slots
  int anpart = 1;                        // parts of a block's atom-list predicate pass
  int32_t *d_abpart = nullptr;           // [nablk][anpart][NBUCKET] item counts per part
  int32_t *d_abcount = nullptr, *d_aboff = nullptr;
  AItem *d_araw = nullptr;               // [nablk][anpart][araw_cap] raw records of the predicate pass
  int araw_cap = 0;
  int32_t *d_arcount = nullptr;          // [nablk][anpart] records per block and part
  int32_t *d_account = nullptr;          // [nablk+1] chunks per block
  AItem *d_aitems = nullptr;
  size_t aitem_cap = 0, naitems = 0;
  int32_t *d_achunk = nullptr;
  int nachunk = 0, achunk_cap = 0;
  long long *d_ajpart = nullptr;         // [row][3] partner rows, fixed point
  long long *d_aipart = nullptr;         // [chunk][64][3] fixed point
  int aipart_cap = 0;
  int32_t *d_ajcnt = nullptr, *d_ajfill = nullptr, *d_ajstart = nullptr, *d_acfirst = nullptr;
  int32_t *d_ajpos = nullptr;

  // ---- accumulators ----
  long long *d_facc = nullptr;     // [3*npt] fixed-point gradient
  double *d_epart = nullptr;       // [EN_COUNT][max_parts] partial energies
  int max_parts = 0;
  double *d_energy = nullptr;      // [MOIL_N_ENERGIES]
  double *d_etot = nullptr;          // [EN_COUNT] per-term totals
  unsigned int *d_ticket = nullptr;  // last-workgroup ticket of k_energies, then the overflow word
  int *d_ovf = nullptr;              // = d_ticket + 1: set by a fixed-point term out of range
  double *d_coor = nullptr;        // [3*npt] staging for host-pointer API
  double *d_dpot = nullptr;        // [3*npt]
  bool have_lists = false;
  long list_gen = 0;               // incremented by every list build
  int32_t *d_scan_ws = nullptr;    // scan workspace (grow-only)
  int32_t *d_status = nullptr;     // [8] sizes / overflow needs of a list build (ST_*)
  size_t scan_ws_cap = 0;

  // ---- parameter gradients (adjoint.hip) ----
  std::vector<double> h_q, h_e12, h_e6;  // current per-atom charges / LJ factors
  std::vector<int32_t> h_sp;              // 1-4 pairs (1-based), [2 totspe]
  int32_t *d_xall = nullptr;              // every exclusion pair (0-based), [2 nxall]
  int nxall = 0;
  double v14 = 0, el14 = 0;               // 1-4 scalings, known after set_params
  bool have_scale14 = false;
  double *d_adj = nullptr;                // [3 npt] host-call outputs dq, de12, de6

  // ---- timing (event pairs per stage, summed on request) ----
  int timing = 0;                  // 0 off, 1 real-space kernels, 2 all stages
  int timing_stride = 1;           // level 1: every timing_stride-th evaluation
  long timing_count = 0;
  long n_timed = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_rec;  // stage, (start, stop)
  size_t ev_next = 0;
  // ---- captured evaluations (api.hip) ----
  bool use_graph = false;          // on with MOIL_GPU_GRAPH=1 in the environment
  EGraph eg;
  hipStream_t gstream = nullptr;   // graph launch stream
  hipEvent_t gev_in = nullptr, gev_out = nullptr;
  long timed_gen = -1;             // list generation of the last timed in-order call
  long param_gen = 0;              // incremented by set_params
  // ---- molecular virial (virial.hip) ----
  int nmol = 0;                    // molecules (PRESS.BLOCK); 0: none set
  int32_t *d_patom = nullptr, *d_ordlist = nullptr;  // [nmol], [npt] (1-based values)
  double *d_ptms = nullptr, *d_massmol = nullptr;    // [npt], [nmol]
  double *d_acom = nullptr;        // [3*npt] centre of mass of each atom's molecule
  double *d_vpart = nullptr;       // [7][vpart_cap] real-space virial partials per chunk
  int vpart_cap = 0;
  double *d_vrec = nullptr, *d_vrel = nullptr;  // [3][blocks] reciprocal partials
  double *d_virial = nullptr;      // [MOIL_N_VIRIAL] host-pointer call's result
};

// exclusive scan of n int32 into out[0..n] (out[n] = total), on stream st
// (lists.hip); workspace grows in the context
int32_t scan_excl(moil_ctx *c, const int32_t *in, int32_t *out, int n, hipStream_t st);

// up to 8 int32 buffers set to a value each, in one launch (lists.hip)
struct ClearSet {
  int32_t *p[8];
  int n[8];
  int32_t v[8];
  int count = 0;
  void add(int32_t *ptr, int len, int32_t val) {
    p[count] = ptr;
    n[count] = len;
    v[count] = val;
    count++;
  }
};
void launch_clear(const ClearSet &cs, hipStream_t st);
#define CHECK_SCAN(x)                        \
  do {                                       \
    const int32_t rc_ = (x);                 \
    if (rc_ != MOIL_GPU_OK) return rc_;      \
  } while (0)
// status words of a list build: raw records needed (if over capacity),
// items and chunks of the water and the atom lists
enum { ST_WRAW = 0, ST_WITEMS, ST_WCHUNK, ST_ARAW, ST_AITEMS, ST_ACHUNK, ST_COUNT };
// chunk table of a list (lists.hip): item runs of one block of about
// `pairs` pairs, min_items..max_items items; cap = table capacity
struct ChunkPar {
  int nblk, pairs, max_items, min_items, cap;
  int nbucket;  // NBUCKET (atom lists) or WNBUCKET (water lists)
  int m2_div;   // pairs counted per listwt2 item: popcount / m2_div
};
int32_t chunk_table(moil_ctx *c, const int32_t *d_boff, ChunkPar P, int32_t *d_ccount,
                    int32_t *d_cfirst, int32_t *d_chunks, int32_t *d_cpairs, int32_t *d_order,
                    hipStream_t st);
bool lists_grow(moil_ctx *c, const int32_t *status);
int32_t lists_finish(moil_ctx *c, const int32_t *status);
bool atoms_grow(moil_ctx *c, const int32_t *status);
int32_t atoms_finish(moil_ctx *c, const int32_t *status);
// list builder (lists.hip)
int32_t lists_init(moil_ctx *c);
int32_t lists_build(moil_ctx *c, const double *d_coor, hipStream_t st);
void lists_free(moil_ctx *c);
// atom lists (atoms.hip)
int32_t atoms_init(moil_ctx *c, const moil_topology *t);
int32_t atoms_build(moil_ctx *c, const double *d_coor, hipStream_t st);
void atoms_free(moil_ctx *c);
// PME set-up (pme.hip)
int32_t pme_init(moil_ctx *c);
void pme_free(moil_ctx *c);
// parameter gradients (adjoint.hip): dq, de12, de6 per atom
int32_t adjoint_run(moil_ctx *c, const double *d_coor, double *d_dq, double *d_de12,
                    double *d_de6, hipStream_t st);
// lists of n systems, their builds queued on the systems' own streams with
// one host wait for all (api.hip)
int32_t build_lists_many(moil_ctx *const *cs, double *const *d_coor, int n, bool squeeze,
                         hipStream_t st, int32_t *d_stat, int32_t *h_stat, hipEvent_t ready,
                         const hipEvent_t *done);
// batched evaluation launched directly on its streams instead of through
// captured graphs (batch.hip)
void batch_set_eager(moil_batch *b, bool on);
// The launches of an evaluation are described by plans (kargs.h).
