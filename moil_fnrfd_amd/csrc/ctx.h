// Library context: device copies of the topology, Ewald tables, pair
// lists and work buffers for one MOIL system.
#pragma once

#include <rocfft/rocfft.h>

#include <vector>

#include "common.h"

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
};

struct moil_ctx {
  moil_params par;
  int npt = 0, totmon = 0, nwat = 0, nb = 0, nangl = 0, ntors = 0, nimp = 0, totspe = 0;
  int nexcl = 0;
  hipStream_t stream = nullptr;   // default stream of the context
  hipStream_t stream2 = nullptr;  // reciprocal space, overlapped with real space
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
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
  int32_t *d_wat_o = nullptr;     // [nwat] O atom (0-based)
  int32_t *d_wat_mono = nullptr;  // [nwat] monomer number (1-based)
  std::vector<int32_t> h_wat_mono, h_wat_o;
  double qq_oo = 0, qq_oh = 0, qq_hh = 0, oo12 = 0, oo6 = 0;  // last-water parameters
  bool has_nonwater = false;

  // ---- Ewald ----
  double beta = 0, e_self = 0, erftbdns = 5000.0;
  int ntab = 0;
  std::vector<double> h_erf;   // [ntab][4]
  double *d_erf = nullptr;
  double *d_eterm = nullptr;   // influence function on the half spectrum
  std::vector<double> bsp[3];
  int nf1 = 0, nf2 = 0, nf3 = 0, nc1 = 0;  // grid dims; nc1 = nf1/2+1
  int32_t *d_pcnt = nullptr;    // [G+1] atoms per stencil-base grid point
  int32_t *d_pstart = nullptr;  // [G+1] exclusive scan of d_pcnt
  int32_t *d_pbsum = nullptr;   // scan block sums
  int32_t *d_abin = nullptr, *d_arank = nullptr;  // [npt] base point, rank in bin
  void *d_aw = nullptr, *d_sw = nullptr;  // [npt][4] (w1 w2 w3 q*fc), atom / bin order
  int32_t *d_sx = nullptr;      // [npt] base x index, bin order
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
  WItem *d_witems = nullptr;
  size_t witem_cap = 0, nwitems = 0;
  int32_t *d_wchunk = nullptr;     // [nchunk][3]: block, begin, end
  int nwchunk = 0;
  int wchunk_items = 256;
  // packed per-slot water coordinates for the real-space kernel
  void *d_wpos = nullptr;          // FP32: 3 float4 per slot; FP64: 9 double per slot
  int32_t *d_slot_o = nullptr;     // [nwat] O atom of the slot

  // ---- atom lists (non-water monomers) ----
  int natblk = 0;
  int32_t *d_aslot2atom = nullptr;
  AItem *d_aitems = nullptr;
  size_t aitem_cap = 0, naitems = 0;
  int32_t *d_achunk = nullptr;
  int nachunk = 0;
  int32_t *d_abcount = nullptr, *d_aboff = nullptr;
  std::vector<int32_t> h_apartner_mono;  // (unused placeholder)

  // ---- accumulators ----
  long long *d_facc = nullptr;     // [3*npt] fixed-point gradient
  double *d_epart = nullptr;       // [EN_COUNT][max_parts] partial energies
  int max_parts = 0;
  double *d_energy = nullptr;      // [MOIL_N_ENERGIES]
  double *d_coor = nullptr;        // [3*npt] staging for host-pointer API
  double *d_dpot = nullptr;        // [3*npt]
  bool have_lists = false;

  // ---- timing (event pairs per stage, summed on request) ----
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_rec;  // stage, (start, stop)
  size_t ev_next = 0;
};

// list builder (lists.hip)
int32_t lists_init(moil_ctx *c);
int32_t lists_build(moil_ctx *c, const double *d_coor, hipStream_t st);
void lists_free(moil_ctx *c);
// real space (realspace.hip)
int32_t realspace_water_pack(moil_ctx *c, const double *d_coor, hipStream_t st);
int32_t realspace_water(moil_ctx *c, const double *d_coor, hipStream_t st);
int32_t realspace_atoms(moil_ctx *c, const double *d_coor, hipStream_t st);
// PME (pme.hip)
int32_t pme_init(moil_ctx *c);
int32_t pme_run(moil_ctx *c, const double *d_coor, hipStream_t st);
void pme_free(moil_ctx *c);
// bonded, exclusion correction, 1-4 (bonded.hip)
int32_t bonded_run(moil_ctx *c, const double *d_coor, hipStream_t st);
