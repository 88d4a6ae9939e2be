// Launch-shape autotuner (moil_gpu_autotune, moil_gpu_get/set_tuning).
//
// The shapes that decide how one evaluation maps onto the GPU:
//   wchunk_pairs     pair budget of one water real-space workgroup (chunk)
//   achunk_pairs     pair budget of one atom-list workgroup
//   water_wg_per_cu  workgroups per CU of the mixed-precision water kernel's
//                    short grid (0: one workgroup per chunk)
//   pme_tile         largest edge of a spreading tile (grid points, 4..8)
// None of them changes the forces: chunks write their own partial rows and
// the forces are summed in fixed point, so the gradient is bit-identical
// for every shape; the energies are FP64 sums per workgroup, which a chunk
// budget regroups, so they agree to summation order (tested).  The tuner times evaluations of
// the caller's coordinates on the context's stream with HIP events,
// changing one shape at a time over a fixed candidate set (two rounds of
// coordinate descent; a change is kept when it is at least 1 % faster),
// and leaves the best shapes applied with lists rebuilt.  Results are kept
// per device and system signature (atom and water counts, grid, box,
// cutoffs, precision) for the life of the process, and, when
// MOIL_GPU_TUNE_CACHE names a file, appended there and read back by later
// processes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "kargs.h"

namespace {

std::mutex g_tune_mu;
std::map<std::string, moil_tuning> g_tune_cache;
bool g_file_read = false;

std::string signature(const moil_ctx *c) {
  char dev[256] = {0};
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess)
    std::snprintf(dev, sizeof(dev), "%s/%d", prop.gcnArchName, prop.multiProcessorCount);
  char buf[512];
  std::snprintf(buf, sizeof(buf),
                "%s npt=%d nwat=%d grid=%dx%dx%d box=%.4f,%.4f,%.4f cut=%.4f,%.4f,%.4f,%.4f "
                "prec=%d terms=%u",
                dev, c->npt, c->nwat, c->nf1, c->nf2, c->nf3, c->par.a, c->par.b, c->par.c,
                c->par.cutvdw2, c->par.cutele2, c->par.cutvbig2, c->par.cutebig2,
                c->par.precision, c->par.terms);
  for (char *p = buf; *p; p++)
    if (*p == '\n' || *p == '|') *p = ' ';
  return buf;
}

// cache file lines: "<signature>|wchunk achunk wg tile us_default us_tuned"
void read_cache_file() {
  if (g_file_read) return;
  g_file_read = true;
  const char *path = std::getenv("MOIL_GPU_TUNE_CACHE");
  if (!path || !path[0]) return;
  FILE *f = std::fopen(path, "r");
  if (!f) return;
  char line[1024];
  while (std::fgets(line, sizeof(line), f)) {
    char *bar = std::strchr(line, '|');
    if (!bar) continue;
    *bar = 0;
    moil_tuning t;
    if (std::sscanf(bar + 1, "%d %d %d %d %lf %lf", &t.wchunk_pairs, &t.achunk_pairs,
                    &t.water_wg_per_cu, &t.pme_tile, &t.us_default, &t.us_tuned) == 6)
      g_tune_cache[line] = t;
  }
  std::fclose(f);
}

void append_cache_file(const std::string &key, const moil_tuning &t) {
  const char *path = std::getenv("MOIL_GPU_TUNE_CACHE");
  if (!path || !path[0]) return;
  FILE *f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%s|%d %d %d %d %.3f %.3f\n", key.c_str(), t.wchunk_pairs, t.achunk_pairs,
               t.water_wg_per_cu, t.pme_tile, t.us_default, t.us_tuned);
  std::fclose(f);
}

int32_t check(const moil_tuning &t) {
  if (t.wchunk_pairs < 256 || t.wchunk_pairs > 65536 || t.achunk_pairs < 256 ||
      t.achunk_pairs > 65536 || t.water_wg_per_cu < 0 || t.water_wg_per_cu > 16 ||
      t.pme_tile < 4 || t.pme_tile > 8) {
    set_error("tuning out of range: chunk budgets 256..65536 pairs, 0..16 workgroups per CU, "
              "PME tile edge 4..8");
    return MOIL_GPU_EINVAL;
  }
  return MOIL_GPU_OK;
}

moil_tuning current(const moil_ctx *c) {
  moil_tuning t;
  t.wchunk_pairs = c->wchunk_pairs;
  t.achunk_pairs = c->achunk_pairs;
  t.water_wg_per_cu = c->water_wg_per_cu;
  t.pme_tile = c->pt_edge > 0 ? c->pt_edge : 8;
  t.us_default = t.us_tuned = 0.0;
  return t;
}

// the shapes into the context; a new tile shape invalidates the atoms'
// tile assignment, so the lists must be built again before an evaluation
int32_t apply(moil_ctx *c, const moil_tuning &t) {
  const int32_t rc = check(t);
  if (rc != MOIL_GPU_OK) return rc;
  c->wchunk_pairs = t.wchunk_pairs;
  c->achunk_pairs = t.achunk_pairs;
  c->water_wg_per_cu = t.water_wg_per_cu;
  c->water_wgs = t.water_wg_per_cu * c->ncu;
  const bool pme_on = (c->par.terms & (MOIL_TERM_VDW | MOIL_TERM_EL)) != 0;
  if (pme_on && c->d_tbuf && t.pme_tile != c->pt_edge) {
    const int nt0[3] = {c->pt_nt[0], c->pt_nt[1], c->pt_nt[2]};
    const int32_t r2 = pme_set_tile(c, t.pme_tile);
    if (r2 != MOIL_GPU_OK) return r2;
    if (nt0[0] != c->pt_nt[0] || nt0[1] != c->pt_nt[1] || nt0[2] != c->pt_nt[2])
      c->have_lists = false;
  }
  return MOIL_GPU_OK;
}

}  // namespace

extern "C" int32_t moil_gpu_get_tuning(moil_ctx *c, moil_tuning *t) {
  if (!c || !t) {
    set_error("null argument");
    return MOIL_GPU_EINVAL;
  }
  *t = current(c);
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_set_tuning(moil_ctx *c, const moil_tuning *t) {
  if (!c || !t) {
    set_error("null argument");
    return MOIL_GPU_EINVAL;
  }
  return apply(c, *t);
}

extern "C" int32_t moil_gpu_autotune(moil_ctx *c, const double *d_coor, int32_t reps,
                                     moil_tuning *out) {
  if (!c || !d_coor) {
    set_error("null argument");
    return MOIL_GPU_EINVAL;
  }
  if (reps <= 0) reps = 20;
  reps = std::min(reps, 1000);
  hipStream_t st = c->stream;
  const std::string key = signature(c);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    read_cache_file();
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      int32_t rc = apply(c, it->second);
      if (rc == MOIL_GPU_OK) rc = moil_gpu_build_lists_dev(c, d_coor, st);
      if (rc == MOIL_GPU_OK && out) *out = it->second;
      return rc;
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) hipEventDestroy(e0);
    set_error("hipEventCreate failed");
    return MOIL_GPU_EHIP;
  }
  int32_t rc = MOIL_GPU_OK;
  // mean time of one evaluation (us) with shapes t: lists built with them,
  // two untimed evaluations, then reps timed ones
  auto time_of = [&](const moil_tuning &t, double *us) -> int32_t {
    int32_t r = apply(c, t);
    if (r == MOIL_GPU_OK) r = moil_gpu_build_lists_dev(c, d_coor, st);
    for (int k = 0; r == MOIL_GPU_OK && k < 2; k++)
      r = moil_gpu_eforce_dev(c, d_coor, c->d_dpot, c->d_energy, st);
    if (r != MOIL_GPU_OK) return r;
    if (hipEventRecord(e0, st) != hipSuccess) return MOIL_GPU_EHIP;
    for (int k = 0; r == MOIL_GPU_OK && k < reps; k++)
      r = moil_gpu_eforce_dev(c, d_coor, c->d_dpot, c->d_energy, st);
    if (r != MOIL_GPU_OK) return r;
    float ms = 0.f;
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      set_error("autotune: event timing failed");
      return MOIL_GPU_EHIP;
    }
    *us = 1000.0 * ms / reps;
    return MOIL_GPU_OK;
  };
  const moil_tuning def = current(c);
  moil_tuning best = def;
  double t_def = 0.0, t_best = 0.0;
  rc = time_of(def, &t_def);
  t_best = t_def;
  const bool pme_on = (c->par.terms & (MOIL_TERM_VDW | MOIL_TERM_EL)) != 0;
  const bool mixed = c->par.precision == MOIL_PREC_MIXED;
  const int wg[] = {0, 2, 3, 4, 5, 6, 8};
  const int wchunk[] = {1024, 2048, 4096, 8192};
  const int achunk[] = {2048, 4096, 8192, 16384};
  const int tile[] = {5, 6, 7, 8};
  for (int pass = 0; rc == MOIL_GPU_OK && pass < 2; pass++) {
    for (int dim = 0; rc == MOIL_GPU_OK && dim < 4; dim++) {
      const int *vals = dim == 0 ? wg : (dim == 1 ? wchunk : (dim == 2 ? achunk : tile));
      const int nv = dim == 0 ? 7 : 4;
      if (dim == 0 && (!mixed || c->nwat == 0)) continue;
      if (dim == 1 && c->nwat == 0) continue;
      if (dim == 2 && 3 * c->nwat == c->npt) continue;  // waters only: no atom lists
      if (dim == 3 && !pme_on) continue;
      for (int v = 0; rc == MOIL_GPU_OK && v < nv; v++) {
        moil_tuning t = best;
        int *field = dim == 0 ? &t.water_wg_per_cu
                              : (dim == 1 ? &t.wchunk_pairs
                                          : (dim == 2 ? &t.achunk_pairs : &t.pme_tile));
        if (*field == vals[v]) continue;
        *field = vals[v];
        double us = 0.0;
        rc = time_of(t, &us);
        if (rc == MOIL_GPU_EINVAL) {  // a tile edge this grid cannot use: skip it
          rc = MOIL_GPU_OK;
          continue;
        }
        if (rc == MOIL_GPU_OK && us < 0.99 * t_best) {
          best = t;
          t_best = us;
        }
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != MOIL_GPU_OK) {
    // leave the context as it was
    if (apply(c, def) == MOIL_GPU_OK) moil_gpu_build_lists_dev(c, d_coor, st);
    return rc;
  }
  best.us_default = t_def;
  best.us_tuned = t_best;
  rc = apply(c, best);
  if (rc == MOIL_GPU_OK) rc = moil_gpu_build_lists_dev(c, d_coor, st);
  if (rc != MOIL_GPU_OK) return rc;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[key] = best;
    append_cache_file(key, best);
  }
  if (out) *out = best;
  return MOIL_GPU_OK;
}
