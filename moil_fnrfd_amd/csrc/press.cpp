// MOIL's division of a system into molecules for the molecular virial
// (COMMON/PRESS.BLOCK), for callers without MOIL's Fortran: the tables
// init_pressure fills (analysis/init_pressure.f:1-77) by its breadth-first
// search over the bonds (create_mol_BFS, analysis/init_pressure.f:79-187).
//   * Molecules are found in order of their smallest atom: starting from
//     the smallest atom not yet in a molecule, every atom reachable over
//     bonds joins it, and its atoms are listed in ascending order (the
//     binary-tree sort of analysis/init_pressure.f:236-309).
//   * A water monomer (MOIL's TIP3/SPC names; its bonds may be missing
//     when its geometry is held by matrix SHAKE) is the molecule of its
//     three consecutive atoms without a search.
//   * The search stops once its start atom is the last atom, as MOIL's
//     loop (do while ipt < npt) does: a last atom that no earlier
//     molecule reached is left in no molecule (pmol 0), which
//     moil_gpu_set_molecules refuses.
// Host only.
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "moil_gpu.h"

void set_error(const std::string &msg);

extern "C" int32_t moil_molecules(const moil_topology *t, int32_t *nmol, int32_t *patom,
                                  int32_t *ordlist, int32_t *pmol) {
  if (!t || !nmol || !patom || !ordlist || !pmol || t->npt <= 0 || !t->poipt ||
      !t->monoflag || (t->nb > 0 && (!t->ib1 || !t->ib2))) {
    set_error("moil_molecules: null or empty topology");
    return MOIL_GPU_EINVAL;
  }
  const int npt = t->npt;
  // monomer of each atom, and the bonds as an adjacency list
  std::vector<int> mono(npt + 1, 0);
  for (int m = 1; m <= t->totmon; m++)
    for (int a = t->poipt[m - 1] + 1; a <= t->poipt[m] && a <= npt; a++) mono[a] = m;
  std::vector<int> deg(npt + 2, 0), adj;
  for (int b = 0; b < t->nb; b++) {
    const int i = t->ib1[b], j = t->ib2[b];
    if (i < 1 || i > npt || j < 1 || j > npt) {
      set_error("moil_molecules: bond atom out of range");
      return MOIL_GPU_EINVAL;
    }
    deg[i + 1]++;
    deg[j + 1]++;
  }
  for (int a = 1; a <= npt; a++) deg[a + 1] += deg[a];
  adj.resize(deg[npt + 1]);
  {
    std::vector<int> fill(deg.begin(), deg.end());
    for (int b = 0; b < t->nb; b++) {
      adj[fill[t->ib1[b]]++] = t->ib2[b];
      adj[fill[t->ib2[b]]++] = t->ib1[b];
    }
  }
  for (int a = 0; a < npt; a++) pmol[a] = 0;
  std::vector<int> list;
  std::vector<char> seen(npt + 1, 0);
  int imol = 0, nord = 0, ipt = 1;
  while (ipt < npt) {
    list.clear();
    const int m = mono[ipt];
    if (m > 0 && (t->monoflag[m - 1] & MOIL_MONO_WATER) && ipt + 2 <= npt) {
      list = {ipt, ipt + 1, ipt + 2};
    } else {
      // the connected component of ipt (its atoms' order is sorted below)
      list.push_back(ipt);
      seen[ipt] = 1;
      for (size_t q = 0; q < list.size(); q++)
        for (int e = deg[list[q]]; e < deg[list[q] + 1]; e++)
          if (!seen[adj[e]]) {
            seen[adj[e]] = 1;
            list.push_back(adj[e]);
          }
      std::sort(list.begin(), list.end());
    }
    imol++;
    for (int a : list) {
      ordlist[nord++] = a;
      pmol[a - 1] = imol;
    }
    patom[imol - 1] = nord;
    // the next start: the smallest atom after ipt in no molecule yet
    int i = ipt + 1;
    while (i <= npt && pmol[i - 1] != 0) i++;
    ipt = i;
  }
  *nmol = imol;
  return MOIL_GPU_OK;
}
