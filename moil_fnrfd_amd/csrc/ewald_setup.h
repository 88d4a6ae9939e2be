// MOIL's ewald_init arithmetic on the host (ewald_setup.cpp): Cody's erfc,
// the Ewald-coefficient bisection and the erfc spline table, each in
// MOIL's operation order.
#pragma once

#include <vector>

// erfc(x) as MOIL's CALERF with JINT = 1 (symm/erfcfun.f:14-288)
double moil_erfc(double x);
// dtol after MOIL's range adjustment (symm/ewald_init.f:78-82)
double moil_ewald_dtol(double dtol);
// find_ewaldcof (symm/ewald_init.f:298-335) for the real-space cutoff rc
double moil_ewald_beta(double cutoff, double dtol);
// fill_erf_table + cubspl (symm/ewald_init.f:418-574): n rows of
// MOIL's erf_arr(4,n) at `density` nodes per unit of beta r
void moil_erfc_table(double density, int n, std::vector<double> &tab);
