// Host side of MOIL's ewald_init (symm/ewald_init.f:8-242) that has to
// reproduce MOIL's numbers to the last bit: the complementary error
// function MOIL evaluates (W. J. Cody's rational Chebyshev approximations,
// "Rational Chebyshev approximations for the error function", Math. Comp.
// 23 (1969) 631-638, the NETLIB SPECFUN routine CALERF that MOIL ships as
// symm/erfcfun.f), the bisection for the Ewald coefficient
// (symm/ewald_init.f:298-335) and the clamped cubic spline of the erfc
// table (fill_erf_table + de Boor's cubspl, "A Practical Guide to
// Splines", symm/ewald_init.f:418-574).
//
// Every expression keeps MOIL's operation order and is compiled without
// floating-point contraction (-ffp-contract=off), so the table, beta and
// everything derived from them are the bytes MOIL computes with the same C
// library's exp.
#include <cmath>
#include <cstdint>
#include <vector>

#include "ewald_setup.h"

namespace {

// Cody's coefficients (the published double-precision set)
// erf on |x| <= 0.46875
const double kA[5] = {3.16112374387056560e00, 1.13864154151050156e02,
                      3.77485237685302021e02, 3.20937758913846947e03,
                      1.85777706184603153e-1};
const double kB[4] = {2.36012909523441209e01, 2.44024637934444173e02,
                      1.28261652607737228e03, 2.84423683343917062e03};
// erfc on 0.46875 <= |x| <= 4
const double kC[9] = {5.64188496988670089e-1, 8.88314979438837594e0,
                      6.61191906371416295e01, 2.98635138197400131e02,
                      8.81952221241769090e02, 1.71204761263407058e03,
                      2.05107837782607147e03, 1.23033935479799725e03,
                      2.15311535474403846e-8};
const double kD[8] = {1.57449261107098347e01, 1.17693950891312499e02,
                      5.37181101862009858e02, 1.62138957456669019e03,
                      3.29079923573345963e03, 4.36261909014324716e03,
                      3.43936767414372164e03, 1.23033935480374942e03};
// erfc on |x| > 4 (rational in 1/x^2)
const double kP[6] = {3.05326634961232344e-1, 3.60344899949804439e-1,
                      1.25781726111229246e-1, 1.60837851487422766e-2,
                      6.58749161529837803e-4, 1.63153871373020978e-2};
const double kQ[5] = {2.56852019228982242e00, 1.87295284992346047e00,
                      5.27905102951428412e-1, 6.05183413124413191e-2,
                      2.33520497626869185e-3};
const double kSqrtPiInv = 5.6418958354775628695e-1;  // 1/sqrt(pi)
const double kThresh = 0.46875;
const double kXsmall = 1.11e-16;  // IEEE double: 1 + x == 1 below it
const double kXbig = 26.543;      // erfc underflows beyond it

// exp(-y^2) as exp(-ys^2) exp(-(y-ys)(y+ys)) with ys = y truncated to a
// multiple of 1/16: the split that keeps the product accurate
inline double gauss_split(double y) {
  const double ys = std::trunc(y * 16.0) / 16.0;
  const double del = (y - ys) * (y + ys);
  return std::exp(-ys * ys) * std::exp(-del);
}

}  // namespace

double moil_erfc(double x) {
  const double y = std::fabs(x);
  double r;
  if (y <= kThresh) {
    // erf by a rational function in x^2, erfc = 1 - erf
    const double ysq = y > kXsmall ? y * y : 0.0;
    double num = kA[4] * ysq, den = ysq;
    for (int i = 0; i < 3; i++) {
      num = (num + kA[i]) * ysq;
      den = (den + kB[i]) * ysq;
    }
    r = x * (num + kA[3]) / (den + kB[3]);
    return 1.0 - r;
  }
  if (y <= 4.0) {
    double num = kC[8] * y, den = y;
    for (int i = 0; i < 7; i++) {
      num = (num + kC[i]) * y;
      den = (den + kD[i]) * y;
    }
    r = (num + kC[7]) / (den + kD[7]);
    r = gauss_split(y) * r;
  } else if (y >= kXbig) {
    r = 0.0;
  } else {
    const double ysq = 1.0 / (y * y);
    double num = kP[5] * ysq, den = ysq;
    for (int i = 0; i < 4; i++) {
      num = (num + kP[i]) * ysq;
      den = (den + kQ[i]) * ysq;
    }
    r = ysq * (num + kP[4]) / (den + kQ[4]);
    r = (kSqrtPiInv - r) / y;
    r = gauss_split(y) * r;
  }
  return x < 0.0 ? 2.0 - r : r;
}

double moil_ewald_dtol(double dtol) {
  // symm/ewald_init.f:53-82: tolerances outside [1e-10, 0.1] replaced
  if (dtol < 1e-10) return 1e-10;
  if (dtol > 0.1) return 5.2e-05;
  return dtol;
}

double moil_ewald_beta(double cutoff, double dtol) {
  // double x until erfc(x rc)/rc < dtol, then bisect 60 more times than
  // the doublings took
  double x = 0.5;
  int steps = 0;
  for (;;) {
    x = 2.0 * x;
    steps++;
    if (moil_erfc(x * cutoff) / cutoff < dtol) break;
  }
  double lo = 0.0, hi = x;
  for (int i = 0; i < steps + 60; i++) {
    x = (lo + hi) / 2;
    if (moil_erfc(x * cutoff) / cutoff >= dtol) lo = x;
    else hi = x;
  }
  return x;
}

void moil_erfc_table(double density, int n, std::vector<double> &tab) {
  // rows i = 0..n-1 at x_i = i / density: tab[4i] = erfc(x_i) and, after
  // the spline, tab[4i+1..3] = first, second and third derivative of the
  // cubic on [x_i, x_{i+1}] (the last row: the end slope, the abscissa
  // step and the last divided difference, as cubspl leaves them)
  tab.assign((size_t)4 * n, 0.0);
  if (n <= 0) return;
  const double del = 1.0 / density;
  const double fac = 2.0 / std::sqrt(3.14159265358979323846);
  std::vector<double> tau(n);
  double *c = tab.data();
  auto C = [c](int j, int i) -> double & { return c[4 * (i - 1) + (j - 1)]; };  // c(j,i)
  C(2, 1) = -fac;
  {
    const double xe = (n - 1) * del;
    C(2, n) = -fac * std::exp(-xe * xe);
  }
  for (int i = 1; i <= n; i++) {
    const double x = del * (i - 1);
    tau[i - 1] = x;
    C(1, i) = moil_erfc(x);
  }
  if (n < 3) return;  // MOIL's table always has thousands of rows
  // de Boor's tridiagonal system for the slopes, both end slopes given:
  // c(3,m) = step, c(4,m) = divided difference; forward elimination
  for (int m = 2; m <= n; m++) {
    C(3, m) = tau[m - 1] - tau[m - 2];
    C(4, m) = (C(1, m) - C(1, m - 1)) / C(3, m);
  }
  C(4, 1) = 1.0;
  C(3, 1) = 0.0;
  for (int m = 2; m <= n - 1; m++) {
    const double g = -C(3, m + 1) / C(4, m - 1);
    C(2, m) = g * C(2, m - 1) + 3.0 * (C(3, m) * C(4, m + 1) + C(3, m + 1) * C(4, m));
    C(4, m) = g * C(3, m - 1) + 2.0 * (C(3, m) + C(3, m + 1));
  }
  // back substitution down to the first row
  for (int j = n - 1; j >= 1; j--) C(2, j) = (C(2, j) - C(3, j) * C(2, j + 1)) / C(4, j);
  // Taylor coefficients of each piece from its end values and slopes
  for (int i = 2; i <= n; i++) {
    const double dtau = C(3, i);
    const double d1 = (C(1, i) - C(1, i - 1)) / dtau;
    const double d3 = C(2, i - 1) + C(2, i) - 2.0 * d1;
    C(3, i - 1) = 2.0 * (d1 - C(2, i - 1) - d3) / dtau;
    C(4, i - 1) = (d3 / dtau) * (6.0 / dtau);
  }
}
