// sns_poisson.h -- four-step FFT plan and launcher of the Poisson solve.
#pragma once
#include "sns_launch.h"
#include "sns_fast.h"

namespace sns {

struct Fft2Plan {
  int n;                         // transform length (jtot) = n1*n2
  int n1, n2;
  int np1, np2;                  // radix passes of the two factors
  unsigned long long code1, code2;
  int npair;                     // complex column pairs, (itot+1)/2
  int G;                         // rows j2 per workgroup in passes A and A'
};

// four-step plan of length n for itot columns; n1 > 0 fixes the first
// factor (false when the factors or the LDS tiles do not fit)
bool fft2_plan(int n, int itot, Fft2Plan &pl, int n1 = 0);

// Thomas factors of the x systems of all wavenumbers (once per instance):
// fac holds (jtot/2+1) tables of poisson_factor_entries(itot) reciprocal
// pivots (sns_lu.h layout), then u'_0 of each wavenumber, the row pfx after
// which the pivots repeat up to row itot-2 and that pivot (LuTabLUr); plu
// the shared row coefficients l and u in the same layout (poisson_lu_coefs,
// 2 tables)
int poisson_factor_entries(int itot);
size_t poisson_factor_doubles(int itot, int jtot);
void poisson_lu_coefs(int itot, const double *pl, const double *pu, double *plu);
void launch_poisson_factor(hipStream_t s, const Geo &g, const Tables &tb, const double *eigen,
                           double *fac);

// Arguments of the Poisson kernels for one instance: ux*, uy* in the layout
// of g, the spectrum psihat (x-aligned, jtot/2+1 rows of itot), the four-step
// intermediate Y, the Thomas factors fac and psi with boundary values.
struct PoisArgs {
  Geo g;
  Tables tb;
  Fft2Plan pl;
  double gamma;
  const double *dtp;
  const double *ux, *uy;
  const double2 *tw;
  const double *fac, *plu;
  double2 *Y, *psihat;
  double *psi;
};

// stage 0: RHS + forward transform -> psihat (x-aligned spectrum)
// stage 1: tri-diagonal solves along x, one per wavenumber (factors in fac)
// stage 2: inverse transform -> psi with boundary values
// stage < 0: all three
void launch_poisson(hipStream_t s, const PoisArgs &a, int stage);
// the same for nz instances of the shape of `shape` (same itot, jtot), their
// arguments in the device array dev
void launch_poisson_batch(hipStream_t s, const PoisArgs &shape, const PoisArgs *dev, int nz,
                          int stage);

}  // namespace sns
