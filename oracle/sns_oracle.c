/*
 * sns_oracle.c -- CPU oracle (test infrastructure, see sns_oracle.h).
 *
 * A clean-room restatement of the reference's hot path for a single rank.
 * Arithmetic follows the reference expression by expression (same operand
 * order, IEEE division where the reference divides, no contraction: the
 * Makefile passes -ffp-contract=off) so that stencil results are bitwise
 * comparable.  Scratch is allocated per call, as in the reference.  Loops
 * over independent rows, columns and wavenumbers run under OpenMP; no
 * reduction is parallelised, so results do not depend on the thread count.
 *
 * The FFT (FFTW3 in the reference, src/fluid/compute_potential.c:53-67,90,149)
 * is not part of the reference sources; it is restated here from its
 * published conventions: forward r2c Y_J = sum_j x_j exp(-2 pi i jJ/n),
 * J = 0..n/2; backward c2r unnormalised.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include "sns_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.1415926535897932
#endif

#define TEMP_XM (+0.5)
#define TEMP_XP (-0.5)

/* index macros with the reference extents (include/arrays/ headers), jsize == jtot */
#define XF(I)  (param->xf[(I)-1])
#define XC(I)  (param->xc[(I)])
#define DXF(I) (param->dxf[(I)-1])
#define DXC(I) (param->dxc[(I)-1])
#define YF(J)  (param->yf[(J)-1])
#define YC(J)  (param->yc[(J)-1])
#define UX(I, J)   (ux[(J)*(itot+1)+((I)-1)])
#define UY(I, J)   (uy[(J)*(itot+2)+(I)])
#define P(I, J)    (p[(J)*(itot+2)+(I)])
#define PSI(I, J)  (psi[(J)*(itot+2)+(I)])
#define TEMP(I, J) (temp[(J)*(itot+2)+(I)])
#define SRCUXA(I, J) (srcuxa[((J)-1)*(itot-1)+((I)-2)])
#define SRCUXB(I, J) (srcuxb[((J)-1)*(itot-1)+((I)-2)])
#define SRCUXG(I, J) (srcuxg[((J)-1)*(itot-1)+((I)-2)])
#define SRCUYA(I, J) (srcuya[((J)-1)*(itot  )+((I)-1)])
#define SRCUYB(I, J) (srcuyb[((J)-1)*(itot  )+((I)-1)])
#define SRCUYG(I, J) (srcuyg[((J)-1)*(itot  )+((I)-1)])
#define TEMPFORCEX(I, J) (tempforcex[((J)-1)*(itot-1)+((I)-2)])
#define SRCTEMPA(I, J) (srctempa[((J)-1)*(itot)+((I)-1)])
#define SRCTEMPB(I, J) (srctempb[((J)-1)*(itot)+((I)-1)])
#define SRCTEMPG(I, J) (srctempg[((J)-1)*(itot)+((I)-1)])

/* build a complex from its parts (C99 6.2.5p13: array of two reals) */
static double complex mkz(double re, double im){
  double complex z = 0;
  ((double *)&z)[0] = re;
  ((double *)&z)[1] = im;
  return z;
}

static void *xcalloc(size_t n, size_t sz){
  void *ptr = calloc(n ? n : 1, sz);
  if(ptr == NULL) abort();
  return ptr;
}

/* Hooks for the y-slab MPI restatement (sns_oracle_mpi.c, which includes
 * this file): the halo exchange, the periodic y solve of an x-aligned
 * increment and the global minimum of the advective bound.  NULL (this
 * library): the single-rank versions. */
static void (*g_halo_rows)(double *base, size_t pitch, int first_col, int cnt, int jtot) = NULL;
static void (*g_solve_y)(int ni, int jtot, double prefactor, double dy, double *qx) = NULL;
static void (*g_min_allreduce)(double *v) = NULL;

/* ---------------------------------------------------------------- param */

/* src/param/init.c:124-195 */
int orc_set_coordinate(orc_param_t *param){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double lx = param->lx;
  const double ly = param->ly;
  {
    int stretch = (int)param->stretch; /* truncation as in init.c:144 */
    if(stretch < 0) stretch = 0;
    if(stretch >= itot/2){
      const double dx = lx/itot;
      for(int i=1; i<=itot+1; i++) XF(i) = 1.*(i-1)*dx;
    }else{
      for(int i=1; i<=itot+1; i++) XF(i) = cos(M_PI*(1.*(i-1)+stretch)/(itot+2.*stretch));
      double lo = XF(1);
      double hi = XF(itot+1);
      for(int i=1; i<=itot+1; i++) XF(i) = lx*(XF(i)-lo)/(hi-lo);
    }
    XF(1) = 0.;
    XF(itot+1) = lx;
  }
  for(int i=1; i<=itot; i++) DXF(i) = XF(i+1)-XF(i);
  XC(0) = XF(1);
  for(int i=1; i<=itot; i++) XC(i) = 0.5*(XF(i)+XF(i+1));
  XC(itot+1) = XF(itot+1);
  for(int i=1; i<=itot+1; i++) DXC(i) = XC(i)-XC(i-1);
  param->dy = ly/jtot;
  for(int j=1; j<=jtot+1; j++) YF(j) = 1.*(j-1)*param->dy;
  for(int j=1; j<=jtot; j++) YC(j) = 0.5*(2*j-1)*param->dy;
  return 0;
}

/* src/param/init.c:197-213 */
int orc_set_rk_coefs(orc_param_t *param){
  param->alpha[0] = 32./60.; param->beta[0] =   0./60.;
  param->alpha[1] = 25./60.; param->beta[1] = -17./60.;
  param->alpha[2] = 45./60.; param->beta[2] = -25./60.;
  for(int k=0; k<3; k++) param->gamma[k] = param->alpha[k]+param->beta[k];
  return 0;
}

/* src/param/decide_dt.c:11-101; the treatment is chosen from
 * param->expimp_wtimes with the reference's own rule (68-97) */
int orc_param_decide_dt(orc_param_t *param, const orc_state_t *s){
  const double safe_factors[2] = {0.7, 0.9};
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double diffusivity = fmax(sqrt(Pr)/sqrt(Ra), 1./sqrt(Pr)/sqrt(Ra));
  const double lx = param->lx;
  const double dy = param->dy;
  const double *ux = s->ux;
  const double *uy = s->uy;
  double dt_adv;
  {
    const double small = 1.e-8;
    dt_adv = 5.e-2;
    for(int j=1; j<=jtot; j++){
      for(int i=2; i<=itot; i++){
        double vel = fmax(fabs(UX(i, j)), small);
        dt_adv = fmin(dt_adv, DXC(i)/vel);
      }
    }
    for(int j=1; j<=jtot; j++){
      for(int i=1; i<=itot; i++){
        double vel = fmax(fabs(UY(i, j)), small);
        dt_adv = fmin(dt_adv, dy/vel);
      }
    }
    if(g_min_allreduce) g_min_allreduce(&dt_adv);  /* MPI_Allreduce MIN over the slabs */
    dt_adv *= safe_factors[0];
  }
  double dt_dif_x, dt_dif_y;
  {
    double dx = lx;
    for(int i=2; i<=itot; i++) dx = fmin(dx, DXC(i));
    dt_dif_x = 0.25/diffusivity*pow(dx, 2.);
    dt_dif_y = 0.25/diffusivity*pow(dy, 2.);
    dt_dif_x *= safe_factors[1];
    dt_dif_y *= safe_factors[1];
  }
  double dt = 1.e-2;
  {
    double optimal_wtime = DBL_MAX;
    int optimal_implicitx = -1;
    int optimal_implicity = -1;
    for(int implicity=0; implicity<2; implicity++){
      for(int implicitx=0; implicitx<2; implicitx++){
        double dt_this =
          implicitx ? implicity ? dt_adv
                                : fmin(dt_adv, dt_dif_y)
                    : implicity ? fmin(dt_adv, dt_dif_x)
                                : fmin(dt_adv, fmin(dt_dif_x, dt_dif_y));
        double niter = 1./dt_this;
        double wtime = param->expimp_wtimes[implicity][implicitx]*niter;
        if(wtime < optimal_wtime){
          dt = dt_this;
          optimal_wtime = wtime;
          optimal_implicitx = implicitx;
          optimal_implicity = implicity;
        }
      }
    }
    param->implicitx = optimal_implicitx;
    param->implicity = optimal_implicity;
  }
  param->dt = dt;
  return 0;
}

/* ------------------------------------------------------ boundary values */

/* single-rank halo exchange (src/parallel/halo.c:6-22): row 0 <- row jsize,
 * row jsize+1 <- row 1, over cnt entries starting at the given column */
static void halo_rows_local(double *base, size_t pitch, int first_col, int cnt, int jtot){
  memcpy(base + 0*pitch + first_col, base + (size_t)jtot*pitch + first_col, sizeof(double)*cnt);
  memcpy(base + (size_t)(jtot+1)*pitch + first_col, base + 1*pitch + first_col, sizeof(double)*cnt);
}


static void halo_rows(double *base, size_t pitch, int first_col, int cnt, int jtot){
  if(g_halo_rows) g_halo_rows(base, pitch, first_col, cnt, jtot);
  else halo_rows_local(base, pitch, first_col, cnt, jtot);
}

/* src/fluid/boundary_conditions.c:6-19 */
int orc_fluid_update_boundaries_ux(const orc_param_t *param, double *ux){
  const int itot = param->itot;
  const int jtot = param->jtot;
  halo_rows(ux, itot+1, 1, itot-1, jtot); /* &UX(2, .) == ux[J*(itot+1)+1] */
  for(int j=0; j<=jtot+1; j++){
    UX(     1, j) = 0.;
    UX(itot+1, j) = 0.;
  }
  return 0;
}

/* src/fluid/boundary_conditions.c:21-37 */
int orc_fluid_update_boundaries_uy(const orc_param_t *param, double *uy){
  const int itot = param->itot;
  const int jtot = param->jtot;
  halo_rows(uy, itot+2, 1, itot, jtot);
  for(int j=0; j<=jtot+1; j++){
    UY(     0, j) = 0.;
    UY(itot+1, j) = 0.;
  }
  return 0;
}

/* src/fluid/boundary_conditions.c:39-55 */
int orc_fluid_update_boundaries_p(const orc_param_t *param, double *p){
  const int itot = param->itot;
  const int jtot = param->jtot;
  halo_rows(p, itot+2, 1, itot, jtot);
  for(int j=0; j<=jtot+1; j++){
    P(     0, j) = P(   1, j);
    P(itot+1, j) = P(itot, j);
  }
  return 0;
}

/* src/temperature/boundary_conditions.c:6-21 */
int orc_temperature_update_boundaries_temp(const orc_param_t *param, double *temp){
  const int itot = param->itot;
  const int jtot = param->jtot;
  halo_rows(temp, itot+2, 1, itot, jtot);
  for(int j=1; j<=jtot; j++){
    TEMP(     0, j) = TEMP_XM;
    TEMP(itot+1, j) = TEMP_XP;
  }
  return 0;
}

/* ------------------------------------------------------------- variants */

/* Alternative operation orders for the same arithmetic (default 0: the
 * reference's).  ORC_VARIANT_DIRECT_DFT sums the Poisson transforms as a
 * direct O(n^2) DFT instead of the mixed-radix FFT; ORC_VARIANT_REVERSED_SOLVE
 * eliminates every tri-diagonal system from its last row to its first;
 * ORC_VARIANT_PIN_GAUGE replaces the last equation of the singular J = 0
 * Poisson system by psi_hat(itot) = 0 instead of letting the rounding-level
 * last pivot set the additive constant (src/linalg.c:36-43).  Tests
 * compare runs with and without them to measure how far last-bit differences
 * travel in a given configuration (its rounding floor). */
static int g_variant = 0;
int orc_set_variant(int bits){
  g_variant = bits;
  return 0;
}

/* --------------------------------------------------------- linear algebra */

/* src/linalg.c:9-30 */
static int thomas_d(int n, double *l, double *c, double *u, double *q){
  u[0] = u[0]/c[0];
  q[0] = q[0]/c[0];
  for(int i=1; i<=n-1; i++){
    double val = c[i]-l[i]*u[i-1];
    if(fabs(val) < DBL_EPSILON){
      u[i] = 0.;
      q[i] = 0.;
    }else{
      val = 1./val;
      u[i] = val* u[i];
      q[i] = val*(q[i]-l[i]*q[i-1]);
    }
  }
  for(int i=n-2; i>=0; i--) q[i] -= u[i]*q[i+1];
  return 0;
}

/* The system with its rows in reverse order (sub- and super-diagonals
 * exchanged) solved by f, the solution written back in the original order. */
static void reversed_copy(int n, const double *l, const double *c, const double *u,
                          double *lr, double *cr, double *ur){
  for(int i=0; i<n; i++){
    lr[i] = u[n-1-i];
    cr[i] = c[n-1-i];
    ur[i] = l[n-1-i];
  }
}

int orc_dgtsv_b(int n, double *l, double *c, double *u, double *q){
  if(!(g_variant & ORC_VARIANT_REVERSED_SOLVE)) return thomas_d(n, l, c, u, q);
  double *lr = xcalloc(n, sizeof(double)), *cr = xcalloc(n, sizeof(double));
  double *ur = xcalloc(n, sizeof(double)), *qr = xcalloc(n, sizeof(double));
  reversed_copy(n, l, c, u, lr, cr, ur);
  for(int i=0; i<n; i++) qr[i] = q[n-1-i];
  thomas_d(n, lr, cr, ur, qr);
  for(int i=0; i<n; i++) q[i] = qr[n-1-i];
  free(lr); free(cr); free(ur); free(qr);
  return 0;
}

/* src/linalg.c:32-50; real matrix, complex right-hand side.  The
 * complex-by-real products and quotients are component-wise (C99 G.5.1),
 * written out explicitly here. */
static int thomas_z(int n, double *l, double *c, double *u, double complex *q){
  double *qr = xcalloc(n, sizeof(double));
  double *qi = xcalloc(n, sizeof(double));
  for(int i=0; i<n; i++){ qr[i] = creal(q[i]); qi[i] = cimag(q[i]); }
  u[0] = u[0]/c[0];
  qr[0] = qr[0]/c[0];
  qi[0] = qi[0]/c[0];
  for(int i=1; i<=n-1; i++){
    double val = c[i]-l[i]*u[i-1];
    if(fabs(val) < DBL_EPSILON){
      u[i] = 0.;
      qr[i] = 0.;
      qi[i] = 0.;
    }else{
      val = 1./val;
      u[i] = val* u[i];
      qr[i] = val*(qr[i]-l[i]*qr[i-1]);
      qi[i] = val*(qi[i]-l[i]*qi[i-1]);
    }
  }
  for(int i=n-2; i>=0; i--){
    qr[i] -= u[i]*qr[i+1];
    qi[i] -= u[i]*qi[i+1];
  }
  for(int i=0; i<n; i++) q[i] = mkz(qr[i], qi[i]);
  free(qr);
  free(qi);
  return 0;
}

int orc_zgtsv_b(int n, double *l, double *c, double *u, double complex *q){
  if(!(g_variant & ORC_VARIANT_REVERSED_SOLVE)) return thomas_z(n, l, c, u, q);
  double *lr = xcalloc(n, sizeof(double)), *cr = xcalloc(n, sizeof(double));
  double *ur = xcalloc(n, sizeof(double));
  double complex *qr = xcalloc(n, sizeof(double complex));
  reversed_copy(n, l, c, u, lr, cr, ur);
  for(int i=0; i<n; i++) qr[i] = q[n-1-i];
  thomas_z(n, lr, cr, ur, qr);
  for(int i=0; i<n; i++) q[i] = qr[n-1-i];
  free(lr); free(cr); free(ur); free(qr);
  return 0;
}

/* src/linalg.c:52-108: periodic system by superposition of two (n-1) systems */
int orc_dgtsv_p(int n, double *l, double *c, double *u, double *q){
  double *l0 = xcalloc(n-1, sizeof(double)), *c0 = xcalloc(n-1, sizeof(double));
  double *u0 = xcalloc(n-1, sizeof(double)), *q0 = xcalloc(n-1, sizeof(double));
  double *l1 = xcalloc(n-1, sizeof(double)), *c1 = xcalloc(n-1, sizeof(double));
  double *u1 = xcalloc(n-1, sizeof(double)), *q1 = xcalloc(n-1, sizeof(double));
  for(int i=0; i<n-1; i++){
    l0[i] = l[i]; c0[i] = c[i]; u0[i] = u[i]; q0[i] = q[i];
  }
  for(int i=0; i<n-1; i++){
    l1[i] = l[i]; c1[i] = c[i]; u1[i] = u[i];
    q1[i] = i ==   0 ? -l[i]
          : i == n-2 ? -u[i]
          : 0.;
  }
  orc_dgtsv_b(n-1, l0, c0, u0, q0);
  orc_dgtsv_b(n-1, l1, c1, u1, q1);
  q[n-1] = (q[n-1]-u[n-1]*q0[0]-l[n-1]*q0[n-2])
         / (c[n-1]+u[n-1]*q1[0]+l[n-1]*q1[n-2]);
  for(int i=0; i<=n-2; i++) q[i] = q0[i]+q[n-1]*q1[i];
  free(l0); free(c0); free(u0); free(q0);
  free(l1); free(c1); free(u1); free(q1);
  return 0;
}

/* threads the OpenMP loops use (1 without OpenMP) */
int orc_num_threads(void){
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ FFT */

/* Mixed-radix decimation-in-time complex DFT, recursive on the smallest
 * prime factor (radix-2 levels in butterfly form).  tw[m] = exp(-2 pi i m /
 * ntop); a sub-transform of length n uses every (ntop/n)-th entry.  sign =
 * -1 forward, +1 backward.  scr holds 2n entries of scratch. */
static void cdft_rec(int n, const double complex *x, ptrdiff_t xs, double complex *y,
                     const double complex *tw, int ntop, int sign, double complex *scr){
  if(n == 1){ y[0] = x[0]; return; }
  int p = 2;
  while(p*p <= n && n % p) p++;
  if(n % p) p = n;
  const int m = n/p;
  const int step = ntop/n;
  double complex *sub = scr;
  for(int q=0; q<p; q++) cdft_rec(m, x + q*xs, xs*p, sub + (size_t)q*m, tw, ntop, sign, scr + n);
  if(p == 2){
    for(int k=0; k<m; k++){
      double complex w = tw[(size_t)k*step];
      if(sign > 0) w = conj(w);
      const double complex a = sub[k];
      const double complex b = w*sub[m+k];
      y[k]   = a+b;
      y[k+m] = a-b;
    }
    return;
  }
  for(int k=0; k<n; k++){
    double complex acc = 0.;
    for(int q=0; q<p; q++){
      long e = ((long)q*k) % n;
      double complex w = tw[e*step];
      if(sign > 0) w = conj(w);
      acc += w*sub[(size_t)q*m + (k % m)];
    }
    y[k] = acc;
  }
}

static double complex *twiddles(int n){
  double complex *tw = xcalloc((size_t)n, sizeof(double complex));
  for(int m=0; m<n; m++){
    double a = 2.*M_PI*(double)m/(double)n;
    tw[m] = mkz(cos(a), -sin(a));
  }
  return tw;
}


static void cdft_direct(int n, const double complex *x, double complex *y,
                        const double complex *tw, int sign){
  for(int k=0; k<n; k++){
    double complex acc = 0.;
    for(int j=0; j<n; j++){
      double complex w = tw[((long)j*k) % n];
      if(sign > 0) w = conj(w);
      acc += w*x[j];
    }
    y[k] = acc;
  }
}

static void cdft(int n, const double complex *x, double complex *y, const double complex *tw,
                 int sign, double complex *scr){
  if(g_variant & ORC_VARIANT_DIRECT_DFT) cdft_direct(n, x, y, tw, sign);
  else cdft_rec(n, x, 1, y, tw, n, sign, scr);
}

/* r2c with precomputed twiddles; work holds 4n entries */
static void rfft_tw(int n, const double *in, double complex *out, const double complex *tw,
                    double complex *work){
  double complex *x = work, *y = work + n;
  for(int j=0; j<n; j++) x[j] = mkz(in[j], 0.);
  cdft(n, x, y, tw, -1, work + 2*(size_t)n);
  for(int k=0; k<=n/2; k++) out[k] = y[k];
  out[0] = mkz(creal(out[0]), 0.);
  if(n % 2 == 0) out[n/2] = mkz(creal(out[n/2]), 0.);
}

/* c2r with precomputed twiddles; work holds 4n entries */
static void irfft_tw(int n, const double complex *in, double *out, const double complex *tw,
                     double complex *work){
  double complex *x = work, *y = work + n;
  for(int k=0; k<=n/2; k++) x[k] = in[k];
  for(int k=n/2+1; k<n; k++) x[k] = conj(in[n-k]);
  x[0] = mkz(creal(x[0]), 0.);
  if(n % 2 == 0) x[n/2] = mkz(creal(x[n/2]), 0.);
  cdft(n, x, y, tw, +1, work + 2*(size_t)n);
  for(int j=0; j<n; j++) out[j] = creal(y[j]);
}

/* r2c: out[J] = sum_j in[j] exp(-2 pi i jJ/n), J = 0..n/2 */
int orc_rfft(int n, const double *in, double complex *out){
  double complex *tw = twiddles(n);
  double complex *work = xcalloc(4*(size_t)n, sizeof(double complex));
  rfft_tw(n, in, out, tw, work);
  free(work); free(tw);
  return 0;
}

/* c2r: out[j] = sum_{J=0}^{n-1} Y_J exp(+2 pi i jJ/n) with the Hermitian
 * extension Y_{n-J} = conj(Y_J); unnormalised.  Imaginary parts of the DC
 * (and even-n Nyquist) entries do not contribute, as with FFTW's c2r. */
int orc_irfft(int n, const double complex *in, double *out){
  double complex *tw = twiddles(n);
  double complex *work = xcalloc(4*(size_t)n, sizeof(double complex));
  irfft_tw(n, in, out, tw, work);
  free(work); free(tw);
  return 0;
}

/* ------------------------------------------------ decomposition, transpose */

/* src/parallel/others.c:6-10 */
int orc_parallel_get_size(int num, int size, int rank){
  return (num+rank)/size;
}

/* src/parallel/others.c:12-19 */
int orc_parallel_get_offset(int num, int size, int rank){
  int offset = 0;
  for(int r=0; r<rank; r++) offset += orc_parallel_get_size(num, size, r);
  return offset;
}

/* src/parallel/transpose.c:6-85, with MPI_Alltoallw simulated in one
 * process.  Rank r holds sendbufs[r] (x-aligned, g_isize x jsize_r) and
 * receives recvbufs[r] (y-aligned, isize_r x g_jsize).  The send type to
 * peer n walks (ii over n's i-range, jj over the sender's rows) with strides
 * (1, g_isize); the receive type from peer n walks (ii over own i-range,
 * jj over n's rows) contiguously in jj with stride g_jsize in ii. */
int orc_parallel_transpose_sim(int nranks, int g_isize, int g_jsize, size_t elemsize,
                               const void *const *sendbufs, void *const *recvbufs){
  for(int me=0; me<nranks; me++){          /* receiver */
    const int isize_me = orc_parallel_get_size(g_isize, nranks, me);
    const int ioffs_me = orc_parallel_get_offset(g_isize, nranks, me);
    for(int n=0; n<nranks; n++){           /* sender */
      const int jsize_n = orc_parallel_get_size(g_jsize, nranks, n);
      const int joffs_n = orc_parallel_get_offset(g_jsize, nranks, n);
      const char *src = (const char *)sendbufs[n];
      char *dst = (char *)recvbufs[me];
      for(int ii=0; ii<isize_me; ii++){
        for(int jj=0; jj<jsize_n; jj++){
          size_t s_idx = (size_t)ioffs_me + ii + (size_t)jj*g_isize;
          size_t d_idx = (size_t)ii*g_jsize + joffs_n + jj;
          memcpy(dst + d_idx*elemsize, src + s_idx*elemsize, elemsize);
        }
      }
    }
  }
  return 0;
}

/* single-rank transpose: recv[i*g_jsize + j] = send[j*g_isize + i] */
static void transpose_d(int g_isize, int g_jsize, const double *send, double *recv){
#pragma omp parallel for
  for(int i=0; i<g_isize; i++)
    for(int j=0; j<g_jsize; j++)
      recv[(size_t)i*g_jsize+j] = send[(size_t)j*g_isize+i];
}

static void transpose_z(int g_isize, int g_jsize, const double complex *send, double complex *recv){
#pragma omp parallel for
  for(int i=0; i<g_isize; i++)
    for(int j=0; j<g_jsize; j++)
      recv[(size_t)i*g_jsize+j] = send[(size_t)j*g_isize+i];
}

/* ---------------------------------------------------------- temperature */

/* src/temperature/force.c:6-32 */
int orc_temperature_compute_force(const orc_param_t *param, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double *temp = s->temp;
  double *tempforcex = s->tempforcex;
  if(param->with_thermal_forcing){
    for(int j=1; j<=jtot; j++){
      for(int i=2; i<=itot; i++){
        double temp_xm = TEMP(i-1, j);
        double temp_xp = TEMP(i  , j);
        TEMPFORCEX(i, j) = 0.5*temp_xm+0.5*temp_xp;
      }
    }
  }else{
    for(int j=1; j<=jtot; j++)
      for(int i=2; i<=itot; i++)
        TEMPFORCEX(i, j) = 0.;
  }
  return 0;
}

/* src/temperature/update.c:11-87 */
static void temp_compute_src(const orc_param_t *param, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const int implicitx = param->implicitx;
  const int implicity = param->implicity;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double *ux = s->ux;
  const double *uy = s->uy;
  const double *temp = s->temp;
  double *srctempa = s->srctempa;
  double *srctempb = s->srctempb;
  double *srctempg = s->srctempg;
  memcpy(srctempb, srctempa, sizeof(double)*itot*jtot);
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double dtdx_xm = (-TEMP(i-1, j  )+TEMP(i  , j  ))/DXC(i  );
      double dtdx_xp = (-TEMP(i  , j  )+TEMP(i+1, j  ))/DXC(i+1);
      double dtdy_ym = (-TEMP(i  , j-1)+TEMP(i  , j  ))/dy;
      double dtdy_yp = (-TEMP(i  , j  )+TEMP(i  , j+1))/dy;
      double adv1;
      {
        double c_xm = DXC(i  )/(2.*DXF(i));
        double c_xp = DXC(i+1)/(2.*DXF(i));
        double ux_xm = UX(i  , j);
        double ux_xp = UX(i+1, j);
        adv1 = -c_xm*ux_xm*dtdx_xm
               -c_xp*ux_xp*dtdx_xp;
      }
      double adv2;
      {
        double uy_ym = UY(i, j  );
        double uy_yp = UY(i, j+1);
        adv2 = -0.5*uy_ym*dtdy_ym
               -0.5*uy_yp*dtdy_yp;
      }
      double dif1 = 1./sqrt(Pr)/sqrt(Ra)*(dtdx_xp-dtdx_xm)/DXF(i);
      double dif2 = 1./sqrt(Pr)/sqrt(Ra)*(dtdy_yp-dtdy_ym)/dy;
      SRCTEMPA(i, j) =
        +(adv1+adv2)
        +(1.-1.*implicitx)*dif1
        +(1.-1.*implicity)*dif2;
      SRCTEMPG(i, j) =
        +(  +1.*implicitx)*dif1
        +(  +1.*implicity)*dif2;
    }
  }
}

/* y-direction periodic Crank-Nicolson solve on an x-aligned increment
 * qx (ni x jtot, row pitch ni), shared by T, ux, uy
 * (src/temperature/update.c:168-220, src/fluid/update_velocity.c:281-333,425-477) */
static void solve_y_periodic_local(int ni, int jtot, double prefactor, double dy, double *qx);
static void solve_y_periodic(int ni, int jtot, double prefactor, double dy, double *qx){
  if(g_solve_y) g_solve_y(ni, jtot, prefactor, dy, qx);
  else solve_y_periodic_local(ni, jtot, prefactor, dy, qx);
}
static void solve_y_periodic_local(int ni, int jtot, double prefactor, double dy, double *qx){
  double *qy = xcalloc((size_t)ni*jtot, sizeof(double));
  transpose_d(ni, jtot, qx, qy);
#pragma omp parallel
  {
  double *l = xcalloc(jtot, sizeof(double)), *c = xcalloc(jtot, sizeof(double));
  double *u = xcalloc(jtot, sizeof(double)), *q = xcalloc(jtot, sizeof(double));
#pragma omp for
  for(int i=0; i<ni; i++){
    for(int j=0; j<jtot; j++){
      double coef_ym = 1./dy/dy;
      double coef_yp = 1./dy/dy;
      double coef_y0 = -coef_ym-coef_yp;
      l[j] =   -prefactor*coef_ym;
      u[j] =   -prefactor*coef_yp;
      c[j] = 1.-prefactor*coef_y0;
    }
    for(int j=0; j<jtot; j++) q[j] = qy[(size_t)i*jtot+j];
    orc_dgtsv_p(jtot, l, c, u, q);
    for(int j=0; j<jtot; j++) qy[(size_t)i*jtot+j] = q[j];
  }
  free(l); free(c); free(u); free(q);
  }
  transpose_d(jtot, ni, qy, qx);
  free(qy);
}

/* src/temperature/update.c:89-231 */
static void temp_update(const orc_param_t *param, int rkstep, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double alpha = param->alpha[rkstep];
  const double beta  = param->beta [rkstep];
  const double gamma = param->gamma[rkstep];
  const double dt = param->dt;
  const double *srctempa = s->srctempa;
  const double *srctempb = s->srctempb;
  const double *srctempg = s->srctempg;
  double *temp = s->temp;
  double *qx = xcalloc((size_t)itot*jtot, sizeof(double));
#define QX(I, J) (qx[((J)-1)*(itot)+((I)-1)])
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      QX(i, j) =
        +alpha*dt*SRCTEMPA(i, j)
        +beta *dt*SRCTEMPB(i, j)
        +gamma*dt*SRCTEMPG(i, j);
    }
  }
  if(param->implicitx){
#pragma omp parallel
    {
    double *l = xcalloc(itot, sizeof(double)), *c = xcalloc(itot, sizeof(double));
    double *u = xcalloc(itot, sizeof(double)), *q = xcalloc(itot, sizeof(double));
#pragma omp for
    for(int j=1; j<=jtot; j++){
      for(int i=1; i<=itot; i++){
        double prefactor = (gamma*dt)/(2.*sqrt(Pr)*sqrt(Ra));
        double coef_xm = 1./DXF(i)/DXC(i  );
        double coef_xp = 1./DXF(i)/DXC(i+1);
        double coef_x0 = -coef_xm-coef_xp;
        l[i-1] =   -prefactor*coef_xm;
        u[i-1] =   -prefactor*coef_xp;
        c[i-1] = 1.-prefactor*coef_x0;
      }
      for(int i=1; i<=itot; i++) q[i-1] = QX(i, j);
      orc_dgtsv_b(itot, l, c, u, q);
      for(int i=1; i<=itot; i++) QX(i, j) = q[i-1];
    }
    free(l); free(c); free(u); free(q);
    }
  }
  if(param->implicity){
    double prefactor = (gamma*dt)/(2.*sqrt(Pr)*sqrt(Ra));
    solve_y_periodic(itot, jtot, prefactor, dy, qx);
  }
#pragma omp parallel for
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      TEMP(i, j) += QX(i, j);
#undef QX
  free(qx);
  orc_temperature_update_boundaries_temp(param, temp);
}

/* src/temperature/update.c:233-239 */
int orc_temperature_update_temp(const orc_param_t *param, int rkstep, orc_state_t *s){
  temp_compute_src(param, s);
  temp_update(param, rkstep, s);
  return 0;
}

/* ---------------------------------------------------------------- fluid */

/* src/fluid/update_velocity.c:11-104 */
static void compute_src_ux(const orc_param_t *param, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const int implicitx = param->implicitx;
  const int implicity = param->implicity;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double *ux = s->ux;
  const double *uy = s->uy;
  const double *p = s->p;
  const double *tempforcex = s->tempforcex;
  double *srcuxa = s->srcuxa;
  double *srcuxb = s->srcuxb;
  double *srcuxg = s->srcuxg;
  memcpy(srcuxb, srcuxa, sizeof(double)*(itot-1)*jtot);
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=2; i<=itot; i++){
      double duxdx_xm = (-UX(i-1, j  )+UX(i  , j  ))/DXF(i-1);
      double duxdx_xp = (-UX(i  , j  )+UX(i+1, j  ))/DXF(i  );
      double duxdy_ym = (-UX(i  , j-1)+UX(i  , j  ))/dy;
      double duxdy_yp = (-UX(i  , j  )+UX(i  , j+1))/dy;
      double adv1;
      {
        double c_xm = DXF(i-1)/(2.*DXC(i));
        double c_xp = DXF(i  )/(2.*DXC(i));
        double ux_xm = 0.5*UX(i-1, j)+0.5*UX(i  , j);
        double ux_xp = 0.5*UX(i  , j)+0.5*UX(i+1, j);
        adv1 = -c_xm*ux_xm*duxdx_xm
               -c_xp*ux_xp*duxdx_xp;
      }
      double adv2;
      {
        double c_xm = DXF(i-1)/(2.*DXC(i));
        double c_xp = DXF(i  )/(2.*DXC(i));
        double uy_ym = c_xm*UY(i-1, j  )+c_xp*UY(i  , j  );
        double uy_yp = c_xm*UY(i-1, j+1)+c_xp*UY(i  , j+1);
        adv2 = -0.5*uy_ym*duxdy_ym
               -0.5*uy_yp*duxdy_yp;
      }
      double dif1 = sqrt(Pr)/sqrt(Ra)*(duxdx_xp-duxdx_xm)/DXC(i);
      double dif2 = sqrt(Pr)/sqrt(Ra)*(duxdy_yp-duxdy_ym)/dy;
      double pre;
      {
        double p_xm = P(i-1, j);
        double p_xp = P(i  , j);
        pre = -(p_xp-p_xm)/DXC(i);
      }
      double tmp = TEMPFORCEX(i, j);
      SRCUXA(i, j) =
        +(adv1+adv2)
        +(1.-1.*implicitx)*dif1
        +(1.-1.*implicity)*dif2
        +tmp;
      SRCUXG(i, j) =
        +(  +1.*implicitx)*dif1
        +(  +1.*implicity)*dif2
        +pre;
    }
  }
}

/* src/fluid/update_velocity.c:106-192 */
static void compute_src_uy(const orc_param_t *param, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const int implicitx = param->implicitx;
  const int implicity = param->implicity;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double *ux = s->ux;
  const double *uy = s->uy;
  const double *p = s->p;
  double *srcuya = s->srcuya;
  double *srcuyb = s->srcuyb;
  double *srcuyg = s->srcuyg;
  memcpy(srcuyb, srcuya, sizeof(double)*itot*jtot);
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double duydx_xm = (-UY(i-1, j  )+UY(i  , j  ))/DXC(i  );
      double duydx_xp = (-UY(i  , j  )+UY(i+1, j  ))/DXC(i+1);
      double duydy_ym = (-UY(i  , j-1)+UY(i  , j  ))/dy;
      double duydy_yp = (-UY(i  , j  )+UY(i  , j+1))/dy;
      double adv1;
      {
        double c_xm = DXC(i  )/(2.*DXF(i));
        double c_xp = DXC(i+1)/(2.*DXF(i));
        double ux_xm = 0.5*UX(i  , j-1)+0.5*UX(i  , j);
        double ux_xp = 0.5*UX(i+1, j-1)+0.5*UX(i+1, j);
        adv1 = -c_xm*ux_xm*duydx_xm
               -c_xp*ux_xp*duydx_xp;
      }
      double adv2;
      {
        double uy_ym = 0.5*UY(i, j-1)+0.5*UY(i, j  );
        double uy_yp = 0.5*UY(i, j  )+0.5*UY(i, j+1);
        adv2 = -0.5*uy_ym*duydy_ym
               -0.5*uy_yp*duydy_yp;
      }
      double dif1 = sqrt(Pr)/sqrt(Ra)*(duydx_xp-duydx_xm)/DXF(i);
      double dif2 = sqrt(Pr)/sqrt(Ra)*(duydy_yp-duydy_ym)/dy;
      double pre;
      {
        double p_ym = P(i, j-1);
        double p_yp = P(i, j  );
        pre = -(p_yp-p_ym)/dy;
      }
      SRCUYA(i, j) =
        +(adv1+adv2)
        +(1.-1.*implicitx)*dif1
        +(1.-1.*implicity)*dif2;
      SRCUYG(i, j) =
        +(  +1.*implicitx)*dif1
        +(  +1.*implicity)*dif2
        +pre;
    }
  }
}

/* src/fluid/update_velocity.c:194-344 */
static void update_ux(const orc_param_t *param, int rkstep, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double alpha = param->alpha[rkstep];
  const double beta  = param->beta [rkstep];
  const double gamma = param->gamma[rkstep];
  const double dt = param->dt;
  const double *srcuxa = s->srcuxa;
  const double *srcuxb = s->srcuxb;
  const double *srcuxg = s->srcuxg;
  double *ux = s->ux;
  double *qx = xcalloc((size_t)(itot+1)*jtot, sizeof(double));
#define QX(I, J) (qx[((J)-1)*(itot+1)+((I)-1)])
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot+1; i++){
      QX(i, j) =
          i ==      1 ? 0.
        : i == itot+1 ? 0.
        : +alpha*dt*SRCUXA(i, j)
          +beta *dt*SRCUXB(i, j)
          +gamma*dt*SRCUXG(i, j);
    }
  }
  if(param->implicitx){
    const int n = itot+1;
#pragma omp parallel
    {
    double *l = xcalloc(n, sizeof(double)), *c = xcalloc(n, sizeof(double));
    double *u = xcalloc(n, sizeof(double)), *q = xcalloc(n, sizeof(double));
#pragma omp for
    for(int j=1; j<=jtot; j++){
      for(int i=1; i<=itot+1; i++){
        if(i == 1 || i == itot+1){
          l[i-1] = 0.;
          u[i-1] = 0.;
          c[i-1] = 1.;
        }else{
          double prefactor = (gamma*dt*sqrt(Pr))/(2.*sqrt(Ra));
          double coef_xm = 1./DXC(i  )/DXF(i-1);
          double coef_xp = 1./DXC(i  )/DXF(i  );
          double coef_x0 = -coef_xm-coef_xp;
          l[i-1] =   -prefactor*coef_xm;
          u[i-1] =   -prefactor*coef_xp;
          c[i-1] = 1.-prefactor*coef_x0;
        }
      }
      for(int i=1; i<=itot+1; i++) q[i-1] = QX(i, j);
      orc_dgtsv_b(n, l, c, u, q);
      for(int i=1; i<=itot+1; i++) QX(i, j) = q[i-1];
    }
    free(l); free(c); free(u); free(q);
    }
  }
  if(param->implicity){
    double prefactor = (gamma*dt*sqrt(Pr))/(2.*sqrt(Ra));
    solve_y_periodic(itot+1, jtot, prefactor, dy, qx);
  }
#pragma omp parallel for
  for(int j=1; j<=jtot; j++)
    for(int i=2; i<=itot; i++)
      UX(i, j) += QX(i, j);
#undef QX
  free(qx);
  orc_fluid_update_boundaries_ux(param, ux);
}

/* src/fluid/update_velocity.c:346-488 */
static void update_uy(const orc_param_t *param, int rkstep, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double alpha = param->alpha[rkstep];
  const double beta  = param->beta [rkstep];
  const double gamma = param->gamma[rkstep];
  const double dt = param->dt;
  const double *srcuya = s->srcuya;
  const double *srcuyb = s->srcuyb;
  const double *srcuyg = s->srcuyg;
  double *uy = s->uy;
  double *qx = xcalloc((size_t)itot*jtot, sizeof(double));
#define QX(I, J) (qx[((J)-1)*(itot)+((I)-1)])
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      QX(i, j) =
        +alpha*dt*SRCUYA(i, j)
        +beta *dt*SRCUYB(i, j)
        +gamma*dt*SRCUYG(i, j);
    }
  }
  if(param->implicitx){
#pragma omp parallel
    {
    double *l = xcalloc(itot, sizeof(double)), *c = xcalloc(itot, sizeof(double));
    double *u = xcalloc(itot, sizeof(double)), *q = xcalloc(itot, sizeof(double));
#pragma omp for
    for(int j=1; j<=jtot; j++){
      for(int i=1; i<=itot; i++){
        double prefactor = (gamma*dt*sqrt(Pr))/(2.*sqrt(Ra));
        double coef_xm = 1./DXF(i)/DXC(i  );
        double coef_xp = 1./DXF(i)/DXC(i+1);
        double coef_x0 = -coef_xm-coef_xp;
        l[i-1] =   -prefactor*coef_xm;
        u[i-1] =   -prefactor*coef_xp;
        c[i-1] = 1.-prefactor*coef_x0;
      }
      for(int i=1; i<=itot; i++) q[i-1] = QX(i, j);
      orc_dgtsv_b(itot, l, c, u, q);
      for(int i=1; i<=itot; i++) QX(i, j) = q[i-1];
    }
    free(l); free(c); free(u); free(q);
    }
  }
  if(param->implicity){
    double prefactor = (gamma*dt*sqrt(Pr))/(2.*sqrt(Ra));
    solve_y_periodic(itot, jtot, prefactor, dy, qx);
  }
#pragma omp parallel for
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      UY(i, j) += QX(i, j);
#undef QX
  free(qx);
  orc_fluid_update_boundaries_uy(param, uy);
}

/* src/fluid/update_velocity.c:490-498 */
int orc_fluid_update_velocity(const orc_param_t *param, int rkstep, orc_state_t *s){
  compute_src_ux(param, s);
  compute_src_uy(param, s);
  update_ux(param, rkstep, s);
  update_uy(param, rkstep, s);
  return 0;
}

/* src/fluid/compute_potential.c:15-168 (single rank) */
int orc_fluid_compute_potential(const orc_param_t *param, int rkstep, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const int nwave = jtot/2+1;
  const double dy = param->dy;
  double *psi = s->psi;
  double *qrx = xcalloc((size_t)itot*jtot, sizeof(double));
  double *qry = xcalloc((size_t)itot*jtot, sizeof(double));
  double complex *qcx = xcalloc((size_t)itot*nwave, sizeof(double complex));
  double complex *qcy = xcalloc((size_t)itot*nwave, sizeof(double complex));
#define QRX(I, J) (qrx[((J)-1)*(itot)+((I)-1)])
#define QCX(I, J) (qcx[((J)-1)*(itot)+((I)-1)])
  const double gamma = param->gamma[rkstep];
  const double dt = param->dt;
  const double *ux = s->ux;
  const double *uy = s->uy;
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double ux_xm = UX(i  , j  );
      double ux_xp = UX(i+1, j  );
      double uy_ym = UY(i  , j  );
      double uy_yp = UY(i  , j+1);
      QRX(i, j) =
        1./(gamma*dt)*(
         +(ux_xp-ux_xm)/DXF(i)
         +(uy_yp-uy_ym)/dy
        );
    }
  }
  transpose_d(itot, jtot, qrx, qry);
  double complex *tw = twiddles(jtot);
#pragma omp parallel
  {
    double complex *work = xcalloc(4*(size_t)jtot, sizeof(double complex));
#pragma omp for
    for(int i=0; i<itot; i++) rfft_tw(jtot, qry + (size_t)i*jtot, qcy + (size_t)i*nwave, tw, work);
    free(work);
  }
  transpose_z(nwave, itot, qcy, qcx);
#pragma omp parallel
  {
    double *l = xcalloc(itot, sizeof(double)), *c = xcalloc(itot, sizeof(double));
    double *u = xcalloc(itot, sizeof(double));
    double complex *q = xcalloc(itot, sizeof(double complex));
#pragma omp for
    for(int j=1; j<=nwave; j++){
      double eigenvalue = -4./pow(dy, 2.)*pow(sin(M_PI*(j-1)/(1.*jtot)), 2.);
      for(int i=1; i<=itot; i++){
        l[i-1] = 1./DXC(i  )/DXF(i);
        u[i-1] = 1./DXC(i+1)/DXF(i);
      }
      for(int i=1; i<=itot; i++) c[i-1] = -l[i-1]-u[i-1]+eigenvalue;
      c[     0] += l[     0];
      c[itot-1] += u[itot-1];
      for(int i=1; i<=itot; i++) q[i-1] = QCX(i, j);
      if(j == 1 && (g_variant & ORC_VARIANT_PIN_GAUGE)){
        /* the singular J = 0 system with its last equation replaced by
         * psi_hat(itot) = 0 */
        l[itot-1] = 0.;
        c[itot-1] = 1.;
        q[itot-1] = 0.;
      }
      orc_zgtsv_b(itot, l, c, u, q);
      for(int i=1; i<=itot; i++) QCX(i, j) = q[i-1];
    }
    free(l); free(c); free(u); free(q);
  }
  transpose_z(itot, nwave, qcx, qcy);
#pragma omp parallel
  {
    double complex *work = xcalloc(4*(size_t)jtot, sizeof(double complex));
#pragma omp for
    for(int i=0; i<itot; i++) irfft_tw(jtot, qcy + (size_t)i*nwave, qry + (size_t)i*jtot, tw, work);
    free(work);
  }
  free(tw);
  transpose_d(jtot, itot, qry, qrx);
#pragma omp parallel for
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      PSI(i, j) = QRX(i, j)/(1.*jtot);
  orc_fluid_update_boundaries_p(param, psi);
#undef QRX
#undef QCX
  free(qrx); free(qry); free(qcx); free(qcy);
  return 0;
}

/* src/fluid/correct_velocity.c:6-42 */
int orc_fluid_correct_velocity(const orc_param_t *param, int rkstep, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double gamma = param->gamma[rkstep];
  const double dt = param->dt;
  const double *psi = s->psi;
  double *ux = s->ux;
  double *uy = s->uy;
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=2; i<=itot; i++){
      double dxi = 1./DXC(i);
      double psi_xp = PSI(i  , j);
      double psi_xm = PSI(i-1, j);
      UX(i, j) -= gamma*dt*(psi_xp-psi_xm)*dxi;
    }
  }
  orc_fluid_update_boundaries_ux(param, ux);
#pragma omp parallel for
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double psi_yp = PSI(i, j  );
      double psi_ym = PSI(i, j-1);
      UY(i, j) -= gamma*dt*(psi_yp-psi_ym)/dy;
    }
  }
  orc_fluid_update_boundaries_uy(param, uy);
  return 0;
}

/* src/fluid/update_pressure.c:7-60 */
int orc_fluid_update_pressure(const orc_param_t *param, int rkstep, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double Re = sqrt(Ra)/sqrt(Pr);
  const double gamma = param->gamma[rkstep];
  const double dt = param->dt;
  const double *psi = s->psi;
  double *p = s->p;
#pragma omp parallel for
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      P(i, j) += PSI(i, j);
  if(param->implicitx){
    for(int j=1; j<=jtot; j++){
      for(int i=1; i<=itot; i++){
        double dpsidx_xm = (-PSI(i-1, j)+PSI(i  , j))/DXC(i  );
        double dpsidx_xp = (-PSI(i  , j)+PSI(i+1, j))/DXC(i+1);
        P(i, j) += -(gamma*dt)/(2.*Re)*(+(dpsidx_xp-dpsidx_xm)/DXF(i));
      }
    }
  }
  if(param->implicity){
    for(int j=1; j<=jtot; j++){
      for(int i=1; i<=itot; i++){
        double dpsidy_ym = (-PSI(i, j-1)+PSI(i, j  ))/dy;
        double dpsidy_yp = (-PSI(i, j  )+PSI(i, j+1))/dy;
        P(i, j) += -(gamma*dt)/(2.*Re)*(+(dpsidy_yp-dpsidy_ym)/dy);
      }
    }
  }
  orc_fluid_update_boundaries_p(param, p);
  return 0;
}

/* --------------------------------------------------------------- driver */

/* src/main.c:79-101 */
int orc_integrate(const orc_param_t *param, orc_state_t *s){
  for(int rkstep=0; rkstep<3; rkstep++){
    orc_fluid_update_boundaries_ux(param, s->ux);
    orc_fluid_update_boundaries_uy(param, s->uy);
    orc_fluid_update_boundaries_p(param, s->p);
    if(param->with_temperature){
      orc_temperature_update_boundaries_temp(param, s->temp);
      orc_temperature_compute_force(param, s);
      orc_temperature_update_temp(param, rkstep, s);
    }
    orc_fluid_update_velocity(param, rkstep, s);
    orc_fluid_compute_potential(param, rkstep, s);
    orc_fluid_correct_velocity(param, rkstep, s);
    orc_fluid_update_pressure(param, rkstep, s);
  }
  return 0;
}

/* src/main.c:28-35 without the schedulers */
int orc_run(orc_param_t *param, orc_state_t *s, int nsteps){
  for(int n=0; n<nsteps; n++){
    orc_param_decide_dt(param, s);
    orc_integrate(param, s);
    param->step += 1;
    param->time += param->dt;
  }
  return 0;
}

/* ---------------------------------------------------------- diagnostics */

/* src/logging.c:36-80 */
int orc_check_divergence(const orc_param_t *param, const orc_state_t *s, double out[2]){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double *ux = s->ux;
  const double *uy = s->uy;
  double divmax = 0.;
  double divsum = 0.;
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double ux_xm = UX(i  , j  );
      double ux_xp = UX(i+1, j  );
      double uy_ym = UY(i  , j  );
      double uy_yp = UY(i  , j+1);
      double div = (ux_xp-ux_xm)/DXF(i)
                  +(uy_yp-uy_ym)/dy;
      divmax = fmax(divmax, fabs(div));
      divsum += div;
    }
  }
  out[0] = divmax;
  out[1] = divsum;
  return 0;
}

/* src/logging.c:82-122 */
int orc_check_momentum(const orc_param_t *param, const orc_state_t *s, double out[2]){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double *ux = s->ux;
  const double *uy = s->uy;
  double moms[2] = {0., 0.};
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot+1; i++)
      moms[0] += UX(i, j)*(DXC(i)*dy);
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      moms[1] += UY(i, j)*(DXF(i)*dy);
  out[0] = moms[0];
  out[1] = moms[1];
  return 0;
}

/* src/logging.c:124-173 */
int orc_check_energy(const orc_param_t *param, const orc_state_t *s, double out[3]){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double dy = param->dy;
  const double *ux = s->ux;
  const double *uy = s->uy;
  const double *temp = s->temp;
  double q[3] = {0., 0., 0.};
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot+1; i++)
      q[0] += 0.5*pow(UX(i, j), 2.)*(DXC(i)*dy);
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      q[1] += 0.5*pow(UY(i, j), 2.)*(DXF(i)*dy);
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      q[2] += 0.5*pow(TEMP(i, j), 2.)*(DXF(i)*dy);
  out[0] = q[0]; out[1] = q[1]; out[2] = q[2];
  return 0;
}

/* src/analyses.c:8-45 */
static double kinetic_dissipation_x(const orc_param_t *param, const double *ux, int i, int j){
  const int itot = param->itot;
  const double dy = param->dy;
  double sxjsxj = 0.;
  if(i != 1){
    double dux = UX(i, j)-UX(i-1, j);
    double duxdx = dux/DXF(i-1);
    sxjsxj += 0.5/DXC(i)*dux*duxdx;
  }
  if(i != itot+1){
    double dux = UX(i+1, j)-UX(i, j);
    double duxdx = dux/DXF(i);
    sxjsxj += 0.5/DXC(i)*dux*duxdx;
  }
  {
    double dux = UX(i, j)-UX(i, j-1);
    double duxdy = dux/dy;
    sxjsxj += 0.5/dy*dux*duxdy;
  }
  {
    double dux = UX(i, j+1)-UX(i, j);
    double duxdy = dux/dy;
    sxjsxj += 0.5/dy*dux*duxdy;
  }
  return sqrt(param->Pr)/sqrt(param->Ra)*sxjsxj;
}

/* src/analyses.c:47-83 */
static double kinetic_dissipation_y(const orc_param_t *param, const double *uy, int i, int j){
  const int itot = param->itot;
  const double dy = param->dy;
  double syjsyj = 0.;
  {
    double c = i == 1 ? 1. : 0.5;
    double duy = UY(i, j)-UY(i-1, j);
    double duydx = duy/DXC(i);
    syjsyj += c/DXF(i)*duy*duydx;
  }
  {
    double c = i == itot ? 1. : 0.5;
    double duy = UY(i+1, j)-UY(i, j);
    double duydx = duy/DXC(i+1);
    syjsyj += c/DXF(i)*duy*duydx;
  }
  {
    double duy = UY(i, j)-UY(i, j-1);
    double duydy = duy/dy;
    syjsyj += 0.5/dy*duy*duydy;
  }
  {
    double duy = UY(i, j+1)-UY(i, j);
    double duydy = duy/dy;
    syjsyj += 0.5/dy*duy*duydy;
  }
  return sqrt(param->Pr)/sqrt(param->Ra)*syjsyj;
}

/* src/analyses.c:85-121 */
static double thermal_dissipation(const orc_param_t *param, const double *temp, int i, int j){
  const int itot = param->itot;
  const double dy = param->dy;
  double riri = 0.;
  {
    double c = i == 1 ? 1. : 0.5;
    double dtemp = TEMP(i, j)-TEMP(i-1, j);
    double dtempdx = dtemp/DXC(i);
    riri += c/DXF(i)*dtemp*dtempdx;
  }
  {
    double c = i == itot ? 1. : 0.5;
    double dtemp = TEMP(i+1, j)-TEMP(i, j);
    double dtempdx = dtemp/DXC(i+1);
    riri += c/DXF(i)*dtemp*dtempdx;
  }
  {
    double dtemp = TEMP(i, j)-TEMP(i, j-1);
    double dtempdy = dtemp/dy;
    riri += 0.5/dy*dtemp*dtempdy;
  }
  {
    double dtemp = TEMP(i, j+1)-TEMP(i, j);
    double dtempdy = dtemp/dy;
    riri += 0.5/dy*dtemp*dtempdy;
  }
  return 1./sqrt(param->Pr)/sqrt(param->Ra)*riri;
}

/* src/logging.c:175-272: wall flux, energy injection, kinetic and thermal
 * dissipation Nusselt numbers */
int orc_check_nusselt(const orc_param_t *param, const orc_state_t *s, double out[4]){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double lx = param->lx;
  const double ly = param->ly;
  const double dy = param->dy;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double *ux = s->ux;
  const double *uy = s->uy;
  const double *temp = s->temp;
  double nu_wall = 0.;
  for(int j=1; j<=jtot; j++){
    double dtdx0 = (TEMP(     1, j)-TEMP(   0, j))/DXC(     1);
    double dtdx1 = (TEMP(itot+1, j)-TEMP(itot, j))/DXC(itot+1);
    nu_wall -= 0.5*(dtdx0+dtdx1)*dy;
  }
  nu_wall /= ly;
  double nu_inje = 0.;
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot+1; i++){
      double integrand = UX(i, j)*0.5*(TEMP(i-1, j)+TEMP(i, j));
      double cellsize = DXC(i)*dy;
      nu_inje += sqrt(Pr)*sqrt(Ra)*integrand*cellsize;
    }
  }
  nu_inje /= lx*ly;
  nu_inje += 1.;
  double nu_eps_k = 0.;
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot+1; i++){
      double cellsize = DXC(i)*dy;
      nu_eps_k += sqrt(Pr)*sqrt(Ra)*kinetic_dissipation_x(param, ux, i, j)*cellsize;
    }
  }
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double cellsize = DXF(i)*dy;
      nu_eps_k += sqrt(Pr)*sqrt(Ra)*kinetic_dissipation_y(param, uy, i, j)*cellsize;
    }
  }
  nu_eps_k /= lx*ly;
  nu_eps_k += 1.;
  double nu_eps_h = 0.;
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double cellsize = DXF(i)*dy;
      nu_eps_h += sqrt(Pr)*sqrt(Ra)*thermal_dissipation(param, temp, i, j)*cellsize;
    }
  }
  nu_eps_h /= lx*ly;
  out[0] = nu_wall; out[1] = nu_inje; out[2] = nu_eps_k; out[3] = nu_eps_h;
  return 0;
}

/* ------------------------------------------------------ initial condition */

uint64_t orc_splitmix64(uint64_t x){
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30))*0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27))*0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

/* T(i,j) = base(i) + amplitude*(2U-1), U in [0,1) from the top 53 bits of
 * splitmix64(seed<<40 ^ ((j-1)*itot + (i-1))); base is the reference's linear
 * conduction profile (src/temperature/init.c:39-47) when linear != 0, else 0 */
int orc_init_temperature(const orc_param_t *param, uint64_t seed, double amplitude, int linear, double *temp){
  const int itot = param->itot;
  const int jtot = param->jtot;
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      uint64_t key = (seed << 40) ^ ((uint64_t)(j-1)*(uint64_t)itot + (uint64_t)(i-1));
      double u = (double)(orc_splitmix64(key) >> 11)*0x1.0p-53;
      double base = linear ? (TEMP_XP-TEMP_XM)*XC(i)+TEMP_XM : 0.;
      TEMP(i, j) = base+amplitude*(2.*u-1.);
    }
  }
  orc_temperature_update_boundaries_temp(param, temp);
  return 0;
}
