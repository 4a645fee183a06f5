/*
 * sns_oracle_mpi.c -- y-slab MPI restatement of the reference's time step.
 *
 * TEST INFRASTRUCTURE AND CPU BASELINE ONLY: a standalone program, never
 * linked into the product.  It runs the oracle's arithmetic (this file
 * includes sns_oracle.c) with the reference's parallel decomposition: each
 * rank owns the rows parallel_get_size/offset give it (src/parallel/
 * others.c:6-19), halo rows come from the neighbouring ranks
 * (src/parallel/halo.c:6-22), the advective bound is reduced over the ranks
 * (src/param/decide_dt.c:28-52) and every y-direction operation -- the
 * Crank-Nicolson y solves (src/temperature/update.c:168-220,
 * src/fluid/update_velocity.c:281-333,425-477) and the Poisson transforms
 * and x solves (src/fluid/compute_potential.c:15-168) -- goes through global
 * transposes with the layout of src/parallel/transpose.c:6-85
 * (recv[i*g_jsize+j] = send[j*g_isize+i]).  The transposes here pack blocks
 * for MPI_Alltoallv instead of building derived datatypes for
 * MPI_Alltoallw.  One thread per rank, as in the reference.
 *
 * Every operation is the single-rank oracle's in the same order, and the
 * only reduction is a minimum, so the fields equal the single-rank
 * oracle's bit for bit for any number of ranks.
 *
 * Usage: mpiexec -n R sns_oracle_mpi itot jtot Ra Pr implicitx implicity seed amplitude
 *                       nsteps [stretch] [pin_gauge] [out_file]
 * Rank 0 prints one JSON line (ranks, steps, seconds of the step loop by
 * MPI_Wtime, cell-steps/s, dt, time); with out_file it writes the interior
 * rows of ux, uy, temp, p (float64, in that order, reference layouts).
 */
#include <mpi.h>
#include <stdio.h>
#include "sns_oracle.c"

static int g_rank, g_size;
static int g_ym, g_yp;

static int slab_size(int num, int r) { return orc_parallel_get_size(num, g_size, r); }
static int slab_off(int num, int r) { return orc_parallel_get_offset(num, g_size, r); }

/* src/parallel/halo.c:6-22: row 0 <- row jsize of the rank below, row
 * jsize+1 <- row 1 of the rank above (periodic in y) */
static void mpi_halo_rows(double *base, size_t pitch, int first_col, int cnt, int jsize){
  MPI_Sendrecv(base + (size_t)jsize*pitch + first_col, cnt, MPI_DOUBLE, g_yp, 0,
               base + first_col, cnt, MPI_DOUBLE, g_ym, 0, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
  MPI_Sendrecv(base + 1*pitch + first_col, cnt, MPI_DOUBLE, g_ym, 1,
               base + (size_t)(jsize+1)*pitch + first_col, cnt, MPI_DOUBLE, g_yp, 1,
               MPI_COMM_WORLD, MPI_STATUS_IGNORE);
}

static void mpi_min(double *v){
  MPI_Allreduce(MPI_IN_PLACE, v, 1, MPI_DOUBLE, MPI_MIN, MPI_COMM_WORLD);
}

/* Global transpose of an nrow x ncol array of es-byte elements whose rows
 * are split over the ranks (this rank: rows slab_off(nrow) .. +slab_size,
 * all ncol columns, row-major) into its columns split over the ranks (this
 * rank: columns slab_off(ncol) .. +slab_size, all nrow rows, one column
 * after the other): recv[c*nrow + r] = send[r*ncol + c] in global indices. */
static void mpi_transpose(int nrow, int ncol, size_t es, const void *send, void *recv){
  const int myr = slab_size(nrow, g_rank), myc = slab_size(ncol, g_rank);
  const int roff = slab_off(nrow, g_rank);
  (void)roff;
  int *scnt = xcalloc(g_size, sizeof(int)), *sdsp = xcalloc(g_size, sizeof(int));
  int *rcnt = xcalloc(g_size, sizeof(int)), *rdsp = xcalloc(g_size, sizeof(int));
  size_t stot = 0, rtot = 0;
  for(int r=0; r<g_size; r++){
    scnt[r] = (int)(es*(size_t)myr*(size_t)slab_size(ncol, r));
    rcnt[r] = (int)(es*(size_t)slab_size(nrow, r)*(size_t)myc);
    sdsp[r] = (int)stot; rdsp[r] = (int)rtot;
    stot += (size_t)scnt[r]; rtot += (size_t)rcnt[r];
  }
  char *sb = xcalloc(stot, 1), *rb = xcalloc(rtot, 1);
  const char *s = send;
  char *d = recv;
  /* block for rank r: its columns, this rank's rows, column after column */
  for(int r=0; r<g_size; r++){
    const int c0 = slab_off(ncol, r), nc = slab_size(ncol, r);
    char *o = sb + sdsp[r];
    for(int c=0; c<nc; c++)
      for(int i=0; i<myr; i++)
        memcpy(o + es*((size_t)c*myr + i), s + es*((size_t)i*ncol + c0 + c), es);
  }
  MPI_Alltoallv(sb, scnt, sdsp, MPI_BYTE, rb, rcnt, rdsp, MPI_BYTE, MPI_COMM_WORLD);
  for(int r=0; r<g_size; r++){
    const int r0 = slab_off(nrow, r), nr = slab_size(nrow, r);
    const char *o = rb + rdsp[r];
    for(int c=0; c<myc; c++)
      for(int i=0; i<nr; i++)
        memcpy(d + es*((size_t)c*nrow + r0 + i), o + es*((size_t)c*nr + i), es);
  }
  free(sb); free(rb); free(scnt); free(sdsp); free(rcnt); free(rdsp);
}

static int g_jtot;  /* global rows */

/* the periodic y solve of an x-aligned increment qx (ni x jsize local rows),
 * as solve_y_periodic_local but on whole columns gathered by transposes */
static void mpi_solve_y(int ni, int jsize, double prefactor, double dy, double *qx){
  (void)jsize;
  const int jtot = g_jtot;
  const int nc = slab_size(ni, g_rank);
  double *qy = xcalloc((size_t)nc*jtot, sizeof(double));
  mpi_transpose(jtot, ni, sizeof(double), qx, qy);
  double *l = xcalloc(jtot, sizeof(double)), *c = xcalloc(jtot, sizeof(double));
  double *u = xcalloc(jtot, sizeof(double)), *q = xcalloc(jtot, sizeof(double));
  for(int i=0; i<nc; i++){
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
  mpi_transpose(ni, jtot, sizeof(double), qy, qx);
  free(qy);
}

/* src/fluid/compute_potential.c:15-168 over y slabs (param->jtot is the
 * slab's row count, g_jtot the global one) */
static void mpi_compute_potential(const orc_param_t *param, int rkstep, orc_state_t *s){
  const int itot = param->itot;
  const int jtot = param->jtot;        /* local rows (the index macros use it) */
  const int gj = g_jtot;
  const int nwave = gj/2+1;
  const double dy = param->dy;
  double *psi = s->psi;
  const int ni = slab_size(itot, g_rank), nw = slab_size(nwave, g_rank);
  const int w0 = slab_off(nwave, g_rank);
  double *qrx = xcalloc((size_t)itot*jtot, sizeof(double));
  double *qry = xcalloc((size_t)ni*gj, sizeof(double));
  double complex *qcy = xcalloc((size_t)ni*nwave, sizeof(double complex));
  double complex *qcx = xcalloc((size_t)nw*itot, sizeof(double complex));
#define QRX(I, J) (qrx[((J)-1)*(itot)+((I)-1)])
  const double gamma = param->gamma[rkstep];
  const double dt = param->dt;
  const double *ux = s->ux;
  const double *uy = s->uy;
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
  mpi_transpose(gj, itot, sizeof(double), qrx, qry);
  double complex *tw = twiddles(gj);
  double complex *work = xcalloc(4*(size_t)gj, sizeof(double complex));
  for(int i=0; i<ni; i++) rfft_tw(gj, qry + (size_t)i*gj, qcy + (size_t)i*nwave, tw, work);
  mpi_transpose(itot, nwave, sizeof(double complex), qcy, qcx);
  {
    double *l = xcalloc(itot, sizeof(double)), *c = xcalloc(itot, sizeof(double));
    double *u = xcalloc(itot, sizeof(double));
    double complex *q = xcalloc(itot, sizeof(double complex));
    for(int jl=0; jl<nw; jl++){
      const int j = w0+jl+1;  /* global 1-based wavenumber row */
      double eigenvalue = -4./pow(dy, 2.)*pow(sin(M_PI*(j-1)/(1.*gj)), 2.);
      for(int i=1; i<=itot; i++){
        l[i-1] = 1./DXC(i  )/DXF(i);
        u[i-1] = 1./DXC(i+1)/DXF(i);
      }
      for(int i=1; i<=itot; i++) c[i-1] = -l[i-1]-u[i-1]+eigenvalue;
      c[     0] += l[     0];
      c[itot-1] += u[itot-1];
      for(int i=0; i<itot; i++) q[i] = qcx[(size_t)jl*itot+i];
      if(j == 1 && (g_variant & ORC_VARIANT_PIN_GAUGE)){
        l[itot-1] = 0.;
        c[itot-1] = 1.;
        q[itot-1] = 0.;
      }
      orc_zgtsv_b(itot, l, c, u, q);
      for(int i=0; i<itot; i++) qcx[(size_t)jl*itot+i] = q[i];
    }
    free(l); free(c); free(u); free(q);
  }
  mpi_transpose(nwave, itot, sizeof(double complex), qcx, qcy);
  for(int i=0; i<ni; i++) irfft_tw(gj, qcy + (size_t)i*nwave, qry + (size_t)i*gj, tw, work);
  free(work);
  free(tw);
  mpi_transpose(itot, gj, sizeof(double), qry, qrx);
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      PSI(i, j) = QRX(i, j)/(1.*gj);
  orc_fluid_update_boundaries_p(param, psi);
#undef QRX
  free(qrx); free(qry); free(qcx); free(qcy);
}

/* src/main.c:79-101 */
static void mpi_integrate(const orc_param_t *param, orc_state_t *s){
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
    mpi_compute_potential(param, rkstep, s);
    orc_fluid_correct_velocity(param, rkstep, s);
    orc_fluid_update_pressure(param, rkstep, s);
  }
}

/* rows joff+1 .. joff+jsize of a full (gj+2)-row array into a slab array */
static double *slab_copy(const double *full, int pitch, int joff, int jsize){
  double *f = xcalloc((size_t)(jsize+2)*pitch, sizeof(double));
  memcpy(f, full + (size_t)joff*pitch, sizeof(double)*(size_t)(jsize+2)*pitch);
  return f;
}

/* gather the interior rows of a slab field on rank 0 and append them to fp */
static void gather_write(FILE *fp, const double *f, int pitch, int jsize, int gj){
  int *cnt = xcalloc(g_size, sizeof(int)), *dsp = xcalloc(g_size, sizeof(int));
  for(int r=0; r<g_size; r++){
    cnt[r] = slab_size(gj, r)*pitch;
    dsp[r] = slab_off(gj, r)*pitch;
  }
  double *all = g_rank == 0 ? xcalloc((size_t)gj*pitch, sizeof(double)) : NULL;
  MPI_Gatherv(f + pitch, jsize*pitch, MPI_DOUBLE, all, cnt, dsp, MPI_DOUBLE, 0, MPI_COMM_WORLD);
  if(g_rank == 0){
    fwrite(all, sizeof(double), (size_t)gj*pitch, fp);
    free(all);
  }
  free(cnt); free(dsp);
}

int main(int argc, char **argv){
  MPI_Init(&argc, &argv);
  MPI_Comm_rank(MPI_COMM_WORLD, &g_rank);
  MPI_Comm_size(MPI_COMM_WORLD, &g_size);
  if(argc < 10){
    if(g_rank == 0)
      fprintf(stderr, "usage: %s itot jtot Ra Pr implicitx implicity seed amplitude nsteps "
                      "[stretch] [pin_gauge] [out_file]\n", argv[0]);
    MPI_Finalize();
    return 2;
  }
  const int itot = atoi(argv[1]), gj = atoi(argv[2]);
  const double Ra = atof(argv[3]), Pr = atof(argv[4]);
  const int imx = atoi(argv[5]), imy = atoi(argv[6]);
  const uint64_t seed = strtoull(argv[7], NULL, 10);
  const double amp = atof(argv[8]);
  const int nsteps = atoi(argv[9]);
  const double stretch = argc > 10 ? atof(argv[10]) : (double)itot;
  const int pin = argc > 11 ? atoi(argv[11]) : 0;
  const char *out = argc > 12 ? argv[12] : NULL;
  g_jtot = gj;
  g_ym = (g_rank-1+g_size)%g_size;
  g_yp = (g_rank+1)%g_size;
  const int jsize = slab_size(gj, g_rank), joff = slab_off(gj, g_rank);
  if(jsize < 1){
    if(g_rank == 0) fprintf(stderr, "more ranks than rows\n");
    MPI_Finalize();
    return 2;
  }
  orc_set_variant(pin ? ORC_VARIANT_PIN_GAUGE : 0);
  /* global parameters: grid, RK coefficients and the initial temperature */
  orc_param_t gp;
  memset(&gp, 0, sizeof gp);
  gp.itot = itot; gp.jtot = gj; gp.lx = 1.; gp.ly = 2.; gp.stretch = stretch;
  gp.Ra = Ra; gp.Pr = Pr; gp.with_temperature = 1; gp.with_thermal_forcing = 1;
  gp.xf = xcalloc(itot+1, sizeof(double)); gp.xc = xcalloc(itot+2, sizeof(double));
  gp.dxf = xcalloc(itot, sizeof(double)); gp.dxc = xcalloc(itot+1, sizeof(double));
  gp.yf = xcalloc(gj+1, sizeof(double)); gp.yc = xcalloc(gj, sizeof(double));
  orc_set_coordinate(&gp);
  orc_set_rk_coefs(&gp);
  for(int a=0; a<2; a++) for(int b=0; b<2; b++) gp.expimp_wtimes[a][b] = DBL_MAX;
  gp.expimp_wtimes[imy][imx] = 1.;
  double *temp_full = xcalloc((size_t)(gj+2)*(itot+2), sizeof(double));
  orc_init_temperature(&gp, seed, amp, 0, temp_full);
  /* this rank's slab: the same parameters with jtot = jsize (dy stays global) */
  orc_param_t lp = gp;
  lp.jtot = jsize;
  orc_state_t st;
  memset(&st, 0, sizeof st);
  st.temp = slab_copy(temp_full, itot+2, joff, jsize);
  free(temp_full);
  st.ux = xcalloc((size_t)(jsize+2)*(itot+1), sizeof(double));
  st.uy = xcalloc((size_t)(jsize+2)*(itot+2), sizeof(double));
  st.p = xcalloc((size_t)(jsize+2)*(itot+2), sizeof(double));
  st.psi = xcalloc((size_t)(jsize+2)*(itot+2), sizeof(double));
  st.srcuxa = xcalloc((size_t)jsize*(itot-1), sizeof(double));
  st.srcuxb = xcalloc((size_t)jsize*(itot-1), sizeof(double));
  st.srcuxg = xcalloc((size_t)jsize*(itot-1), sizeof(double));
  st.tempforcex = xcalloc((size_t)jsize*(itot-1), sizeof(double));
  st.srcuya = xcalloc((size_t)jsize*itot, sizeof(double));
  st.srcuyb = xcalloc((size_t)jsize*itot, sizeof(double));
  st.srcuyg = xcalloc((size_t)jsize*itot, sizeof(double));
  st.srctempa = xcalloc((size_t)jsize*itot, sizeof(double));
  st.srctempb = xcalloc((size_t)jsize*itot, sizeof(double));
  st.srctempg = xcalloc((size_t)jsize*itot, sizeof(double));
  g_halo_rows = mpi_halo_rows;
  g_solve_y = mpi_solve_y;
  g_min_allreduce = mpi_min;
  orc_fluid_update_boundaries_ux(&lp, st.ux);
  orc_fluid_update_boundaries_uy(&lp, st.uy);
  orc_fluid_update_boundaries_p(&lp, st.p);
  orc_temperature_update_boundaries_temp(&lp, st.temp);
  MPI_Barrier(MPI_COMM_WORLD);
  const double t0 = MPI_Wtime();
  for(int n=0; n<nsteps; n++){
    orc_param_decide_dt(&lp, &st);
    mpi_integrate(&lp, &st);
    lp.step += 1;
    lp.time += lp.dt;
  }
  MPI_Barrier(MPI_COMM_WORLD);
  const double el = MPI_Wtime()-t0;
  if(g_rank == 0){
    printf("{\"ranks\": %d, \"itot\": %d, \"jtot\": %d, \"steps\": %d, \"seconds\": %.6f, "
           "\"cell_steps_per_s\": %.6e, \"dt\": %.17g, \"time\": %.17g}\n",
           g_size, itot, gj, nsteps, el, (double)itot*gj*nsteps/el, lp.dt, lp.time);
    fflush(stdout);
  }
  if(out){
    FILE *fp = g_rank == 0 ? fopen(out, "wb") : NULL;
    if(g_rank == 0 && !fp){ perror(out); MPI_Abort(MPI_COMM_WORLD, 1); }
    gather_write(fp, st.ux, itot+1, jsize, gj);
    gather_write(fp, st.uy, itot+2, jsize, gj);
    gather_write(fp, st.temp, itot+2, jsize, gj);
    gather_write(fp, st.p, itot+2, jsize, gj);
    if(g_rank == 0) fclose(fp);
  }
  MPI_Finalize();
  return 0;
}
