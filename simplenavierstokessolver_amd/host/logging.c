/*
 * Monitoring (reference: src/logging.c:13-281, src/analyses.c:8-121): the
 * divergence, momentum, quadratic-energy and Nusselt diagnostics are
 * evaluated on the device (snsg_diagnostics) and appended to output/log/*.dat.
 * logging_compute evaluates the same quantities on host arrays with the
 * reference's sequential loops.
 */
#include <stdio.h>
#include <math.h>
#include "common.h"
#include "param.h"
#include "parallel.h"
#include "fluid.h"
#include "temperature.h"
#include "fileio.h"
#include "logging.h"

/* src/analyses.c:8-45 */
static double kinetic_dissipation_x(const param_t *param, const double *ux, int i, int j){
  const int itot = param->itot;
  const double *dxf = param->dxf, *dxc = param->dxc;
  const double dy = param->dy;
  double s = 0.;
  if(i != 1){ double d = UX(i, j)-UX(i-1, j); double dd = d/DXF(i-1); s += 0.5/DXC(i)*d*dd; }
  if(i != itot+1){ double d = UX(i+1, j)-UX(i, j); double dd = d/DXF(i); s += 0.5/DXC(i)*d*dd; }
  { double d = UX(i, j)-UX(i, j-1); double dd = d/dy; s += 0.5/dy*d*dd; }
  { double d = UX(i, j+1)-UX(i, j); double dd = d/dy; s += 0.5/dy*d*dd; }
  return sqrt(param->Pr)/sqrt(param->Ra)*s;
}

/* src/analyses.c:47-83 */
static double kinetic_dissipation_y(const param_t *param, const double *uy, int i, int j){
  const int itot = param->itot;
  const double *dxf = param->dxf, *dxc = param->dxc;
  const double dy = param->dy;
  double s = 0.;
  { double c = i == 1 ? 1. : 0.5; double d = UY(i, j)-UY(i-1, j); double dd = d/DXC(i); s += c/DXF(i)*d*dd; }
  { double c = i == itot ? 1. : 0.5; double d = UY(i+1, j)-UY(i, j); double dd = d/DXC(i+1); s += c/DXF(i)*d*dd; }
  { double d = UY(i, j)-UY(i, j-1); double dd = d/dy; s += 0.5/dy*d*dd; }
  { double d = UY(i, j+1)-UY(i, j); double dd = d/dy; s += 0.5/dy*d*dd; }
  return sqrt(param->Pr)/sqrt(param->Ra)*s;
}

/* src/analyses.c:85-121 */
static double thermal_dissipation(const param_t *param, const double *temp, int i, int j){
  const int itot = param->itot;
  const double *dxf = param->dxf, *dxc = param->dxc;
  const double dy = param->dy;
  double s = 0.;
  { double c = i == 1 ? 1. : 0.5; double d = TEMP(i, j)-TEMP(i-1, j); double dd = d/DXC(i); s += c/DXF(i)*d*dd; }
  { double c = i == itot ? 1. : 0.5; double d = TEMP(i+1, j)-TEMP(i, j); double dd = d/DXC(i+1); s += c/DXF(i)*d*dd; }
  { double d = TEMP(i, j)-TEMP(i, j-1); double dd = d/dy; s += 0.5/dy*d*dd; }
  { double d = TEMP(i, j+1)-TEMP(i, j); double dd = d/dy; s += 0.5/dy*d*dd; }
  return 1./sqrt(param->Pr)/sqrt(param->Ra)*s;
}

int logging_compute(const param_t *param, const double *ux, const double *uy, const double *temp, double out[11]){
  const int itot = param->itot;
  const int jtot = param->jtot;
  const double *dxf = param->dxf, *dxc = param->dxc;
  const double dy = param->dy, lx = param->lx, ly = param->ly;
  const double Ra = param->Ra, Pr = param->Pr;
  /* src/logging.c:36-80 */
  double divmax = 0., divsum = 0.;
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      double div = (UX(i+1, j)-UX(i, j))/DXF(i)+(UY(i, j+1)-UY(i, j))/dy;
      divmax = fmax(divmax, fabs(div));
      divsum += div;
    }
  }
  /* src/logging.c:82-122 */
  double mx = 0., my = 0.;
  for(int j=1; j<=jtot; j++) for(int i=1; i<=itot+1; i++) mx += UX(i, j)*(DXC(i)*dy);
  for(int j=1; j<=jtot; j++) for(int i=1; i<=itot; i++) my += UY(i, j)*(DXF(i)*dy);
  /* src/logging.c:124-173 */
  double kx = 0., ky = 0., h = 0.;
  for(int j=1; j<=jtot; j++) for(int i=1; i<=itot+1; i++) kx += 0.5*pow(UX(i, j), 2.)*(DXC(i)*dy);
  for(int j=1; j<=jtot; j++) for(int i=1; i<=itot; i++) ky += 0.5*pow(UY(i, j), 2.)*(DXF(i)*dy);
  if(temp) for(int j=1; j<=jtot; j++) for(int i=1; i<=itot; i++) h += 0.5*pow(TEMP(i, j), 2.)*(DXF(i)*dy);
  /* src/logging.c:175-272 */
  double nu_wall = 0., nu_inje = 0., nu_eps_k = 0., nu_eps_h = 0.;
  if(temp){
    for(int j=1; j<=jtot; j++){
      double dtdx0 = (TEMP(1, j)-TEMP(0, j))/DXC(1);
      double dtdx1 = (TEMP(itot+1, j)-TEMP(itot, j))/DXC(itot+1);
      nu_wall -= 0.5*(dtdx0+dtdx1)*dy;
    }
    nu_wall /= ly;
    for(int j=1; j<=jtot; j++){
      for(int i=1; i<=itot+1; i++){
        double integrand = UX(i, j)*0.5*(TEMP(i-1, j)+TEMP(i, j));
        double cellsize = DXC(i)*dy;
        nu_inje += sqrt(Pr)*sqrt(Ra)*integrand*cellsize;
      }
    }
    nu_inje /= lx*ly;
    nu_inje += 1.;
    for(int j=1; j<=jtot; j++)
      for(int i=1; i<=itot; i++)
        nu_eps_h += sqrt(Pr)*sqrt(Ra)*thermal_dissipation(param, temp, i, j)*(DXF(i)*dy);
    nu_eps_h /= lx*ly;
  }
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot+1; i++)
      nu_eps_k += sqrt(Pr)*sqrt(Ra)*kinetic_dissipation_x(param, ux, i, j)*(DXC(i)*dy);
  for(int j=1; j<=jtot; j++)
    for(int i=1; i<=itot; i++)
      nu_eps_k += sqrt(Pr)*sqrt(Ra)*kinetic_dissipation_y(param, uy, i, j)*(DXF(i)*dy);
  nu_eps_k /= lx*ly;
  nu_eps_k += 1.;
  out[0] = divmax; out[1] = divsum; out[2] = mx; out[3] = my;
  out[4] = kx; out[5] = ky; out[6] = h;
  out[7] = nu_wall; out[8] = nu_inje; out[9] = nu_eps_k; out[10] = nu_eps_h;
  return 0;
}

int sns_logging_compute_raw(const int itot, const int jtot, const double lx, const double ly,
                            const double Ra, const double Pr, const double *dxf, const double *dxc,
                            const double dy, const double *ux, const double *uy, const double *temp,
                            double out[11]){
  param_t param;
  param.itot = itot;
  param.jtot = jtot;
  param.lx = lx;
  param.ly = ly;
  param.Ra = Ra;
  param.Pr = Pr;
  param.dxf = (double *)dxf;
  param.dxc = (double *)dxc;
  param.dy = dy;
  return logging_compute(&param, ux, uy, temp, out);
}

static FILE *open_log(const param_t *param, const char *name){
  return fileio_fopen(name, param->step == 0 ? "w" : "a");
}

int logging(param_t *param, const parallel_t *parallel, const fluid_t *fluid, const temperature_t *temperature){
  (void)parallel;
  double q[11];
  common_check_gpu(snsg_diagnostics(param->gpu, fluid->ux, fluid->uy,
                                    param->with_temperature ? temperature->temp : NULL, q),
                   "logging");
  FILE *fp;
  char scheme[32];
  snprintf(scheme, sizeof scheme, "x: %3s, y: %3s", param->implicitx ? "IMP" : "EXP", param->implicity ? "IMP" : "EXP");
  printf("step %8d, time %7.1f, dt %.2e (%s)\n", param->step, param->time, param->dt, scheme);
  if((fp = open_log(param, FILEIO_LOG "/progress.dat"))){
    fprintf(fp, "step %8d, time %7.1f, dt %.2e (%s)\n", param->step, param->time, param->dt, scheme);
    fileio_fclose(fp);
  }
  if((fp = open_log(param, FILEIO_LOG "/divergence.dat"))){
    fprintf(fp, "%8.2f % .1e % .1e\n", param->time, q[0], q[1]);
    fileio_fclose(fp);
  }
  if((fp = open_log(param, FILEIO_LOG "/momentum.dat"))){
    fprintf(fp, "%8.2f % 18.15e % 18.15e\n", param->time, q[2], q[3]);
    fileio_fclose(fp);
  }
  if((fp = open_log(param, FILEIO_LOG "/energy.dat"))){
    fprintf(fp, "%8.2f % 18.15e % 18.15e % 18.15e\n", param->time, q[4], q[5], q[6]);
    fileio_fclose(fp);
  }
  if((fp = open_log(param, FILEIO_LOG "/nusselt.dat"))){
    fprintf(fp, "%8.2f % 18.15e % 18.15e % 18.15e % 18.15e\n", param->time, q[7], q[8], q[9], q[10]);
    fileio_fclose(fp);
  }
  return 0;
}
