/*
 * Run parameters, grid, RK coefficients, time step and treatment choice
 * (reference: src/param/init.c, decide_dt.c, estimate_cost.c).
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>
#include <limits.h>
#include <float.h>
#include <math.h>
#include "common.h"
#include "param.h"
#include "parallel.h"
#include "fluid.h"
#include "temperature.h"
#include "fileio.h"

static snsg_ctx *g_current_gpu = NULL;

snsg_ctx *param_current_gpu(void){
  if(g_current_gpu == NULL){
    fprintf(stderr, "no GPU instance: call param_init first\n");
    abort();
  }
  return g_current_gpu;
}

static int verbose(void){
  const char *v = getenv("sns_quiet");
  return !(v != NULL && 0 == strcmp(v, "true"));
}

#define PRINTF_MAIN(...) do { if(verbose()) printf(__VA_ARGS__); } while(0)

/* ---------------------------------------------------------------------------
 * Configuration from environment variables.  The variable names, defaults
 * and printed values are the reference's interface (src/param/init.c:37-121);
 * the loader is one table walked by one function.  Each entry names a
 * param_t member by its offset; every default is held as a double.
 * ------------------------------------------------------------------------- */

typedef enum { CFG_BOOL, CFG_INT, CFG_SEED, CFG_DOUBLE } cfg_kind;

typedef struct {
  const char *name;   /* environment variable */
  cfg_kind kind;
  size_t member;      /* offsetof(param_t, ...) */
  double fallback;    /* value when the variable is unset (bool: 0 / 1) */
} cfg_entry;

#define CFG(name, kind, member, fallback) {name, kind, offsetof(param_t, member), fallback}

static const cfg_entry g_config[] = {
  CFG("with_temperature",     CFG_BOOL,   with_temperature,     1.),
  CFG("with_thermal_forcing", CFG_BOOL,   with_thermal_forcing, 1.),
  CFG("timemax",              CFG_DOUBLE, timemax,              2.0e+2),
  CFG("wtimemax",             CFG_DOUBLE, wtimemax,             6.0e+2),
  CFG("log_rate",             CFG_DOUBLE, log.rate,             2.0e-1),
  CFG("log_after",            CFG_DOUBLE, log.after,            0.0e+0),
  CFG("save_rate",            CFG_DOUBLE, save.rate,            2.0e+1),
  CFG("save_after",           CFG_DOUBLE, save.after,           0.0e+0),
  CFG("stat_rate",            CFG_DOUBLE, stat.rate,            1.0e-1),
  CFG("stat_after",           CFG_DOUBLE, stat.after,           1.0e+2),
  CFG("itot",                 CFG_INT,    itot,                 128),
  CFG("jtot",                 CFG_INT,    jtot,                 192),
  CFG("ly",                   CFG_DOUBLE, ly,                   2.0e+0),
  CFG("stretch",              CFG_DOUBLE, stretch,              3.0e+0),
  CFG("Ra",                   CFG_DOUBLE, Ra,                   1.0e+8),
  CFG("Pr",                   CFG_DOUBLE, Pr,                   1.0e+1),
  /* additions of this build: initial condition and device */
  CFG("ic_seed",              CFG_SEED,   ic_seed,              0),
  CFG("ic_amplitude",         CFG_DOUBLE, ic_amplitude,         5.0e-2),
  CFG("ic_linear",            CFG_BOOL,   ic_linear,            1.),
  CFG("sns_device",           CFG_INT,    device,               0),
};

/* a boolean variable counts when it starts with "true" or "false"; anything
 * else leaves the default */
static int parse_bool(const char *text, bool *out){
  if(text == NULL) return 0;
  if(0 == strncmp(text, "true", 4)){ *out = true; return 1; }
  if(0 == strncmp(text, "false", 5)){ *out = false; return 1; }
  return 0;
}

/* decimal integer, saturated to the range of int */
static int parse_int(const char *text){
  long v = strtol(text, NULL, 10);
  if(v < INT_MIN) return INT_MIN;
  if(v > INT_MAX) return INT_MAX;
  return (int)v;
}

static void load_entry(param_t *param, const cfg_entry *e){
  const char *text = getenv(e->name);
  void *dst = (char *)param+e->member;
  const char *note = " (default)";
  switch(e->kind){
    case CFG_BOOL: {
      bool v = e->fallback != 0.;
      if(parse_bool(text, &v)) note = "";
      *(bool *)dst = v;
      PRINTF_MAIN("  %s: %s%s\n", e->name, v ? "true" : "false", note);
      break;
    }
    case CFG_INT:
    case CFG_SEED: {
      int v = (int)e->fallback;
      if(text != NULL){ v = parse_int(text); note = ""; }
      if(e->kind == CFG_INT) *(int *)dst = v;
      else *(uint64_t *)dst = (uint64_t)v;
      PRINTF_MAIN("  %s: %d%s\n", e->name, v, note);
      break;
    }
    case CFG_DOUBLE: {
      double v = e->fallback;
      if(text != NULL){ v = strtod(text, NULL); note = ""; }
      *(double *)dst = v;
      PRINTF_MAIN("  %s: %.3e%s\n", e->name, v, note);
      break;
    }
  }
}

/* a copy of an environment variable's text, or NULL when it is unset */
static char *env_copy(const char *name){
  const char *text = getenv(name);
  if(text == NULL) return NULL;
  const size_t n = strlen(text)+1;
  char *copy = common_calloc(n, sizeof(char));
  memcpy(copy, text, n);
  return copy;
}

static int load_config(param_t *param){
  PRINTF_MAIN("------- parameters are loaded -------\n");
  param->dirname_restart = env_copy("dirname_restart");
  param->load_flow_field = param->dirname_restart != NULL;
  if(param->load_flow_field){
    PRINTF_MAIN("  flow fields are loaded: %s\n", param->dirname_restart);
  }else{
    PRINTF_MAIN("  flow fields are initialised\n");
  }
  for(size_t k = 0; k < sizeof g_config/sizeof g_config[0]; k++){
    load_entry(param, &g_config[k]);
  }
  param->lx = LX;
  /* addition of this build: a diffusive treatment forced by implicitx /
   * implicity ("true" or "1" selects implicit; either variable forces both) */
  {
    const char *sx = getenv("implicitx");
    const char *sy = getenv("implicity");
    bool vx = false, vy = false;
    param->force_treatment = (sx != NULL || sy != NULL);
    param->forced_implicitx = (parse_bool(sx, &vx) && vx) || (sx != NULL && 0 == strcmp(sx, "1"));
    param->forced_implicity = (parse_bool(sy, &vy) && vy) || (sy != NULL && 0 == strcmp(sy, "1"));
    if(param->force_treatment){
      PRINTF_MAIN("  treatment forced: x %s, y %s\n",
          param->forced_implicitx ? "IMP" : "EXP", param->forced_implicity ? "IMP" : "EXP");
    }
  }
  PRINTF_MAIN("-------------------------------------\n");
  return 0;
}

/* src/param/init.c:124-195, on caller-provided arrays */
int sns_set_coordinate(const int itot, const int jtot, const double lx, const double ly, const double stretch_d,
                       double *xf, double *xc, double *dxf, double *dxc, double *yf, double *yc, double *dy_out){
  {
    int stretch = stretch_d;  /* truncated to an integer */
    if(stretch < 0) stretch = 0;
    if(stretch >= itot/2){
      const double dx = lx/itot;
      for(int i=1; i<=itot+1; i++) XF(i) = 1.*(i-1)*dx;
    }else{
      for(int i=1; i<=itot+1; i++) XF(i) = cos(M_PI*(1.*(i-1)+stretch)/(itot+2.*stretch));
      const double lo = XF(1);
      const double hi = XF(itot+1);
      for(int i=1; i<=itot+1; i++) XF(i) = lx*(XF(i)-lo)/(hi-lo);
    }
    XF(     1) = 0.;
    XF(itot+1) = lx;
  }
  for(int i=1; i<=itot; i++) DXF(i) = XF(i+1)-XF(i);
  XC(0) = XF(1);
  for(int i=1; i<=itot; i++) XC(i) = 0.5*(XF(i)+XF(i+1));
  XC(itot+1) = XF(itot+1);
  for(int i=1; i<=itot+1; i++) DXC(i) = XC(i)-XC(i-1);
  const double dy = ly/jtot;
  for(int j=1; j<=jtot+1; j++) YF(j) = 1.*(j-1)*dy;
  for(int j=1; j<=jtot; j++) YC(j) = 0.5*(2*j-1)*dy;
  *dy_out = dy;
  return 0;
}

static int set_coordinate(param_t *param){
  const int itot = param->itot;
  const int jtot = param->jtot;
  param->xf  = common_calloc(1, XF_MEMSIZE);
  param->xc  = common_calloc(1, XC_MEMSIZE);
  param->dxf = common_calloc(1, DXF_MEMSIZE);
  param->dxc = common_calloc(1, DXC_MEMSIZE);
  param->yf  = common_calloc(1, YF_MEMSIZE);
  param->yc  = common_calloc(1, YC_MEMSIZE);
  return sns_set_coordinate(itot, jtot, param->lx, param->ly, param->stretch, param->xf, param->xc,
                            param->dxf, param->dxc, param->yf, param->yc, &param->dy);
}

/* src/param/init.c:197-213 */
int sns_rk_coefs(double alpha[3], double beta[3], double gamma[3]){
  alpha[0] =  32./60.;
  beta [0] =   0./60.;
  alpha[1] =  25./60.;
  beta [1] = -17./60.;
  alpha[2] =  45./60.;
  beta [2] = -25./60.;
  for(int rkstep=0; rkstep<RKSTEPMAX; rkstep++) gamma[rkstep] = alpha[rkstep]+beta[rkstep];
  return 0;
}

static int set_rk_coefs(param_t *param){
  double a[3], b[3], g[3];
  sns_rk_coefs(a, b, g);
  for(int k=0; k<RKSTEPMAX; k++){
    param->rkcoefs[k].alpha = a[k];
    param->rkcoefs[k].beta  = b[k];
    param->rkcoefs[k].gamma = g[k];
  }
  return 0;
}

/* src/param/init.c:215-235 */
static double compute_next_schedule(const double rate, const double after, const double time){
  const double small = 1.e-8;
  if(time < after){
    return rate*ceil((after+small)/rate);
  }
  return rate*ceil((time+small)/rate);
}

static int create_gpu(param_t *param){
  snsg_config cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.itot = param->itot;
  cfg.jtot = param->jtot;
  cfg.ly = param->ly;
  cfg.Ra = param->Ra;
  cfg.Pr = param->Pr;
  cfg.with_temperature = param->with_temperature;
  cfg.with_thermal_forcing = param->with_thermal_forcing;
  cfg.implicitx = param->implicitx;
  cfg.implicity = param->implicity;
  for(int k=0; k<RKSTEPMAX; k++){
    cfg.alpha[k] = param->rkcoefs[k].alpha;
    cfg.beta[k]  = param->rkcoefs[k].beta;
    cfg.gamma[k] = param->rkcoefs[k].gamma;
  }
  int rc = snsg_create(param->device, &cfg, param->dxf, param->dxc, &param->gpu);
  common_check_gpu(rc, "snsg_create");
  g_current_gpu = param->gpu;
  return 0;
}

/* src/param/init.c:237-259 */
param_t *param_init(void){
  param_t *param = common_calloc(1, sizeof(param_t));
  load_config(param);
  set_coordinate(param);
  if(param->load_flow_field){
    fileio_r_0d_serial(param->dirname_restart, "step", sizeof(int),    &(param->step));
    fileio_r_0d_serial(param->dirname_restart, "time", sizeof(double), &(param->time));
  }else{
    param->step = 0;
    param->time = 0.;
  }
  set_rk_coefs(param);
  param->log.next  = compute_next_schedule(param->log.rate , param->log.after, param->time);
  param->save.next = compute_next_schedule(param->save.rate, param->save.after, param->time);
  param->stat.next = compute_next_schedule(param->stat.rate, param->stat.after, param->time);
  if(param->force_treatment){
    param->implicitx = param->forced_implicitx;
    param->implicity = param->forced_implicity;
  }
  create_gpu(param);
  return param;
}

int param_finalise(param_t *param){
  if(param->gpu != NULL){
    if(g_current_gpu == param->gpu) g_current_gpu = NULL;
    snsg_destroy(param->gpu);
  }
  common_free(param->dirname_restart);
  common_free(param->xf);
  common_free(param->xc);
  common_free(param->dxf);
  common_free(param->dxc);
  common_free(param->yf);
  common_free(param->yc);
  common_free(param);
  return 0;
}

/* src/param/decide_dt.c:11-101.  The advective bound is a device
 * reduction (min is exact, so it equals the sequential loop); the rest is
 * evaluated on the host with the reference's operations. */
int param_decide_dt(param_t *param, const parallel_t *parallel, const fluid_t *fluid){
  (void)parallel;
  const double safe_factors[2] = {0.7, 0.9};
  const int itot = param->itot;
  const double Ra = param->Ra;
  const double Pr = param->Pr;
  const double diffusivity = fmax(sqrt(Pr)/sqrt(Ra), 1./sqrt(Pr)/sqrt(Ra));
  const double lx = param->lx;
  const double *dxc = param->dxc;
  const double dy = param->dy;
  double dt_adv;
  common_check_gpu(snsg_param_decide_dt(param->gpu, fluid->ux, fluid->uy, &dt_adv), "param_decide_dt");
  dt_adv *= safe_factors[0];
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
  common_check_gpu(snsg_set_treatment(param->gpu, param->implicitx, param->implicity), "set_treatment");
  common_check_gpu(snsg_set_dt(param->gpu, param->dt), "set_dt");
  return 0;
}

/* src/param/estimate_cost.c:11-42: run substage 0 on fresh fields until the
 * wall-time budget is used up; on the GPU every iteration ends with a
 * stream synchronisation so that the count reflects completed work */
static int emulate(param_t *param, const parallel_t *parallel, const double wtimemax){
  int niter;
  fluid_t *fluid = fluid_init(param, parallel);
  temperature_t *temperature = temperature_init(param, parallel);
  common_check_gpu(snsg_set_treatment(param->gpu, param->implicitx, param->implicity), "set_treatment");
  common_check_gpu(snsg_set_dt(param->gpu, param->dt), "set_dt");
  const double t0 = parallel_get_wtime(MPI_MAX);
  for(niter=0; ; niter++){
    const int rkstep = 0;
    fluid_update_boundaries_ux(param, parallel, fluid->ux);
    fluid_update_boundaries_uy(param, parallel, fluid->uy);
    fluid_update_boundaries_p(param, parallel, fluid->p);
    if(param->with_temperature){
      temperature_update_boundaries_temp(param, parallel, temperature->temp);
      temperature_compute_force(param, parallel, temperature);
      temperature_update_temp(param, parallel, rkstep, fluid, temperature);
    }
    fluid_update_velocity(param, parallel, rkstep, fluid, temperature);
    fluid_compute_potential(param, parallel, rkstep, fluid);
    fluid_correct_velocity(param, parallel, rkstep, fluid);
    fluid_update_pressure(param, parallel, rkstep, fluid);
    common_check_gpu(snsg_synchronize(param->gpu), "emulate");
    if(parallel_get_wtime(MPI_MIN)-t0 > wtimemax) break;
  }
  fluid_finalise(fluid);
  temperature_finalise(temperature);
  return niter;
}

/* src/param/estimate_cost.c:44-74.  With a forced treatment (environment
 * variables implicitx / implicity) the cost table makes the reference's
 * selection rule pick that treatment and no emulation runs. */
int param_estimate_cost(param_t *param, const parallel_t *parallel){
  if(param->force_treatment){
    for(int y=0; y<2; y++)
      for(int x=0; x<2; x++)
        param->expimp_wtimes[y][x] = DBL_MAX;
    param->expimp_wtimes[param->forced_implicity][param->forced_implicitx] = 1.;
    return 0;
  }
  const char *s = getenv("estimate_wtime");
  const double wtimemax = s ? strtod(s, NULL) : 1.;
  const double dt = 1.e-8;
  PRINTF_MAIN("------- Check optimal diffusive term treatment -------\n");
  for(int implicity=0; implicity<2; implicity++){
    for(int implicitx=0; implicitx<2; implicitx++){
      param->dt = dt;
      param->implicitx = implicitx;
      param->implicity = implicity;
      int niter = emulate(param, parallel, wtimemax);
      if(niter < 1) niter = 1;
      param->expimp_wtimes[implicity][implicitx] = wtimemax/(1.*niter);
      PRINTF_MAIN("%s(x)-%s(y): %8d iterations in %.1e [s]\n",
          implicitx ? "IMP" : "EXP", implicity ? "IMP" : "EXP", niter, wtimemax);
    }
  }
  PRINTF_MAIN("------------------------------------------------------\n");
  return 0;
}
