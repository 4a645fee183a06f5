/*
 * Run parameters (reference: include/param.h).  param_t keeps the reference's
 * members; the coordinate arrays are host arrays with the reference extents.
 * Added members: the GPU instance (gpu), the forced diffusive treatment
 * (force_treatment, from the environment variables implicitx / implicity),
 * the initial-condition controls and the device index.
 */
#ifndef PARAM_H
#define PARAM_H

#include <stdbool.h>
#include <stdint.h>

#include "structure.h"
#include "sns_arrays.h"
#include "sns_hip.h"

#define RKSTEPMAX 3
#define LX 1.

typedef struct rkcoef_t_ {
  double alpha;
  double beta;
  double gamma;
} rkcoef_t;

typedef struct schedule_t_ {
  double rate;
  double after;
  double next;
} schedule_t;

struct param_t_ {
  /* restart / initialise */
  bool load_flow_field;
  char *dirname_restart;
  /* fluid-temperature coupling */
  bool with_temperature;
  bool with_thermal_forcing;
  /* domain sizes */
  int itot, jtot;
  double lx, ly;
  /* grid sizes (host arrays) */
  double stretch;
  double *xf, *xc;
  double *dxf, *dxc;
  double *yf, *yc;
  double dy;
  /* non-dimensional params */
  double Ra, Pr;
  /* temporal integration */
  rkcoef_t rkcoefs[3];
  double time, dt;
  int step, implicitx, implicity;
  double expimp_wtimes[2][2];
  /* when to stop, when to write log, etc. */
  double timemax, wtimemax;
  schedule_t log, save, stat;
  /* --- additions of the GPU build --- */
  int force_treatment;          /* 1: implicitx / implicity fixed by the environment */
  int forced_implicitx, forced_implicity;
  uint64_t ic_seed;             /* seed of the counter-hash temperature noise */
  double ic_amplitude;          /* noise amplitude A: T += A*(2U-1) */
  bool ic_linear;               /* add the conduction profile (reference default) */
  int device;
  snsg_ctx *gpu;                /* GPU instance holding device metrics and scratch */
};

extern param_t *param_init(void);
extern int param_finalise(param_t *param);

extern int param_decide_dt(param_t *param, const parallel_t *parallel, const fluid_t *fluid);
extern int param_estimate_cost(param_t *param, const parallel_t *parallel);

/* grid and RK coefficients on caller-provided arrays (used by param_init) */
extern int sns_set_coordinate(const int itot, const int jtot, const double lx, const double ly, const double stretch,
                              double *xf, double *xc, double *dxf, double *dxc, double *yf, double *yc, double *dy);
extern int sns_rk_coefs(double alpha[3], double beta[3], double gamma[3]);

/* the GPU instance used by the parallel_* helpers (the one param_init created) */
extern snsg_ctx *param_current_gpu(void);

#endif
