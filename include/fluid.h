/*
 * Velocity, pressure and scalar potential (reference: include/fluid.h).  The
 * arrays are device pointers with the reference layouts (sns_arrays.h);
 * fluid_download copies them to host arrays of the same layout.
 */
#ifndef FLUID_H
#define FLUID_H

#include "structure.h"
#include "sns_arrays.h"

struct fluid_t_ {
  double *ux, *uy;
  double *p, *psi;
  double *srcuxa, *srcuxb, *srcuxg;
  double *srcuya, *srcuyb, *srcuyg;
};

extern fluid_t *fluid_init(const param_t *param, const parallel_t *parallel);
extern int fluid_finalise(fluid_t *fluid);

/* boundary condition and halo update */
extern int fluid_update_boundaries_ux(const param_t *param, const parallel_t *parallel, double *ux);
extern int fluid_update_boundaries_uy(const param_t *param, const parallel_t *parallel, double *uy);
extern int fluid_update_boundaries_p(const param_t *param, const parallel_t *parallel, double *p);

extern int fluid_update_velocity(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid, const temperature_t *temperature);
extern int fluid_compute_potential(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid);
extern int fluid_correct_velocity(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid);
extern int fluid_update_pressure(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid);

/* copy ux, uy, p (and psi when psi != NULL) to host arrays of the reference layout */
extern int fluid_download(const param_t *param, const fluid_t *fluid, double *ux, double *uy, double *p, double *psi);

#endif
