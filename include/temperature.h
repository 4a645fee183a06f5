/* Temperature field (reference: include/temperature.h); arrays are device pointers. */
#ifndef TEMPERATURE_H
#define TEMPERATURE_H

#include "structure.h"
#include "sns_arrays.h"

struct temperature_t_ {
  double *temp;
  double *tempforcex;
  double *srctempa, *srctempb, *srctempg;
};

extern temperature_t *temperature_init(const param_t *param, const parallel_t *parallel);
extern int temperature_finalise(temperature_t *temperature);

extern int temperature_compute_force(const param_t *param, const parallel_t *parallel, temperature_t *temperature);
extern int temperature_update_temp(const param_t *param, const parallel_t *parallel, const int rkstep, const fluid_t *fluid, temperature_t *temperature);

extern int temperature_update_boundaries_temp(const param_t *param, const parallel_t *parallel, double *temp);

/* seeded initial temperature on a host array of the reference layout (interior only) */
#include <stdint.h>
extern int sns_initial_temperature(const int itot, const int jtot, const double *xc, const uint64_t seed,
                                   const double amplitude, const int linear, double *temp);

/* copy temp to a host array of the reference layout */
extern int temperature_download(const param_t *param, const temperature_t *temperature, double *temp);

#define TEMP_XM +0.5
#define TEMP_XP -0.5

#endif
