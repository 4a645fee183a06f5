/* Monitoring output (reference: include/logging.h, src/logging.c, src/analyses.c). */
#ifndef LOGGING_H
#define LOGGING_H
#include "structure.h"
extern int logging(param_t *param, const parallel_t *parallel, const fluid_t *fluid, const temperature_t *temperature);
/* the five quantities of src/logging.c on host arrays of the reference layout:
 * out = {divmax, divsum, mom_x, mom_y, K_x, K_y, H, Nu_wall, Nu_inje, Nu_eps_k, Nu_eps_h} */
extern int logging_compute(const param_t *param, const double *ux, const double *uy, const double *temp, double out[11]);
/* the same on plain arrays (no param_t needed) */
extern int sns_logging_compute_raw(const int itot, const int jtot, const double lx, const double ly,
                                   const double Ra, const double Pr, const double *dxf, const double *dxc,
                                   const double dy, const double *ux, const double *uy, const double *temp,
                                   double out[11]);
#endif
