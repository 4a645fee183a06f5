/*
 * Fluid fields on the GPU (reference: src/fluid/init.c, finalise.c,
 * boundary_conditions.c, update_velocity.c, compute_potential.c,
 * correct_velocity.c, update_pressure.c).  Every function keeps the
 * reference's signature and semantics and enqueues the corresponding
 * kernels through the sns_hip.h C-ABI.
 */
#include <stdio.h>
#include <stdlib.h>
#include "common.h"
#include "param.h"
#include "parallel.h"
#include "fluid.h"
#include "temperature.h"
#include "fileio.h"

static double *dev_alloc(const param_t *param, size_t bytes){
  void *p = NULL;
  common_check_gpu(snsg_malloc(param->gpu, bytes, &p), "fluid_init");
  return (double *)p;
}

/* src/fluid/init.c:8-29 */
static int allocate(const param_t *param, const parallel_t *parallel, fluid_t **fluid){
  const int itot = param->itot;
  const int jsize = parallel_get_size(param->jtot, parallel->mpisize, parallel->mpirank);
  *fluid = common_calloc(1, sizeof(fluid_t));
  (*fluid)->ux  = dev_alloc(param, UX_MEMSIZE);
  (*fluid)->uy  = dev_alloc(param, UY_MEMSIZE);
  (*fluid)->p   = dev_alloc(param, P_MEMSIZE);
  (*fluid)->psi = dev_alloc(param, PSI_MEMSIZE);
  (*fluid)->srcuxa = dev_alloc(param, SRCUXA_MEMSIZE);
  (*fluid)->srcuxb = dev_alloc(param, SRCUXA_MEMSIZE);
  (*fluid)->srcuxg = dev_alloc(param, SRCUXA_MEMSIZE);
  (*fluid)->srcuya = dev_alloc(param, SRCUYA_MEMSIZE);
  (*fluid)->srcuyb = dev_alloc(param, SRCUYA_MEMSIZE);
  (*fluid)->srcuyg = dev_alloc(param, SRCUYA_MEMSIZE);
  return 0;
}

/* load one restart array into a host buffer of the halo'd layout and upload it */
static void load_field(const param_t *param, const parallel_t *parallel, const char *name,
                       int ux_like, double *dev, size_t bytes){
  double *host = common_calloc(1, bytes);
  int rc = ux_like ? fileio_r_ux_like_parallel(param->dirname_restart, name, param, parallel, host)
                   : fileio_r_p_like_parallel(param->dirname_restart, name, param, parallel, host);
  if(rc != 0){
    fprintf(stderr, "restart: cannot load %s/%s.npy\n", param->dirname_restart, name);
    abort();
  }
  common_check_gpu(snsg_upload(param->gpu, dev, host, bytes), "restart upload");
  common_free(host);
}

/* src/fluid/init.c:31-83: zero velocity and pressure, or the restart files */
static int init_or_load(const param_t *param, const parallel_t *parallel, fluid_t *fluid){
  const int itot = param->itot;
  const int jsize = parallel_get_size(param->jtot, parallel->mpisize, parallel->mpirank);
  if(param->load_flow_field){
    load_field(param, parallel, "ux", 1, fluid->ux, UX_MEMSIZE);
    load_field(param, parallel, "uy", 0, fluid->uy, UY_MEMSIZE);
    load_field(param, parallel, "p",  0, fluid->p,  P_MEMSIZE);
  }
  fluid_update_boundaries_ux(param, parallel, fluid->ux);
  fluid_update_boundaries_uy(param, parallel, fluid->uy);
  fluid_update_boundaries_p(param, parallel, fluid->p);
  return 0;
}

fluid_t *fluid_init(const param_t *param, const parallel_t *parallel){
  fluid_t *fluid = NULL;
  allocate(param, parallel, &fluid);
  init_or_load(param, parallel, fluid);
  return fluid;
}

/* src/fluid/finalise.c:5-18 */
int fluid_finalise(fluid_t *fluid){
  snsg_ctx *gpu = param_current_gpu();
  double *ptrs[] = {fluid->ux, fluid->uy, fluid->p, fluid->psi, fluid->srcuxa, fluid->srcuxb,
                    fluid->srcuxg, fluid->srcuya, fluid->srcuyb, fluid->srcuyg};
  for(size_t k=0; k<sizeof(ptrs)/sizeof(ptrs[0]); k++) snsg_free(gpu, ptrs[k]);
  common_free(fluid);
  return 0;
}

/* src/fluid/boundary_conditions.c:6-55 */
int fluid_update_boundaries_ux(const param_t *param, const parallel_t *parallel, double *ux){
  (void)parallel;
  common_check_gpu(snsg_fluid_update_boundaries_ux(param->gpu, ux), "fluid_update_boundaries_ux");
  return 0;
}

int fluid_update_boundaries_uy(const param_t *param, const parallel_t *parallel, double *uy){
  (void)parallel;
  common_check_gpu(snsg_fluid_update_boundaries_uy(param->gpu, uy), "fluid_update_boundaries_uy");
  return 0;
}

int fluid_update_boundaries_p(const param_t *param, const parallel_t *parallel, double *p){
  (void)parallel;
  common_check_gpu(snsg_fluid_update_boundaries_p(param->gpu, p), "fluid_update_boundaries_p");
  return 0;
}

/* src/fluid/update_velocity.c:490-498.  SRCUXB <- SRCUXA and SRCUYB <- SRCUYA
 * (lines 32, 126) are pointer exchanges; the kernels then overwrite the A arrays. */
int fluid_update_velocity(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid, const temperature_t *temperature){
  (void)parallel;
  double *t;
  t = fluid->srcuxb; fluid->srcuxb = fluid->srcuxa; fluid->srcuxa = t;
  t = fluid->srcuyb; fluid->srcuyb = fluid->srcuya; fluid->srcuya = t;
  int rc = snsg_fluid_update_velocity(param->gpu, rkstep, fluid->ux, fluid->uy, fluid->p,
      temperature->tempforcex, fluid->srcuxa, fluid->srcuxb, fluid->srcuxg,
      fluid->srcuya, fluid->srcuyb, fluid->srcuyg);
  common_check_gpu(rc, "fluid_update_velocity");
  return 0;
}

/* src/fluid/compute_potential.c:15-168 */
int fluid_compute_potential(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid){
  (void)parallel;
  common_check_gpu(snsg_fluid_compute_potential(param->gpu, rkstep, fluid->ux, fluid->uy, fluid->psi),
                   "fluid_compute_potential");
  return 0;
}

/* src/fluid/correct_velocity.c:6-42 */
int fluid_correct_velocity(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid){
  (void)parallel;
  common_check_gpu(snsg_fluid_correct_velocity(param->gpu, rkstep, fluid->psi, fluid->ux, fluid->uy),
                   "fluid_correct_velocity");
  return 0;
}

/* src/fluid/update_pressure.c:7-60 */
int fluid_update_pressure(const param_t *param, const parallel_t *parallel, const int rkstep, fluid_t *fluid){
  (void)parallel;
  common_check_gpu(snsg_fluid_update_pressure(param->gpu, rkstep, fluid->psi, fluid->p),
                   "fluid_update_pressure");
  return 0;
}

int fluid_download(const param_t *param, const fluid_t *fluid, double *ux, double *uy, double *p, double *psi){
  const int itot = param->itot;
  const int jsize = param->jtot;
  snsg_ctx *gpu = param->gpu;
  if(ux) common_check_gpu(snsg_download(gpu, ux, fluid->ux, UX_MEMSIZE), "fluid_download");
  if(uy) common_check_gpu(snsg_download(gpu, uy, fluid->uy, UY_MEMSIZE), "fluid_download");
  if(p) common_check_gpu(snsg_download(gpu, p, fluid->p, P_MEMSIZE), "fluid_download");
  if(psi) common_check_gpu(snsg_download(gpu, psi, fluid->psi, PSI_MEMSIZE), "fluid_download");
  return 0;
}
