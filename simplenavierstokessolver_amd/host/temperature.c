/*
 * Temperature field on the GPU (reference: src/temperature/init.c,
 * finalise.c, force.c, update.c, boundary_conditions.c).
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
  common_check_gpu(snsg_malloc(param->gpu, bytes, &p), "temperature_init");
  return (double *)p;
}

static uint64_t splitmix64(uint64_t x){
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30))*0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27))*0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

/* Initial temperature (replaces the rand() field of src/temperature/init.c:39-47,
 * which depends on the rank count): T(i,j) = base(i) + A*(2U-1), U in [0,1)
 * from the top 53 bits of splitmix64(seed<<40 ^ ((j-1)*itot + (i-1))); base is
 * the conduction profile (TEMP_XP-TEMP_XM)*XC(i)+TEMP_XM when ic_linear, else 0. */
int sns_initial_temperature(const int itot, const int jtot, const double *xc, const uint64_t seed,
                            const double amplitude, const int linear, double *temp){
  for(int j=1; j<=jtot; j++){
    for(int i=1; i<=itot; i++){
      uint64_t key = (seed << 40) ^ ((uint64_t)(j-1)*(uint64_t)itot + (uint64_t)(i-1));
      double u = (double)(splitmix64(key) >> 11)*0x1.0p-53;
      double base = linear ? (TEMP_XP-TEMP_XM)*XC(i)+TEMP_XM : 0.;
      TEMP(i, j) = base+amplitude*(2.*u-1.);
    }
  }
  return 0;
}

static void initial_temperature(const param_t *param, double *temp){
  sns_initial_temperature(param->itot, param->jtot, param->xc, param->ic_seed, param->ic_amplitude,
                          param->ic_linear, temp);
}

/* src/temperature/init.c:8-62 */
temperature_t *temperature_init(const param_t *param, const parallel_t *parallel){
  const int itot = param->itot;
  const int jsize = parallel_get_size(param->jtot, parallel->mpisize, parallel->mpirank);
  temperature_t *temperature = common_calloc(1, sizeof(temperature_t));
  temperature->temp       = dev_alloc(param, TEMP_MEMSIZE);
  temperature->tempforcex = dev_alloc(param, TEMPFORCEX_MEMSIZE);
  temperature->srctempa   = dev_alloc(param, SRCTEMPA_MEMSIZE);
  temperature->srctempb   = dev_alloc(param, SRCTEMPA_MEMSIZE);
  temperature->srctempg   = dev_alloc(param, SRCTEMPA_MEMSIZE);
  double *host = common_calloc(1, TEMP_MEMSIZE);
  if(param->load_flow_field){
    if(fileio_r_p_like_parallel(param->dirname_restart, "temp", param, parallel, host) != 0){
      fprintf(stderr, "restart: cannot load %s/temp.npy\n", param->dirname_restart);
      abort();
    }
  }else{
    initial_temperature(param, host);
  }
  common_check_gpu(snsg_upload(param->gpu, temperature->temp, host, TEMP_MEMSIZE), "temperature_init");
  common_free(host);
  temperature_update_boundaries_temp(param, parallel, temperature->temp);
  return temperature;
}

/* src/temperature/finalise.c */
int temperature_finalise(temperature_t *temperature){
  snsg_ctx *gpu = param_current_gpu();
  snsg_free(gpu, temperature->temp);
  snsg_free(gpu, temperature->tempforcex);
  snsg_free(gpu, temperature->srctempa);
  snsg_free(gpu, temperature->srctempb);
  snsg_free(gpu, temperature->srctempg);
  common_free(temperature);
  return 0;
}

/* src/temperature/force.c:6-32 */
int temperature_compute_force(const param_t *param, const parallel_t *parallel, temperature_t *temperature){
  (void)parallel;
  common_check_gpu(snsg_temperature_compute_force(param->gpu, temperature->temp, temperature->tempforcex),
                   "temperature_compute_force");
  return 0;
}

/* src/temperature/update.c:233-239 (SRCTEMPB <- SRCTEMPA at line 31 is a pointer exchange) */
int temperature_update_temp(const param_t *param, const parallel_t *parallel, const int rkstep, const fluid_t *fluid, temperature_t *temperature){
  (void)parallel;
  double *t = temperature->srctempb;
  temperature->srctempb = temperature->srctempa;
  temperature->srctempa = t;
  int rc = snsg_temperature_update_temp(param->gpu, rkstep, fluid->ux, fluid->uy, temperature->temp,
      temperature->srctempa, temperature->srctempb, temperature->srctempg);
  common_check_gpu(rc, "temperature_update_temp");
  return 0;
}

/* src/temperature/boundary_conditions.c:6-21 */
int temperature_update_boundaries_temp(const param_t *param, const parallel_t *parallel, double *temp){
  (void)parallel;
  common_check_gpu(snsg_temperature_update_boundaries_temp(param->gpu, temp), "temperature_update_boundaries_temp");
  return 0;
}

int temperature_download(const param_t *param, const temperature_t *temperature, double *temp){
  const int itot = param->itot;
  const int jsize = param->jtot;
  common_check_gpu(snsg_download(param->gpu, temp, temperature->temp, TEMP_MEMSIZE), "temperature_download");
  return 0;
}
