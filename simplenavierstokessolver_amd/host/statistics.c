/*
 * Time-averaged statistics (reference: src/statistics/init.c, collect.c,
 * output.c, finalise.c).  Accumulation runs on the device
 * (snsg_statistics_collect); the output downloads the sums and writes them
 * with the sample count and the run parameters.
 */
#include <stdio.h>
#include <string.h>
#include "common.h"
#include "param.h"
#include "parallel.h"
#include "fluid.h"
#include "temperature.h"
#include "statistics.h"
#include "fileio.h"

static double *dev_zeros(const param_t *param, size_t bytes){
  void *p = NULL;
  common_check_gpu(snsg_malloc(param->gpu, bytes, &p), "statistics_init");
  return (double *)p;
}

/* src/statistics/init.c */
statistics_t *statistics_init(const param_t *param, const parallel_t *parallel){
  (void)parallel;
  const size_t nu = sizeof(double)*(size_t)(param->itot+1)*(param->jtot+2);
  const size_t nc = sizeof(double)*(size_t)(param->itot+2)*(param->jtot+2);
  statistics_t *s = common_calloc(1, sizeof(statistics_t));
  s->ux1 = dev_zeros(param, nu);
  s->ux2 = dev_zeros(param, nu);
  s->uy1 = dev_zeros(param, nc);
  s->uy2 = dev_zeros(param, nc);
  s->temp1 = dev_zeros(param, nc);
  s->temp2 = dev_zeros(param, nc);
  s->num = 0;
  return s;
}

/* src/statistics/collect.c:69-75 */
int statistics_collect(param_t *param, const parallel_t *parallel, const fluid_t *fluid, const temperature_t *temperature, statistics_t *statistics){
  (void)parallel;
  common_check_gpu(snsg_statistics_collect(param->gpu, fluid->ux, fluid->uy,
                                           param->with_temperature ? temperature->temp : NULL,
                                           statistics->ux1, statistics->ux2,
                                           statistics->uy1, statistics->uy2,
                                           statistics->temp1, statistics->temp2),
                   "statistics_collect");
  statistics->num += 1;
  return 0;
}

static void save_param(const char dirname[], const param_t *param, const int num){
  fileio_w_0d_serial(dirname, "num",  NPYIO_INT,    sizeof(int),    &num);
  fileio_w_0d_serial(dirname, "step", NPYIO_INT,    sizeof(int),    &(param->step));
  fileio_w_0d_serial(dirname, "itot", NPYIO_INT,    sizeof(int),    &(param->itot));
  fileio_w_0d_serial(dirname, "jtot", NPYIO_INT,    sizeof(int),    &(param->jtot));
  fileio_w_0d_serial(dirname, "time", NPYIO_DOUBLE, sizeof(double), &(param->time));
  fileio_w_0d_serial(dirname, "lx",   NPYIO_DOUBLE, sizeof(double), &(param->lx  ));
  fileio_w_0d_serial(dirname, "ly",   NPYIO_DOUBLE, sizeof(double), &(param->ly  ));
  fileio_w_0d_serial(dirname, "Ra",   NPYIO_DOUBLE, sizeof(double), &(param->Ra  ));
  fileio_w_0d_serial(dirname, "Pr",   NPYIO_DOUBLE, sizeof(double), &(param->Pr  ));
  fileio_w_1d_serial(dirname, "xf",   NPYIO_DOUBLE, sizeof(double), param->itot+1, param->xf);
  fileio_w_1d_serial(dirname, "xc",   NPYIO_DOUBLE, sizeof(double), param->itot+2, param->xc);
  fileio_w_1d_serial(dirname, "yf",   NPYIO_DOUBLE, sizeof(double), param->jtot+1, param->yf);
  fileio_w_1d_serial(dirname, "yc",   NPYIO_DOUBLE, sizeof(double), param->jtot  , param->yc);
}

static void save_field(const param_t *param, const parallel_t *parallel, const char dirname[],
                       const char name[], const double *dev, int ux_like){
  const size_t n = (size_t)(param->itot+(ux_like ? 1 : 2))*(param->jtot+2);
  double *host = common_calloc(n, sizeof(double));
  common_check_gpu(snsg_download(param->gpu, host, dev, n*sizeof(double)), "statistics_output");
  if(ux_like) fileio_w_ux_like_parallel(dirname, name, param, parallel, host);
  else        fileio_w_uy_like_parallel(dirname, name, param, parallel, host);
  common_free(host);
}

/* src/statistics/output.c */
int statistics_output(const param_t *param, const parallel_t *parallel, const statistics_t *statistics){
  char dirname[64];
  snprintf(dirname, sizeof dirname, "%s/step%010d", FILEIO_STAT, param->step);
  fileio_mkdir_by_main_process(dirname, parallel);
  save_param(dirname, param, statistics->num);
  save_field(param, parallel, dirname, "ux1", statistics->ux1, 1);
  save_field(param, parallel, dirname, "uy1", statistics->uy1, 0);
  save_field(param, parallel, dirname, "ux2", statistics->ux2, 1);
  save_field(param, parallel, dirname, "uy2", statistics->uy2, 0);
  if(param->with_temperature){
    save_field(param, parallel, dirname, "temp1", statistics->temp1, 0);
    save_field(param, parallel, dirname, "temp2", statistics->temp2, 0);
  }
  return 0;
}

/* src/statistics/finalise.c */
int statistics_finalise(statistics_t *statistics){
  snsg_ctx *gpu = param_current_gpu();
  snsg_free(gpu, statistics->ux1);
  snsg_free(gpu, statistics->ux2);
  snsg_free(gpu, statistics->uy1);
  snsg_free(gpu, statistics->uy2);
  snsg_free(gpu, statistics->temp1);
  snsg_free(gpu, statistics->temp2);
  common_free(statistics);
  return 0;
}
