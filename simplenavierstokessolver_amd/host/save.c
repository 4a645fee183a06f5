/* Checkpoint writer (reference: src/save.c:13-68); fields are downloaded first. */
#include <stdio.h>
#include <string.h>
#include "common.h"
#include "param.h"
#include "parallel.h"
#include "fluid.h"
#include "temperature.h"
#include "save.h"
#include "fileio.h"

int save(param_t *param, const parallel_t *parallel, const fluid_t *fluid, const temperature_t *temperature){
  const int itot = param->itot;
  const int jsize = param->jtot;
  char dirname[64];
  snprintf(dirname, sizeof dirname, "%s/step%010d", FILEIO_SAVE, param->step);
  fileio_mkdir_by_main_process(dirname, parallel);
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
  double *ux = common_calloc(1, UX_MEMSIZE);
  double *uy = common_calloc(1, UY_MEMSIZE);
  double *p  = common_calloc(1, P_MEMSIZE);
  fluid_download(param, fluid, ux, uy, p, NULL);
  fileio_w_ux_like_parallel(dirname, "ux", param, parallel, ux);
  fileio_w_uy_like_parallel(dirname, "uy", param, parallel, uy);
  fileio_w_p_like_parallel (dirname, "p",  param, parallel, p );
  if(param->with_temperature){
    double *temp = common_calloc(1, TEMP_MEMSIZE);
    temperature_download(param, temperature, temp);
    fileio_w_p_like_parallel(dirname, "temp", param, parallel, temp);
    common_free(temp);
  }
  common_free(ux);
  common_free(uy);
  common_free(p);
  return 0;
}
