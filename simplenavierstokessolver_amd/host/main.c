/*
 * Driver (reference: src/main.c).  Same time loop: decide dt, integrate three
 * RK substages through the reference-named functions, log / save / collect
 * statistics on their schedules, stop at timemax or the wall-time limit,
 * save the fields and the statistics at the end.
 */
#include <stdio.h>
#include <sys/stat.h>
#include "param.h"
#include "parallel.h"
#include "fluid.h"
#include "temperature.h"
#include "save.h"
#include "logging.h"
#include "statistics.h"
#include "fileio.h"

/* src/main.c:79-101 */
static int integrate(const param_t *param, const parallel_t *parallel, fluid_t *fluid, temperature_t *temperature){
  for(int rkstep=0; rkstep<RKSTEPMAX; rkstep++){
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
  }
  return 0;
}

int main(void){
  double wtimes[2] = {0., 0.};
  wtimes[0] = parallel_get_wtime(MPI_MIN);
  mkdir("output", 0777);
  mkdir(FILEIO_SAVE, 0777);
  mkdir(FILEIO_LOG, 0777);
  mkdir(FILEIO_STAT, 0777);
  param_t       *param       = param_init();
  parallel_t    *parallel    = parallel_init();
  fluid_t       *fluid       = fluid_init(param, parallel);
  temperature_t *temperature = temperature_init(param, parallel);
  statistics_t  *statistics  = statistics_init(param, parallel);
  param_estimate_cost(param, parallel);
  for(;;){
    param_decide_dt(param, parallel, fluid);
    integrate(param, parallel, fluid, temperature);
    param->step += 1;
    param->time += param->dt;
    if(param->log.next < param->time){
      logging(param, parallel, fluid, temperature);
      param->log.next += param->log.rate;
    }
    if(param->save.next < param->time){
      save(param, parallel, fluid, temperature);
      param->save.next += param->save.rate;
    }
    if(param->stat.next < param->time){
      statistics_collect(param, parallel, fluid, temperature, statistics);
      param->stat.next += param->stat.rate;
    }
    if(param->time > param->timemax){
      break;
    }
    wtimes[1] = parallel_get_wtime(MPI_MAX);
    if(wtimes[1]-wtimes[0] > param->wtimemax){
      break;
    }
  }
  printf("elapsed: %.2f [s]\n", wtimes[1]-wtimes[0]);
  save(param, parallel, fluid, temperature);
  statistics_output(param, parallel, statistics);
  statistics_finalise(statistics);
  temperature_finalise(temperature);
  fluid_finalise(fluid);
  parallel_finalise(parallel);
  param_finalise(param);
  return 0;
}
