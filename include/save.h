/* Checkpoint writer (reference: include/save.h, src/save.c). */
#ifndef SAVE_H
#define SAVE_H
#include "structure.h"
extern int save(param_t *param, const parallel_t *parallel, const fluid_t *fluid, const temperature_t *temperature);
#endif
