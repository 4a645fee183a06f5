/* Forward declarations of the solver's structures (reference: include/structure.h). */
#ifndef STRUCTURE_H
#define STRUCTURE_H

typedef struct param_t_       param_t;
typedef struct parallel_t_    parallel_t;
typedef struct fluid_t_       fluid_t;
typedef struct temperature_t_ temperature_t;

#endif
