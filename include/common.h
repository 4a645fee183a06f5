/* Allocation helpers with fail-fast error handling (reference: include/common.h,
 * src/common.c:8-21, where a failed calloc aborts the run). */
#ifndef COMMON_H
#define COMMON_H

#include <stdlib.h>

#if !defined(M_PI)
#define M_PI 3.1415926535897932
#endif

extern void *common_calloc(size_t count, size_t size);
extern void common_free(void *ptr);
/* abort the run with a message when a device call failed (rc != 0) */
extern void common_check_gpu(int rc, const char *what);

#endif
