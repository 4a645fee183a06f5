/*
 * NPY file I/O in the reference's format (reference: include/fileio.h,
 * src/fileio.c, src/simple_npyio.c).  2-D fields are stored (jtot, itot+1)
 * for ux-like and (jtot, itot+2) for uy/p-like arrays, C order, float64,
 * without the y-halo rows; the header is NPY 1.0 padded to 64 bytes.
 * The *_parallel functions take HOST arrays of the halo'd layout.
 */
#ifndef FILEIO_H
#define FILEIO_H

#include <stdio.h>
#include <stddef.h>
#include "param.h"
#include "parallel.h"

#define FILEIO_SAVE "output/save"
#define FILEIO_LOG  "output/log"
#define FILEIO_STAT "output/stat"

extern FILE *fileio_fopen(const char * restrict path, const char * restrict mode);
extern int fileio_fclose(FILE *stream);
extern int fileio_mkdir_by_main_process(const char dirname[], const parallel_t *parallel);

#define NPYIO_INT    "'int32'"
#define NPYIO_DOUBLE "'float64'"

extern int fileio_w_0d_serial(const char dirname[], const char dsetname[], const char dtype[], const size_t size, const void *data);
extern int fileio_r_0d_serial(const char dirname[], const char dsetname[],                     const size_t size,       void *data);
extern int fileio_w_1d_serial(const char dirname[], const char dsetname[], const char dtype[], const size_t size, const size_t nitems, const void *data);
extern int fileio_r_ux_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel,       double *data);
extern int fileio_w_ux_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, const double *data);
extern int fileio_r_uy_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel,       double *data);
extern int fileio_w_uy_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, const double *data);
extern int fileio_r_p_like_parallel (const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel,       double *data);
extern int fileio_w_p_like_parallel (const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, const double *data);

/* NPY header writer / reader: returns the header size in bytes, 0 on failure */
extern size_t sns_npy_w_header(FILE *fp, const size_t ndim, const size_t *shape, const char dtype[]);
extern size_t sns_npy_r_header(FILE *fp, size_t *ndim, size_t shape[8], char dtype[16]);

#endif
