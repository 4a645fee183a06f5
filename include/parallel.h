/*
 * Domain decomposition (reference: include/parallel.h:1-28).  The prototypes
 * are the reference's, MPI types included, so that reference callers
 * (src/main.c:18,56, src/param/estimate_cost.c:19,34) compile unchanged.
 * The single-GPU build runs one rank: mpisize == 1, mpirank == 0, and the
 * y-slab arithmetic is the reference's (src/parallel/others.c:6-19).  The
 * MPI arguments are accepted and need no communicator:
 *   - parallel_get_wtime: the min / max over one rank is the local clock;
 *   - parallel_transpose: an LDS-tiled transpose of device buffers with
 *     elements of dtype bytes (mpi_dtype is not needed for the size);
 *   - parallel_update_halo_*: a device row copy; the element size comes
 *     from mpi_dtype (MPI_DOUBLE, MPI_FLOAT, MPI_INT, MPI_C_DOUBLE_COMPLEX,
 *     MPI_DOUBLE_COMPLEX), other types abort with a message.
 * Only <mpi.h> is needed to compile against this header; the library does
 * not call into MPI, so it runs with or without MPI_Init.
 */
#ifndef PARALLEL_H
#define PARALLEL_H

#include <stddef.h>
#include <mpi.h>

#include "structure.h"

struct parallel_t_ {
  int mpisize, mpirank;
  int ymrank, yprank;
};

extern parallel_t *parallel_init(void);
extern int parallel_finalise(parallel_t *parallel);
extern int parallel_get_size(const int num, const int size, const int rank);
extern int parallel_get_offset(const int num, const int size, const int rank);
extern double parallel_get_wtime(const MPI_Op op);

/* global transpose of device buffers: recvbuf[i*g_jsize+j] = sendbuf[j*g_isize+i] */
extern int parallel_transpose(const int g_isize, const int g_jsize, const size_t dtype, const MPI_Datatype mpi_dtype, const void *sendbuf, void *recvbuf);

/* periodic halo exchange on one rank: copies cnt elements from sendbuf to
 * recvbuf (device buffers) */
extern int parallel_update_halo_ym(const parallel_t *parallel, const int cnt, const MPI_Datatype dtype, const void *sendbuf, void *recvbuf);
extern int parallel_update_halo_yp(const parallel_t *parallel, const int cnt, const MPI_Datatype dtype, const void *sendbuf, void *recvbuf);

#endif
