/*
 * Domain decomposition (reference: include/parallel.h).  The single-GPU build
 * runs one rank: mpisize == 1, mpirank == 0, and the y-slab arithmetic below
 * is the reference's (src/parallel/others.c:6-19).  The reference's MPI
 * collectives become device operations of the current GPU instance:
 * parallel_transpose is an LDS-tiled transpose of device buffers and the
 * halo exchanges copy one device row to another.  MPI types in the
 * reference's signatures are replaced by element sizes in bytes.
 */
#ifndef PARALLEL_H
#define PARALLEL_H

#include <stddef.h>
#include "structure.h"

struct parallel_t_ {
  int mpisize, mpirank;
  int ymrank, yprank;
};

extern parallel_t *parallel_init(void);
extern int parallel_finalise(parallel_t *parallel);
extern int parallel_get_size(const int num, const int size, const int rank);
extern int parallel_get_offset(const int num, const int size, const int rank);
/* wall clock in seconds (reference: MPI_Wtime reduced over ranks) */
extern double parallel_get_wtime(void);

/* global transpose of device buffers: recvbuf[i*g_jsize+j] = sendbuf[j*g_isize+i] */
extern int parallel_transpose(const int g_isize, const int g_jsize, const size_t dtypesize,
                              const void *sendbuf, void *recvbuf);

/* periodic halo exchange on one rank: copies cnt elements of dtypesize bytes
 * from sendbuf to recvbuf (device buffers) */
extern int parallel_update_halo_ym(const parallel_t *parallel, const int cnt, const size_t dtypesize,
                                   const void *sendbuf, void *recvbuf);
extern int parallel_update_halo_yp(const parallel_t *parallel, const int cnt, const size_t dtypesize,
                                   const void *sendbuf, void *recvbuf);

#endif
