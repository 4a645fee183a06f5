/*
 * One-rank decomposition and the device replacements of the reference's MPI
 * calls (src/parallel/{init,others,transpose,halo}.c).  The MPI arguments of
 * the reference's prototypes are accepted; nothing here calls into MPI.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include "common.h"
#include "param.h"
#include "parallel.h"
#include "sns_hip.h"

/* src/parallel/init.c:7-27 with mpisize = 1: both neighbours are rank 0 */
parallel_t *parallel_init(void){
  parallel_t *parallel = common_calloc(1, sizeof(parallel_t));
  parallel->mpisize = 1;
  parallel->mpirank = 0;
  parallel->ymrank = 0;
  parallel->yprank = 0;
  return parallel;
}

int parallel_finalise(parallel_t *parallel){
  common_free(parallel);
  return 0;
}

/* src/parallel/others.c:6-10 */
int parallel_get_size(const int num, const int size, const int rank){
  return (num+rank)/size;
}

/* src/parallel/others.c:12-19 */
int parallel_get_offset(const int num, const int size, const int rank){
  int offset = 0;
  for(int r=0; r<rank; r++){
    offset += parallel_get_size(num, size, r);
  }
  return offset;
}

/* src/parallel/others.c: MPI_Wtime reduced with op over the ranks; over one
 * rank the reduction is the local value */
double parallel_get_wtime(const MPI_Op op){
  (void)op;
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec+1.e-9*(double)ts.tv_nsec;
}

/* src/parallel/transpose.c:6-85 on one rank, as a device transpose */
int parallel_transpose(const int g_isize, const int g_jsize, const size_t dtype, const MPI_Datatype mpi_dtype, const void *sendbuf, void *recvbuf){
  (void)mpi_dtype;
  int rc = snsg_transpose(param_current_gpu(), g_isize, g_jsize, dtype, sendbuf, recvbuf);
  common_check_gpu(rc, "parallel_transpose");
  return rc;
}

/* bytes of one element of an MPI type the halo exchange may be given */
static size_t mpi_type_bytes(const MPI_Datatype dtype){
  if(dtype == MPI_DOUBLE) return sizeof(double);
  if(dtype == MPI_FLOAT) return sizeof(float);
  if(dtype == MPI_INT) return sizeof(int);
  if(dtype == MPI_C_DOUBLE_COMPLEX || dtype == MPI_DOUBLE_COMPLEX) return 2*sizeof(double);
  fprintf(stderr, "parallel_update_halo: unsupported MPI datatype %ld\n", (long)dtype);
  abort();
}

/* src/parallel/halo.c:6-22: with one rank the neighbour is the rank itself */
static int halo_copy(const int cnt, const MPI_Datatype dtype, const void *sendbuf, void *recvbuf){
  snsg_ctx *gpu = param_current_gpu();
  /* device-to-device copy of one row through a 1 x cnt transpose */
  int rc = snsg_transpose(gpu, cnt, 1, mpi_type_bytes(dtype), sendbuf, recvbuf);
  common_check_gpu(rc, "parallel_update_halo");
  return rc;
}

int parallel_update_halo_ym(const parallel_t *parallel, const int cnt, const MPI_Datatype dtype, const void *sendbuf, void *recvbuf){
  (void)parallel;
  return halo_copy(cnt, dtype, sendbuf, recvbuf);
}

int parallel_update_halo_yp(const parallel_t *parallel, const int cnt, const MPI_Datatype dtype, const void *sendbuf, void *recvbuf){
  (void)parallel;
  return halo_copy(cnt, dtype, sendbuf, recvbuf);
}
