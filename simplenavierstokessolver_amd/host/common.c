/* Allocation with fail-fast error handling (reference: src/common.c:8-21). */
#include <stdio.h>
#include <stdlib.h>
#include "common.h"
#include "sns_hip.h"

void *common_calloc(size_t count, size_t size){
  void *ptr = calloc(count ? count : 1, size ? size : 1);
  if(ptr == NULL){
    fprintf(stderr, "memory allocation error\n");
    abort();
  }
  return ptr;
}

void common_free(void *ptr){
  free(ptr);
}

void common_check_gpu(int rc, const char *what){
  if(rc != 0){
    fprintf(stderr, "GPU call failed in %s: %s\n", what, snsg_last_error());
    abort();
  }
}
