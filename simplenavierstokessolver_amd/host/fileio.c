/*
 * NPY I/O in the reference's on-disk format (src/fileio.c, src/simple_npyio.c):
 * NPY 1.0 header "{'descr':...,'fortran_order':False,'shape':(a,b,)}" padded
 * with spaces and a newline to a multiple of 64 bytes; 2-D fields hold the
 * interior rows only (the y-halo rows are skipped, src/fileio.c:216,253).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <errno.h>
#include <sys/stat.h>
#include "common.h"
#include "fileio.h"

static const char magic[] = "\x93NUMPY";

FILE *fileio_fopen(const char * restrict path, const char * restrict mode){
  FILE *stream = fopen(path, mode);
  if(stream == NULL){
    char msg[1024];
    snprintf(msg, sizeof msg, "%s:%d, %s", __FILE__, __LINE__, path);
    perror(msg);
  }
  return stream;
}

int fileio_fclose(FILE *stream){
  if(fclose(stream) != 0) perror("fileio_fclose");
  return 0;
}

/* continue if the directory exists (src/fileio.c:63-78) */
int fileio_mkdir_by_main_process(const char dirname[], const parallel_t *parallel){
  if(parallel->mpirank == 0){
    if(mkdir(dirname, 0777) != 0 && errno != EEXIST){
      perror(dirname);
    }
  }
  return 0;
}

size_t sns_npy_w_header(FILE *fp, const size_t ndim, const size_t *shape, const char dtype[]){
  char dict[512];
  char shp[256] = "(";
  for(size_t n=0; n<ndim; n++){
    char tmp[32];
    snprintf(tmp, sizeof tmp, "%zu,", shape[n]);
    strncat(shp, tmp, sizeof(shp)-strlen(shp)-1);
  }
  strncat(shp, ")", sizeof(shp)-strlen(shp)-1);
  snprintf(dict, sizeof dict, "{'descr':%s,'fortran_order':False,'shape':%s}", dtype, shp);
  const size_t ndict = strlen(dict);
  const size_t except = 6+1+1+2+ndict;
  size_t total = 0;
  while(total <= except) total += 64;
  const size_t npad = total-except;
  const uint16_t hlen = (uint16_t)(ndict+npad);
  const uint8_t ver[2] = {1, 0};
  const uint8_t hl[2] = {(uint8_t)(hlen & 0xff), (uint8_t)(hlen >> 8)};
  if(fwrite(magic, 1, 6, fp) != 6) return 0;
  if(fwrite(ver, 1, 2, fp) != 2) return 0;
  if(fwrite(hl, 1, 2, fp) != 2) return 0;
  if(fwrite(dict, 1, ndict, fp) != ndict) return 0;
  for(size_t k=0; k+1<npad; k++) fputc(' ', fp);
  fputc('\n', fp);
  return total;
}

/* parse the header; dtype receives the descr value including quotes */
size_t sns_npy_r_header(FILE *fp, size_t *ndim, size_t shape[8], char dtype[16]){
  uint8_t pre[10];
  if(fread(pre, 1, 8, fp) != 8) return 0;
  if(memcmp(pre, magic, 6) != 0) return 0;
  size_t hlen, hsize;
  if(pre[6] == 1){
    if(fread(pre+8, 1, 2, fp) != 2) return 0;
    hlen = (size_t)pre[8] | ((size_t)pre[9] << 8);
    hsize = 10+hlen;
  }else{
    uint8_t b[4];
    if(fread(b, 1, 4, fp) != 4) return 0;
    hlen = (size_t)b[0] | ((size_t)b[1] << 8) | ((size_t)b[2] << 16) | ((size_t)b[3] << 24);
    hsize = 12+hlen;
  }
  char *h = common_calloc(hlen+1, 1);
  if(fread(h, 1, hlen, fp) != hlen){ common_free(h); return 0; }
  /* strip spaces outside quotes */
  size_t w = 0;
  int inq = 0;
  for(size_t r=0; r<hlen; r++){
    if(h[r] == '\'' || h[r] == '"') inq = !inq;
    if(!inq && (h[r] == ' ' || h[r] == '\n')) continue;
    h[w++] = h[r];
  }
  h[w] = 0;
  const char *d = strstr(h, "'descr':");
  const char *f = strstr(h, "'fortran_order':");
  const char *s = strstr(h, "'shape':(");
  if(!d || !f || !s){ common_free(h); return 0; }
  d += strlen("'descr':");
  size_t k = 0;
  dtype[k++] = *d++;
  while(*d && *d != '\'' && k < 14) dtype[k++] = *d++;
  dtype[k++] = '\'';
  dtype[k] = 0;
  if(0 != strncmp(f+strlen("'fortran_order':"), "False", 5)){ common_free(h); return 0; }
  s += strlen("'shape':(");
  *ndim = 0;
  while(*s && *s != ')' && *ndim < 8){
    shape[(*ndim)++] = (size_t)strtoull(s, (char **)&s, 10);
    if(*s == ',') s++;
  }
  common_free(h);
  return hsize;
}

static char *npy_name(const char dirname[], const char dsetname[]){
  size_t n = strlen(dirname)+strlen(dsetname)+6;
  char *fname = common_calloc(n, 1);
  snprintf(fname, n, "%s/%s.npy", dirname, dsetname);
  return fname;
}

static int is_double(const char *dt){
  return 0 == strcmp(dt, "'float64'") || 0 == strcmp(dt, "'<f8'");
}

/* src/fileio.c:109-132 */
int fileio_r_0d_serial(const char dirname[], const char dsetname[], const size_t size, void *data){
  char *fname = npy_name(dirname, dsetname);
  FILE *fp = fileio_fopen(fname, "rb");
  int rc = 1;
  if(fp != NULL){
    size_t ndim, shape[8];
    char dtype[16];
    if(sns_npy_r_header(fp, &ndim, shape, dtype) != 0 && fread(data, size, 1, fp) == 1) rc = 0;
    fclose(fp);
  }
  common_free(fname);
  return rc;
}

/* src/fileio.c:134-148 */
int fileio_w_0d_serial(const char dirname[], const char dsetname[], const char dtype[], const size_t size, const void *data){
  char *fname = npy_name(dirname, dsetname);
  FILE *fp = fileio_fopen(fname, "wb");
  if(fp != NULL){
    if(sns_npy_w_header(fp, 0, NULL, dtype) != 0) fwrite(data, size, 1, fp);
    fclose(fp);
  }
  common_free(fname);
  return 0;
}

/* src/fileio.c:150-164 */
int fileio_w_1d_serial(const char dirname[], const char dsetname[], const char dtype[], const size_t size, const size_t nitems, const void *data){
  char *fname = npy_name(dirname, dsetname);
  FILE *fp = fileio_fopen(fname, "wb");
  if(fp != NULL){
    const size_t shape[1] = {nitems};
    if(sns_npy_w_header(fp, 1, shape, dtype) != 0) fwrite(data, size, nitems, fp);
    fclose(fp);
  }
  common_free(fname);
  return 0;
}

/* src/fileio.c:166-224: header check, then the interior rows into data+shape[1] */
static int r_2d(const char dirname[], const char dsetname[], const size_t shape[2], double *data){
  char *fname = npy_name(dirname, dsetname);
  FILE *fp = fileio_fopen(fname, "rb");
  int rc = 1;
  if(fp != NULL){
    size_t ndim, shp[8];
    char dtype[16];
    const size_t hs = sns_npy_r_header(fp, &ndim, shp, dtype);
    if(hs == 0){
      fprintf(stderr, "%s: npy header read failed\n", fname);
    }else if(ndim != 2 || shp[0] != shape[0] || shp[1] != shape[1]){
      printf("%s: shape should be (%zu, %zu)\n", fname, shape[0], shape[1]);
    }else if(!is_double(dtype)){
      printf("%s: dtype should be 'float64', not %s\n", fname, dtype);
    }else if(fread(data+shape[1], sizeof(double), shape[0]*shape[1], fp) == shape[0]*shape[1]){
      rc = 0;
    }
    fclose(fp);
  }
  common_free(fname);
  return rc;
}

/* src/fileio.c:226-261 */
static int w_2d(const char dirname[], const char dsetname[], const size_t shape[2], const double *data){
  char *fname = npy_name(dirname, dsetname);
  FILE *fp = fileio_fopen(fname, "wb");
  int rc = 1;
  if(fp != NULL){
    if(sns_npy_w_header(fp, 2, shape, NPYIO_DOUBLE) != 0 &&
       fwrite(data+shape[1], sizeof(double), shape[0]*shape[1], fp) == shape[0]*shape[1]) rc = 0;
    fclose(fp);
  }
  common_free(fname);
  return rc;
}

/* src/fileio.c:263-345 (one rank: offset 0, count = whole array) */
int fileio_r_ux_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, double *data){
  (void)parallel;
  const size_t shape[2] = {(size_t)param->jtot, (size_t)param->itot+1};
  return r_2d(dirname, dsetname, shape, data);
}
int fileio_w_ux_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, const double *data){
  (void)parallel;
  const size_t shape[2] = {(size_t)param->jtot, (size_t)param->itot+1};
  return w_2d(dirname, dsetname, shape, data);
}
int fileio_r_uy_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, double *data){
  (void)parallel;
  const size_t shape[2] = {(size_t)param->jtot, (size_t)param->itot+2};
  return r_2d(dirname, dsetname, shape, data);
}
int fileio_w_uy_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, const double *data){
  (void)parallel;
  const size_t shape[2] = {(size_t)param->jtot, (size_t)param->itot+2};
  return w_2d(dirname, dsetname, shape, data);
}
int fileio_r_p_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, double *data){
  return fileio_r_uy_like_parallel(dirname, dsetname, param, parallel, data);
}
int fileio_w_p_like_parallel(const char dirname[], const char dsetname[], const param_t *param, const parallel_t *parallel, const double *data){
  return fileio_w_uy_like_parallel(dirname, dsetname, param, parallel, data);
}
