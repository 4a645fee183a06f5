// sns_tridiag.h -- workgroup-cooperative tri-diagonal solver for gfx950.
//
// The reference solves every line with a sequential Thomas sweep
// (src/linalg.c:9-50).  Here one workgroup of NT threads solves one line of
// up to NT*M unknowns held in registers:
//
//  1. each thread eliminates its chunk of M consecutive rows so that every
//     unknown of the chunk is expressed through the chunk's first and last
//     unknowns (the partition scheme of Laszlo, Giles & Appleyard,
//     "Manycore algorithms for batch scalar and block tridiagonal solvers",
//     ACM TOMS 42(4), 2016);
//  2. the 2*NT chunk-boundary unknowns form a tri-diagonal system with unit
//     diagonal, solved by parallel cyclic reduction in LDS (log2(2*NT)
//     steps, two barriers each);
//  3. each thread recovers its interior unknowns.
//
// Several right-hand sides (NR) share one matrix (complex data = 2 RHS).
// Rows past the end of the line must be identity rows (a=0, b=1, c=0, d=0);
// a[0] of the first row and c of the last row must be zero.
//
// Rounding differs from the sequential sweep at the level of a few ulps for
// the diagonally dominant systems solved here.
#pragma once
#include <hip/hip_runtime.h>

namespace sns {

template <int NT, int NR>
struct TridiagLds {
  double af[NT], al[NT], bf[NT], bl[NT], cf[NT], cl[NT];
  double df[NR][NT], dl[NR][NT];
};

template <int NT, int M, int NR>
__device__ __forceinline__ void block_tridiag(double (&a)[M], double (&b)[M], double (&c)[M],
                                              double (&d)[NR][M], TridiagLds<NT, NR> &s) {
  static_assert(M >= 4, "chunk must hold at least four rows");
  const int t = threadIdx.x;
  // 1. chunk elimination, forward part: rows i >= 1 become
  //    a_i x_0 + x_i + c_i x_{i+1} = d_i
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const double r = 1.0 / b[i];
    a[i] *= r;
    c[i] *= r;
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][i] *= r;
  }
#pragma unroll
  for (int i = 2; i < M; i++) {
    const double r = 1.0 / (b[i] - a[i] * c[i - 1]);
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][i] = r * (d[k][i] - a[i] * d[k][i - 1]);
    a[i] = -r * a[i] * a[i - 1];
    c[i] = r * c[i];
  }
  // backward part: rows 1..M-2 become x_i = d_i - a_i x_0 - c_i x_{M-1};
  // row 0 becomes a_0 x_{-1} + x_0 + c_0 x_{M-1} = d_0
#pragma unroll
  for (int i = M - 3; i >= 1; i--) {
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][i] -= c[i] * d[k][i + 1];
    a[i] -= c[i] * a[i + 1];
    c[i] = -c[i] * c[i + 1];
  }
  {
    const double r = 1.0 / (1.0 - c[0] * a[1]);
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][0] = r * (d[k][0] - c[0] * d[k][1]);
    a[0] = r * a[0];
    c[0] = -r * c[0] * c[1];
  }
  // 2. reduced system: unknown 2t = first row of chunk t, 2t+1 = last row
  double ra[2] = {a[0], a[M - 1]};
  double rb[2] = {1.0, 1.0};
  double rc[2] = {c[0], c[M - 1]};
  double rd[2][NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    rd[0][k] = d[k][0];
    rd[1][k] = d[k][M - 1];
  }
  for (int stride = 1; stride < 2 * NT; stride <<= 1) {
    s.af[t] = ra[0]; s.al[t] = ra[1];
    s.bf[t] = rb[0]; s.bl[t] = rb[1];
    s.cf[t] = rc[0]; s.cl[t] = rc[1];
#pragma unroll
    for (int k = 0; k < NR; k++) { s.df[k][t] = rd[0][k]; s.dl[k][t] = rd[1][k]; }
    __syncthreads();
    double na[2], nb[2], nc[2], nd[2][NR];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int u = 2 * t + e;
      const int um = u - stride, up = u + stride;
      double am = 0.0, bm = 1.0, cm = 0.0, dm[NR];
      double ap = 0.0, bp = 1.0, cp = 0.0, dp[NR];
#pragma unroll
      for (int k = 0; k < NR; k++) { dm[k] = 0.0; dp[k] = 0.0; }
      if (um >= 0) {
        const int tt = um >> 1;
        if (um & 1) {
          am = s.al[tt]; bm = s.bl[tt]; cm = s.cl[tt];
#pragma unroll
          for (int k = 0; k < NR; k++) dm[k] = s.dl[k][tt];
        } else {
          am = s.af[tt]; bm = s.bf[tt]; cm = s.cf[tt];
#pragma unroll
          for (int k = 0; k < NR; k++) dm[k] = s.df[k][tt];
        }
      }
      if (up < 2 * NT) {
        const int tt = up >> 1;
        if (up & 1) {
          ap = s.al[tt]; bp = s.bl[tt]; cp = s.cl[tt];
#pragma unroll
          for (int k = 0; k < NR; k++) dp[k] = s.dl[k][tt];
        } else {
          ap = s.af[tt]; bp = s.bf[tt]; cp = s.cf[tt];
#pragma unroll
          for (int k = 0; k < NR; k++) dp[k] = s.df[k][tt];
        }
      }
      const double k1 = ra[e] / bm;
      const double k2 = rc[e] / bp;
      na[e] = -am * k1;
      nc[e] = -cp * k2;
      nb[e] = rb[e] - cm * k1 - ap * k2;
#pragma unroll
      for (int k = 0; k < NR; k++) nd[e][k] = rd[e][k] - dm[k] * k1 - dp[k] * k2;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 2; e++) {
      ra[e] = na[e]; rb[e] = nb[e]; rc[e] = nc[e];
#pragma unroll
      for (int k = 0; k < NR; k++) rd[e][k] = nd[e][k];
    }
  }
  // 3. interior recovery
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const double x0 = rd[0][k] / rb[0];
    const double x1 = rd[1][k] / rb[1];
#pragma unroll
    for (int i = 1; i < M - 1; i++) d[k][i] = d[k][i] - a[i] * x0 - c[i] * x1;
    d[k][0] = x0;
    d[k][M - 1] = x1;
  }
}

}  // namespace sns
