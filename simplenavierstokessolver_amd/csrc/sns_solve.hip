// sns_solve.hip -- line solvers, transposes and the Poisson chain (gfx950).
#include "sns_launch.h"
#include "sns_tridiag.h"
#include "sns_fast.h"
#include "sns_fused.h"
#include "sns_lu.h"

namespace sns {

// ------------------------------------------------------------ helpers

template <int KIND>
__device__ __forceinline__ double cn_prefactor(const Geo &g, double gamma, double dt) {
  // T:     (gamma*dt)/(2.*sqrt(Pr)*sqrt(Ra))   src/temperature/update.c:142
  // ux,uy: (gamma*dt*sqrt(Pr))/(2.*sqrt(Ra))   src/fluid/update_velocity.c:254,399
  return KIND == KIND_TEMP ? (gamma * dt) / g.two_sqrtPr_sqrtRa
                           : (gamma * dt * g.sqrtPr) / g.two_sqrtRa;
}

template <int KIND>
__device__ __forceinline__ void store_field_bc(const Geo &g, double *f, int i, int j, double v);

// boundary-complete stores (same rules as sns_stencil.hip)
template <>
__device__ __forceinline__ void store_field_bc<KIND_UX>(const Geo &g, double *f, int i, int j,
                                                        double v) {
  f[ux_idx(g, i, j)] = v;
  if (j == 1) f[ux_idx(g, i, g.jtot + 1)] = v;
  if (j == g.jtot) f[ux_idx(g, i, 0)] = v;
  if (i == 2) {
    f[ux_idx(g, 1, j)] = 0.;
    if (j == 1) f[ux_idx(g, 1, g.jtot + 1)] = 0.;
    if (j == g.jtot) f[ux_idx(g, 1, 0)] = 0.;
  }
  if (i == g.itot) {
    f[ux_idx(g, g.itot + 1, j)] = 0.;
    if (j == 1) f[ux_idx(g, g.itot + 1, g.jtot + 1)] = 0.;
    if (j == g.jtot) f[ux_idx(g, g.itot + 1, 0)] = 0.;
  }
}
template <int WALLKIND>
__device__ __forceinline__ void store_c_bc(const Geo &g, double *f, int i, int j, double v) {
  f[c_idx(g, i, j)] = v;
  if (j == 1) f[c_idx(g, i, g.jtot + 1)] = v;
  if (j == g.jtot) f[c_idx(g, i, 0)] = v;
  for (int side = 0; side < 2; side++) {
    const bool hit = side == 0 ? (i == 1) : (i == g.itot);
    if (!hit) continue;
    const int iw = side == 0 ? 0 : g.itot + 1;
    const double w = WALLKIND == 0 ? 0. : (WALLKIND == 1 ? v : (side == 0 ? +0.5 : -0.5));
    f[c_idx(g, iw, j)] = w;
    if (WALLKIND != 2) {
      if (j == 1) f[c_idx(g, iw, g.jtot + 1)] = w;
      if (j == g.jtot) f[c_idx(g, iw, 0)] = w;
    }
  }
}
template <>
__device__ __forceinline__ void store_field_bc<KIND_UY>(const Geo &g, double *f, int i, int j,
                                                        double v) {
  store_c_bc<0>(g, f, i, j, v);
}
template <>
__device__ __forceinline__ void store_field_bc<KIND_TEMP>(const Geo &g, double *f, int i, int j,
                                                          double v) {
  store_c_bc<2>(g, f, i, j, v);
}

template <int KIND>
__device__ __forceinline__ double field_at(const Geo &g, const double *f, int i, int j) {
  return KIND == KIND_UX ? f[ux_idx(g, i, j)] : f[c_idx(g, i, j)];
}

// ------------------------------------------------- Crank-Nicolson in x

// One workgroup per row j (src/temperature/update.c:123-167,
// src/fluid/update_velocity.c:230-280,380-424): the increment, from the RK
// sources or from qin, is loaded coalesced and staged in LDS in the chunk
// order of the substitutions, solved with the LU tables of k_lu_factor (the
// fused predictor's solve, so the two paths return the same bits), and added
// to the field with boundary values, or stored to qout.
template <int KIND, int NT, int M>
__global__ void __launch_bounds__(NT) k_cnx_lu(CnxLuArgs a) {
  extern __shared__ double cn_lds[];
  double *row = cn_lds;                 // NT*(M+1): chunk t at t*(M+1)
  double *scr = cn_lds + NT * (M + 1);  // scan scratch of one system
  const Geo &g = a.g;
  const int j = blockIdx.x + 1;
  const int n = KIND == KIND_UX ? g.itot - 1 : g.itot;
  const int i0 = KIND == KIND_UX ? 2 : 1;
  const double dt = *a.dtp;
  const RK rk = a.rk;
  const size_t rowoff = (size_t)(j - 1) * n;
  for (int r = threadIdx.x; r < n; r += NT) {
    const size_t s = rowoff + r;
    row[r + r / M] = a.qin ? a.qin[s]
                           : rk_incr(rk.alpha, rk.beta, rk.gamma, dt, a.srca[s], a.srcb[s],
                                     a.srcg[s]);
  }
  __syncthreads();
  double d[1][M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    d[0][k] = r < n ? row[r + r / M] : 0.;
  }
  const LuTabC<M> tab = {a.lu, a.lu + a.ntm, a.lu + 2 * a.ntm, (int)a.lu_hdr[0], n,
                         a.lu_hdr[1], a.lu_hdr[2], a.lu_hdr[3]};
  lu_solve_shared<NT, M, 1>(d, tab, scr);
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r < n) row[r + r / M] = d[0][k];
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += NT) {
    const double x = row[r + r / M];
    if (a.qout) {
      a.qout[rowoff + r] = x;
    } else {
      const int i = r + i0;
      store_field_bc<KIND>(g, a.field, i, j, field_at<KIND>(g, a.field, i, j) + x);
    }
  }
}

template <int KIND, int NT>
static void cnx_lu_launch(hipStream_t s, const CnxLuArgs &a) {
  constexpr int M = 8;
  const size_t lds = sizeof(double) * (NT * (M + 1) + 4 * (NT / 64));
  static std::atomic<unsigned> attr{0};
  lds_limit_once((const void *)k_cnx_lu<KIND, NT, M>, lds, attr);
  hipLaunchKernelGGL((k_cnx_lu<KIND, NT, M>), dim3(a.g.jtot), dim3(NT), lds, s, a);
}

template <int KIND>
static void cnx_lu_dispatch(hipStream_t s, const CnxLuArgs &a) {
  switch (lu_row_threads(a.g.itot)) {
    case 64: cnx_lu_launch<KIND, 64>(s, a); break;
    case 128: cnx_lu_launch<KIND, 128>(s, a); break;
    case 256: cnx_lu_launch<KIND, 256>(s, a); break;
    case 512: cnx_lu_launch<KIND, 512>(s, a); break;
    default: cnx_lu_launch<KIND, 1024>(s, a); break;
  }
}

void launch_cnx_lu(hipStream_t s, int kind, const CnxLuArgs &a) {
  if (kind == KIND_UX) cnx_lu_dispatch<KIND_UX>(s, a);
  else if (kind == KIND_UY) cnx_lu_dispatch<KIND_UY>(s, a);
  else cnx_lu_dispatch<KIND_TEMP>(s, a);
}

// ------------------------------------------------- Crank-Nicolson in y

// Periodic solve of one y-aligned row of length n = jtot in place, by the
// reference's superposition of two (n-1) systems (src/linalg.c:52-108): the
// row's own right-hand side and the matrix-only right-hand side
// [-l, 0, .., 0, -u] share one elimination (NR = 2).
// Coefficients: src/temperature/update.c:188-197, src/fluid/update_velocity.c:302-310
template <int KIND, int NT, int M, class S>
__global__ void __launch_bounds__(NT) k_cny_rows(S src) {
  const YArgs ya = src.get();
  const Geo &g = ya.g;
  const RK rk = ya.rk;
  __shared__ TridiagLds<NT, 2> lds;
  __shared__ double ends[4];
  const int n = g.jtot;
  double *q = ya.qy + (size_t)blockIdx.x * n;
  const double dt = *ya.dtp;
  const double pf = cn_prefactor<KIND>(g, rk.gamma, dt);
  const double coef_ym = 1. / g.dy / g.dy;
  const double coef_yp = 1. / g.dy / g.dy;
  const double coef_y0 = -coef_ym - coef_yp;
  const double l = -pf * coef_ym;
  const double u = -pf * coef_yp;
  const double cc = 1. - pf * coef_y0;
  const double qlast = q[n - 1];
  const int m = n - 1;  // unknowns of the reduced (non-periodic) systems
  double a[M], b[M], c[M], d[2][M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r < m) {
      a[k] = (r == 0) ? 0. : l;
      b[k] = cc;
      c[k] = (r == m - 1) ? 0. : u;
      d[0][k] = q[r];
      d[1][k] = (r == 0) ? -l : (r == m - 1 ? -u : 0.);
      if (m == 1) d[1][k] = -l;  // single unknown: both couplings on row 0
    } else {
      a[k] = 0.; b[k] = 1.; c[k] = 0.; d[0][k] = 0.; d[1][k] = 0.;
    }
  }
  block_tridiag<NT, M, 2>(a, b, c, d, lds);
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r == 0) { ends[0] = d[0][k]; ends[2] = d[1][k]; }
    if (r == m - 1) { ends[1] = d[0][k]; ends[3] = d[1][k]; }
  }
  __syncthreads();
  const double xl = (qlast - u * ends[0] - l * ends[1]) / (cc + u * ends[2] + l * ends[3]);
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r < m) q[r] = d[0][k] + xl * d[1][k];
  }
  if (threadIdx.x == 0) q[n - 1] = xl;
}

template <int KIND, int NT, int M, class S>
static void cny_launch(hipStream_t s, S src, int nrows, int nz) {
  hipLaunchKernelGGL((k_cny_rows<KIND, NT, M, S>), dim3(nrows, 1, nz), dim3(NT), 0, s, src);
}

template <int KIND, class S>
static void cny_dispatch(hipStream_t s, const Geo &g, S src, int nrows, int nz) {
  const int m = g.jtot - 1;
  if (m <= 64 * 4) cny_launch<KIND, 64, 4>(s, src, nrows, nz);
  else if (m <= 256 * 4) cny_launch<KIND, 256, 4>(s, src, nrows, nz);
  else if (m <= 256 * 8) cny_launch<KIND, 256, 8>(s, src, nrows, nz);
  else if (m <= 256 * 16) cny_launch<KIND, 256, 16>(s, src, nrows, nz);
  else if (m <= 256 * 32) cny_launch<KIND, 256, 32>(s, src, nrows, nz);
  else cny_launch<KIND, 512, 32>(s, src, nrows, nz);  // jtot <= 16384 (snsg_create)
}

template <class S>
static void cny_kind(hipStream_t s, const Geo &g, int kind, S src, int nrows, int nz) {
  if (kind == KIND_UX) cny_dispatch<KIND_UX>(s, g, src, nrows, nz);
  else if (kind == KIND_UY) cny_dispatch<KIND_UY>(s, g, src, nrows, nz);
  else cny_dispatch<KIND_TEMP>(s, g, src, nrows, nz);
}

void launch_cny_rows(hipStream_t s, const Geo &g, int kind, RK rk, const double *dtp, int nrows,
                     double *qy) {
  YArgs a{};
  a.g = g;
  a.rk = rk;
  a.dtp = dtp;
  a.n = nrows;
  a.qy = qy;
  cny_kind(s, g, kind, ArgOne<YArgs>{a}, nrows, 1);
}

// ------------------------------------------------------------ transpose

// recv[i*g_jsize + j] = send[j*g_isize + i] through a padded LDS tile, so
// both the read and the write are row-contiguous (the permutation of
// src/parallel/transpose.c:6-85 on one rank).
template <typename T>
__global__ void __launch_bounds__(256)
    k_transpose(int g_isize, int g_jsize, const T *__restrict__ send, T *__restrict__ recv) {
  __shared__ T tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < g_isize && j < g_jsize) tile[r][tx] = send[(size_t)j * g_isize + i];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + r;
    if (i < g_isize && j < g_jsize) recv[(size_t)i * g_jsize + j] = tile[tx][r];
  }
}

void launch_transpose(hipStream_t s, int g_isize, int g_jsize, size_t elemsize, const void *send,
                      void *recv) {
  const dim3 gr((g_isize + 31) / 32, (g_jsize + 31) / 32);
  if (elemsize == 4)
    hipLaunchKernelGGL(k_transpose<uint32_t>, gr, dim3(256), 0, s, g_isize, g_jsize,
                       (const uint32_t *)send, (uint32_t *)recv);
  else if (elemsize == 8)
    hipLaunchKernelGGL(k_transpose<uint2>, gr, dim3(256), 0, s, g_isize, g_jsize,
                       (const uint2 *)send, (uint2 *)recv);
  else
    hipLaunchKernelGGL(k_transpose<uint4>, gr, dim3(256), 0, s, g_isize, g_jsize,
                       (const uint4 *)send, (uint4 *)recv);
}

// ------------------------------------------- y solve after the predictor

// DIR 0: qy[i*jtot + j] = qf[j*n + i]; DIR 1: qf[j*n + i] = qy[i*jtot + j]
// (k_transpose's tiling, for instance blockIdx.z)
template <int DIR, class S>
__global__ void __launch_bounds__(256) k_ytrans(S src) {
  const YArgs a = src.get();
  __shared__ double tile[32][33];
  const int isz = DIR == 0 ? a.n : a.g.jtot, jsz = DIR == 0 ? a.g.jtot : a.n;
  const double *__restrict__ send = DIR == 0 ? a.qf : a.qy;
  double *__restrict__ recv = DIR == 0 ? a.qy : a.qf;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    if (i < isz && j < jsz) tile[r][tx] = send[(size_t)j * isz + i];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = j0 + tx, i = i0 + r;
    if (i < isz && j < jsz) recv[(size_t)i * jsz + j] = tile[tx][r];
  }
}

// out = in + q (x-aligned increment after the y solve), with boundary values
template <int KIND, class S>
__global__ void k_add_out(S src) {
  const YArgs a = src.get();
  const Geo &g = a.g;
  const int n = a.n;
  const int i0 = KIND == KIND_UX ? 2 : 1;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y + 1;
  if (ii >= n) return;
  const int i = ii + i0;
  const double d = a.qf[(size_t)(j - 1) * n + ii];
  if (KIND == KIND_UX) put_ux(g, a.out, i, j, a.in[ux_idx(g, i, j)] + d);
  else if (KIND == KIND_UY) put_c<0>(g, a.out, i, j, a.in[c_idx(g, i, j)] + d);
  else put_c<2>(g, a.out, i, j, a.in[c_idx(g, i, j)] + d);
}

template <class S>
static void ysolve(hipStream_t s, int kind, const YArgs &shape, S src, int nz) {
  const Geo &g = shape.g;
  const int n = shape.n;
  const dim3 t0((n + 31) / 32, (g.jtot + 31) / 32, nz), t1((g.jtot + 31) / 32, (n + 31) / 32, nz);
  hipLaunchKernelGGL((k_ytrans<0, S>), t0, dim3(256), 0, s, src);
  cny_kind(s, g, kind, src, n, nz);
  hipLaunchKernelGGL((k_ytrans<1, S>), t1, dim3(256), 0, s, src);
  const dim3 gr((n + 255) / 256, g.jtot, nz);
  if (kind == KIND_UX) hipLaunchKernelGGL((k_add_out<KIND_UX, S>), gr, dim3(256), 0, s, src);
  else if (kind == KIND_UY) hipLaunchKernelGGL((k_add_out<KIND_UY, S>), gr, dim3(256), 0, s, src);
  else hipLaunchKernelGGL((k_add_out<KIND_TEMP, S>), gr, dim3(256), 0, s, src);
}

void launch_ysolve(hipStream_t s, int kind, const YArgs &a) {
  ysolve(s, kind, a, ArgOne<YArgs>{a}, 1);
}

void launch_ysolve_batch(hipStream_t s, int kind, const YArgs &shape, const YArgs *dev, int nz) {
  ysolve(s, kind, shape, arg_batch<YArgs>(dev), nz);
}

}  // namespace sns
