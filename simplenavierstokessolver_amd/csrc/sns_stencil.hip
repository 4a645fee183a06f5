// sns_stencil.hip -- point-wise and stencil kernels of the hot path (gfx950).
//
// Each kernel restates one reference loop nest; the boundary values that the
// reference refreshes with a separate call right after the loop are written
// by the same thread that produces the interior value they copy, so a field
// leaves every kernel with its halos and walls already consistent.
#include "sns_launch.h"

namespace sns {

static constexpr int kBlock = 256;

static inline dim3 grid2d(int ni, int nj) { return dim3((ni + kBlock - 1) / kBlock, nj); }

// Store v at interior point (i, j) of a field of the given kind together with
// every boundary entry that the reference derives from it:
//   ux  (src/fluid/boundary_conditions.c:6-19)  halo rows for i=2..itot, UX(1,.)=UX(itot+1,.)=0
//   uy  (:21-37)  halo rows for i=1..itot, UY(0,.)=UY(itot+1,.)=0
//   p   (:39-55)  halo rows, P(0,.)=P(1,.), P(itot+1,.)=P(itot,.) incl. halo rows
//   T   (src/temperature/boundary_conditions.c:6-21) halo rows, T walls on j=1..jtot only
template <int KIND>
__device__ __forceinline__ void store_bc(const Geo &g, double *f, int i, int j, double v);

template <>
__device__ __forceinline__ void store_bc<KIND_UX>(const Geo &g, double *f, int i, int j,
                                                  double v) {
  f[ux_idx(g, i, j)] = v;
  if (j == 1) f[ux_idx(g, i, g.jtot + 1)] = v;
  if (j == g.jtot) f[ux_idx(g, i, 0)] = v;
  if (i == 2 || i == g.itot) {
    const int iw = (i == 2) ? 1 : g.itot + 1;
    f[ux_idx(g, iw, j)] = 0.;
    if (j == 1) f[ux_idx(g, iw, g.jtot + 1)] = 0.;
    if (j == g.jtot) f[ux_idx(g, iw, 0)] = 0.;
    if (i == 2 && i == g.itot) {  // itot == 2: both walls
      f[ux_idx(g, g.itot + 1, j)] = 0.;
      if (j == 1) f[ux_idx(g, g.itot + 1, g.jtot + 1)] = 0.;
      if (j == g.jtot) f[ux_idx(g, g.itot + 1, 0)] = 0.;
    }
  }
}

template <int WALLKIND>  // 0: zero walls (uy), 1: Neumann copies (p), 2: temperature walls
__device__ __forceinline__ void store_bc_c(const Geo &g, double *f, int i, int j, double v) {
  f[c_idx(g, i, j)] = v;
  if (j == 1) f[c_idx(g, i, g.jtot + 1)] = v;
  if (j == g.jtot) f[c_idx(g, i, 0)] = v;
  for (int side = 0; side < 2; side++) {
    const bool hit = side == 0 ? (i == 1) : (i == g.itot);
    if (!hit) continue;
    const int iw = side == 0 ? 0 : g.itot + 1;
    double w;
    if (WALLKIND == 0) w = 0.;
    else if (WALLKIND == 1) w = v;
    else w = side == 0 ? +0.5 : -0.5;
    f[c_idx(g, iw, j)] = w;
    if (WALLKIND != 2) {
      if (j == 1) f[c_idx(g, iw, g.jtot + 1)] = w;
      if (j == g.jtot) f[c_idx(g, iw, 0)] = w;
    }
  }
}

template <>
__device__ __forceinline__ void store_bc<KIND_UY>(const Geo &g, double *f, int i, int j,
                                                  double v) {
  store_bc_c<0>(g, f, i, j, v);
}
template <>
__device__ __forceinline__ void store_bc<KIND_TEMP>(const Geo &g, double *f, int i, int j,
                                                    double v) {
  store_bc_c<2>(g, f, i, j, v);
}

// ------------------------------------------------------------ boundaries

__global__ void k_bc_ux(Geo g, double *ux) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 && t <= g.itot) {
    ux[ux_idx(g, t, 0)] = ux[ux_idx(g, t, g.jtot)];
    ux[ux_idx(g, t, g.jtot + 1)] = ux[ux_idx(g, t, 1)];
  }
  if (t <= g.jtot + 1) {
    ux[ux_idx(g, 1, t)] = 0.;
    ux[ux_idx(g, g.itot + 1, t)] = 0.;
  }
}

// mode 0: uy (zero walls), 1: p (Neumann), 2: temperature (Dirichlet, j=1..jtot)
__global__ void k_bc_c(Geo g, double *f, int mode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 1 && t <= g.itot) {
    f[c_idx(g, t, 0)] = f[c_idx(g, t, g.jtot)];
    f[c_idx(g, t, g.jtot + 1)] = f[c_idx(g, t, 1)];
  }
  if (t <= g.jtot + 1) {
    if (mode == 0) {
      f[c_idx(g, 0, t)] = 0.;
      f[c_idx(g, g.itot + 1, t)] = 0.;
    } else if (mode == 1) {
      // the source row of a halo row is the interior row it copies
      const int js = (t == 0) ? g.jtot : (t == g.jtot + 1 ? 1 : t);
      f[c_idx(g, 0, t)] = f[c_idx(g, 1, js)];
      f[c_idx(g, g.itot + 1, t)] = f[c_idx(g, g.itot, js)];
    } else if (t >= 1 && t <= g.jtot) {
      f[c_idx(g, 0, t)] = +0.5;
      f[c_idx(g, g.itot + 1, t)] = -0.5;
    }
  }
}

static inline int bc_blocks(const Geo &g) {
  const int n = (g.itot + 2 > g.jtot + 2) ? g.itot + 2 : g.jtot + 2;
  return (n + kBlock - 1) / kBlock;
}

void launch_bc_ux(hipStream_t s, const Geo &g, double *ux) {
  hipLaunchKernelGGL(k_bc_ux, dim3(bc_blocks(g)), dim3(kBlock), 0, s, g, ux);
}
void launch_bc_uy(hipStream_t s, const Geo &g, double *uy) {
  hipLaunchKernelGGL(k_bc_c, dim3(bc_blocks(g)), dim3(kBlock), 0, s, g, uy, 0);
}
void launch_bc_p(hipStream_t s, const Geo &g, double *p) {
  hipLaunchKernelGGL(k_bc_c, dim3(bc_blocks(g)), dim3(kBlock), 0, s, g, p, 1);
}
void launch_bc_temp(hipStream_t s, const Geo &g, double *temp) {
  hipLaunchKernelGGL(k_bc_c, dim3(bc_blocks(g)), dim3(kBlock), 0, s, g, temp, 2);
}

// ------------------------------------------------------------ temperature

// src/temperature/force.c:6-32
__global__ void k_force(Geo g, int with_thermal_forcing, const double *__restrict__ temp,
                        double *__restrict__ tf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int j = blockIdx.y + 1;
  if (i > g.itot) return;
  double v = 0.;
  if (with_thermal_forcing) {
    const double temp_xm = temp[c_idx(g, i - 1, j)];
    const double temp_xp = temp[c_idx(g, i, j)];
    v = 0.5 * temp_xm + 0.5 * temp_xp;
  }
  tf[sx_idx(g, i, j)] = v;
}

void launch_force(hipStream_t s, const Geo &g, int with_thermal_forcing, const double *temp,
                  double *tempforcex) {
  if (g.itot < 2) return;
  hipLaunchKernelGGL(k_force, grid2d(g.itot - 1, g.jtot), dim3(kBlock), 0, s, g,
                     with_thermal_forcing, temp, tempforcex);
}

// src/temperature/update.c:11-87 (the copy SRCTEMPB <- SRCTEMPA of line 31
// is a pointer exchange done by the caller)
__global__ void k_src_temp(Geo g, Treat tr, const double *__restrict__ ux,
                           const double *__restrict__ uy, const double *__restrict__ temp,
                           double *__restrict__ srca, double *__restrict__ srcg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > g.itot) return;
  double a, gg;
  temp_src(g, ux, uy, temp, i, j, tr.w_ex_x, tr.w_ex_y, tr.w_im_x, tr.w_im_y, a, gg);
  srca[sy_idx(g, i, j)] = a;
  srcg[sy_idx(g, i, j)] = gg;
}

void launch_src_temp(hipStream_t s, const Geo &g, const Treat &tr, const double *ux,
                     const double *uy, const double *temp, double *srca, double *srcg) {
  hipLaunchKernelGGL(k_src_temp, grid2d(g.itot, g.jtot), dim3(kBlock), 0, s, g, tr, ux, uy,
                     temp, srca, srcg);
}

// ------------------------------------------------------------------ fluid

// src/fluid/update_velocity.c:11-192 (both momentum components in one pass)
__global__ void k_src_vel(Geo g, Treat tr, const double *__restrict__ ux,
                          const double *__restrict__ uy, const double *__restrict__ p,
                          const double *__restrict__ tf, double *__restrict__ srcuxa,
                          double *__restrict__ srcuxg, double *__restrict__ srcuya,
                          double *__restrict__ srcuyg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > g.itot) return;
  double a, gg;
  if (i >= 2) {
    ux_src(g, ux, uy, p, tf[sx_idx(g, i, j)], i, j, tr.w_ex_x, tr.w_ex_y, tr.w_im_x, tr.w_im_y,
           a, gg);
    srcuxa[sx_idx(g, i, j)] = a;
    srcuxg[sx_idx(g, i, j)] = gg;
  }
  uy_src(g, ux, uy, p, i, j, tr.w_ex_x, tr.w_ex_y, tr.w_im_x, tr.w_im_y, a, gg);
  srcuya[sy_idx(g, i, j)] = a;
  srcuyg[sy_idx(g, i, j)] = gg;
}

void launch_src_vel(hipStream_t s, const Geo &g, const Treat &tr, const double *ux,
                    const double *uy, const double *p, const double *tempforcex, double *srcuxa,
                    double *srcuxg, double *srcuya, double *srcuyg) {
  hipLaunchKernelGGL(k_src_vel, grid2d(g.itot, g.jtot), dim3(kBlock), 0, s, g, tr, ux, uy, p,
                     tempforcex, srcuxa, srcuxg, srcuya, srcuyg);
}

// ------------------------------------------------------- increments

// first interior index and count of unknowns per row for each field kind
__host__ __device__ inline int kind_i0(int kind) { return kind == KIND_UX ? 2 : 1; }
__host__ __device__ inline int kind_n(const Geo &g, int kind) {
  return kind == KIND_UX ? g.itot - 1 : g.itot;
}

template <int KIND>
__device__ __forceinline__ void store_field(const Geo &g, double *f, int i, int j, double v) {
  store_bc<KIND>(g, f, i, j, v);
}
template <int KIND>
__device__ __forceinline__ double load_field(const Geo &g, const double *f, int i, int j) {
  return KIND == KIND_UX ? f[ux_idx(g, i, j)] : f[c_idx(g, i, j)];
}

// src/temperature/update.c:114-122,222-229; src/fluid/update_velocity.c:219-229,334-342,
// 371-379,478-486
template <int KIND>
__global__ void k_incr(Geo g, RK rk, const double *__restrict__ dtp,
                       const double *__restrict__ srca, const double *__restrict__ srcb,
                       const double *__restrict__ srcg, double *__restrict__ field,
                       double *__restrict__ q) {
  const int n = KIND == KIND_UX ? g.itot - 1 : g.itot;
  const int i0 = KIND == KIND_UX ? 2 : 1;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y + 1;
  if (ii >= n) return;
  const int i = ii + i0;
  const size_t s = (size_t)(j - 1) * n + ii;
  const double dt = *dtp;
  const double d = rk_incr(rk.alpha, rk.beta, rk.gamma, dt, srca[s], srcb[s], srcg[s]);
  if (q) {
    q[s] = d;
  } else {
    store_field<KIND>(g, field, i, j, load_field<KIND>(g, field, i, j) + d);
  }
}

template <int KIND>
__global__ void k_addq(Geo g, const double *__restrict__ q, double *__restrict__ field) {
  const int n = KIND == KIND_UX ? g.itot - 1 : g.itot;
  const int i0 = KIND == KIND_UX ? 2 : 1;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y + 1;
  if (ii >= n) return;
  const int i = ii + i0;
  const size_t s = (size_t)(j - 1) * n + ii;
  store_field<KIND>(g, field, i, j, load_field<KIND>(g, field, i, j) + q[s]);
}

void launch_incr(hipStream_t s, const Geo &g, int kind, RK rk, const double *dtp,
                 const double *srca, const double *srcb, const double *srcg, double *field,
                 double *q) {
  const dim3 gr = grid2d(kind_n(g, kind), g.jtot);
  if (kind == KIND_UX)
    hipLaunchKernelGGL(k_incr<KIND_UX>, gr, dim3(kBlock), 0, s, g, rk, dtp, srca, srcb, srcg,
                       field, q);
  else if (kind == KIND_UY)
    hipLaunchKernelGGL(k_incr<KIND_UY>, gr, dim3(kBlock), 0, s, g, rk, dtp, srca, srcb, srcg,
                       field, q);
  else
    hipLaunchKernelGGL(k_incr<KIND_TEMP>, gr, dim3(kBlock), 0, s, g, rk, dtp, srca, srcb, srcg,
                       field, q);
}

void launch_addq(hipStream_t s, const Geo &g, int kind, const double *q, double *field) {
  const dim3 gr = grid2d(kind_n(g, kind), g.jtot);
  if (kind == KIND_UX)
    hipLaunchKernelGGL(k_addq<KIND_UX>, gr, dim3(kBlock), 0, s, g, q, field);
  else if (kind == KIND_UY)
    hipLaunchKernelGGL(k_addq<KIND_UY>, gr, dim3(kBlock), 0, s, g, q, field);
  else
    hipLaunchKernelGGL(k_addq<KIND_TEMP>, gr, dim3(kBlock), 0, s, g, q, field);
}

// ------------------------------------------------- correction, pressure

// src/fluid/correct_velocity.c:6-42
__global__ void k_correct(Geo g, double gamma, const double *__restrict__ dtp,
                          const double *__restrict__ psi, double *__restrict__ ux,
                          double *__restrict__ uy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > g.itot) return;
  const double dt = *dtp;
  if (i >= 2) {
    const double dxi = 1. / DXC(g, i);
    const double psi_xp = psi[c_idx(g, i, j)];
    const double psi_xm = psi[c_idx(g, i - 1, j)];
    const double v = ux[ux_idx(g, i, j)] - gamma * dt * (psi_xp - psi_xm) * dxi;
    store_bc<KIND_UX>(g, ux, i, j, v);
  }
  {
    const double psi_yp = psi[c_idx(g, i, j)];
    const double psi_ym = psi[c_idx(g, i, j - 1)];
    const double v = uy[c_idx(g, i, j)] - gamma * dt * (psi_yp - psi_ym) / g.dy;
    store_bc<KIND_UY>(g, uy, i, j, v);
  }
}

void launch_correct(hipStream_t s, const Geo &g, double gamma, const double *dtp,
                    const double *psi, double *ux, double *uy) {
  hipLaunchKernelGGL(k_correct, grid2d(g.itot, g.jtot), dim3(kBlock), 0, s, g, gamma, dtp, psi,
                     ux, uy);
}

// src/fluid/update_pressure.c:7-60
__global__ void k_pressure(Geo g, Treat tr, double gamma, const double *__restrict__ dtp,
                           const double *__restrict__ psi, double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > g.itot) return;
  const double dt = *dtp;
  const double psi_c = psi[c_idx(g, i, j)];
  double v = p[c_idx(g, i, j)] + psi_c;
  if (tr.implicitx) {
    const double dpsidx_xm = (-psi[c_idx(g, i - 1, j)] + psi_c) / DXC(g, i);
    const double dpsidx_xp = (-psi_c + psi[c_idx(g, i + 1, j)]) / DXC(g, i + 1);
    v += -(gamma * dt) / g.two_Re * (+(dpsidx_xp - dpsidx_xm) / DXF(g, i));
  }
  if (tr.implicity) {
    const double dpsidy_ym = (-psi[c_idx(g, i, j - 1)] + psi_c) / g.dy;
    const double dpsidy_yp = (-psi_c + psi[c_idx(g, i, j + 1)]) / g.dy;
    v += -(gamma * dt) / g.two_Re * (+(dpsidy_yp - dpsidy_ym) / g.dy);
  }
  store_bc_c<1>(g, p, i, j, v);
}

void launch_pressure(hipStream_t s, const Geo &g, const Treat &tr, double gamma,
                     const double *dtp, const double *psi, double *p) {
  hipLaunchKernelGGL(k_pressure, grid2d(g.itot, g.jtot), dim3(kBlock), 0, s, g, tr, gamma, dtp,
                     psi, p);
}

// ------------------------------------------------------------ dt limit

// src/param/decide_dt.c:25-45.  min() is exact and order-free, so the
// parallel reduction returns the same value as the sequential loop.
__global__ void __launch_bounds__(kBlock) k_dt_reduce(Geo g, const double *__restrict__ ux,
                                                      const double *__restrict__ uy,
                                                      unsigned long long *out) {
  const double small = 1.e-8;
  double m = 5.e-2;
  const long long nrow = g.itot;  // covers ux (i=2..itot) and uy (i=1..itot)
  const long long total = nrow * g.jtot;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(idx / nrow) + 1;
    const int i = (int)(idx % nrow) + 1;
    if (i >= 2) {
      const double vel = fmax(fabs(ux[ux_idx(g, i, j)]), small);
      m = fmin(m, DXC(g, i) / vel);
    }
    const double vel = fmax(fabs(uy[c_idx(g, i, j)]), small);
    m = fmin(m, g.dy / vel);
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  // positive doubles order like their bit patterns
  if (threadIdx.x == 0) atomicMin(out, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

void launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n) {
  hipLaunchKernelGGL(k_fill_u64, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, p, v, n);
}

// Plain device copy of n 16-byte words (grid-stride, one load and one store
// per word).
__global__ void __launch_bounds__(kBlock) k_copy(const double2 *__restrict__ src,
                                                 double2 *__restrict__ dst, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

void launch_copy(hipStream_t s, const void *src, void *dst, size_t n16) {
  size_t blocks = (n16 + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(kBlock), 0, s,
                     (const double2 *)src, (double2 *)dst, n16);
}

// Tiled device copy: each 256-thread block moves 256*U 16-byte words, every
// thread issuing its U loads before its U stores (U independent loads in
// flight per thread); NT selects nontemporal loads and stores.  The shape
// sweep of this kernel measures the device's practical copy bandwidth.
typedef double v2d __attribute__((ext_vector_type(2)));

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_copy_tile(const v2d *__restrict__ src,
                                                  v2d *__restrict__ dst, size_t n) {
  const size_t base = (size_t)blockIdx.x * (256 * U) + threadIdx.x;
  if (base + (size_t)(U - 1) * 256 < n) {
    v2d v[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      if constexpr (NT) v[k] = __builtin_nontemporal_load(src + base + k * 256);
      else v[k] = src[base + k * 256];
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
      if constexpr (NT) __builtin_nontemporal_store(v[k], dst + base + k * 256);
      else dst[base + k * 256] = v[k];
    }
  } else {
    for (int k = 0; k < U; k++) {
      const size_t e = base + (size_t)k * 256;
      if (e < n) dst[e] = src[e];
    }
  }
}

template <int U, bool NT>
static void copy_tile_launch(hipStream_t s, const void *src, void *dst, size_t n16) {
  const size_t blocks = (n16 + 256 * U - 1) / (256 * U);
  hipLaunchKernelGGL((k_copy_tile<U, NT>), dim3((unsigned)(blocks ? blocks : 1)), dim3(256), 0,
                     s, (const v2d *)src, (v2d *)dst, n16);
}

bool launch_copy_shaped(hipStream_t s, const void *src, void *dst, size_t n16, int loads,
                        bool nontemporal) {
  if ((n16 + 255) / 256 > 0x7fffffffull) return false;
#define SNS_COPY_CASE(U)                                                       \
  case U:                                                                      \
    if (nontemporal) copy_tile_launch<U, true>(s, src, dst, n16);             \
    else copy_tile_launch<U, false>(s, src, dst, n16);                        \
    return true;
  switch (loads) {
    SNS_COPY_CASE(1)
    SNS_COPY_CASE(2)
    SNS_COPY_CASE(4)
    SNS_COPY_CASE(8)
    default: return false;
  }
#undef SNS_COPY_CASE
}

void launch_dt_reduce(hipStream_t s, const Geo &g, const double *ux, const double *uy,
                      unsigned long long *out) {
  const long long total = (long long)g.itot * g.jtot;
  long long blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_dt_reduce, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, ux, uy, out);
}

// ------------------------------------------------ initial temperature

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Interior T(i, j) = A*(2U-1), U = top 53 bits of
// splitmix64(seed<<40 ^ ((j-1)*itot + (i-1))) times 2^-53: the seeded
// noise of the host layer's initial condition (which replaces the rand()
// field of src/temperature/init.c:39-47), evaluated per cell.
__global__ void k_init_temp(Geo g, double *temp, uint64_t seed, double amplitude) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > g.itot) return;
  const uint64_t key = (seed << 40) ^ ((uint64_t)(j - 1) * (uint64_t)g.itot + (uint64_t)(i - 1));
  const double u = (double)(splitmix64(key) >> 11) * 0x1.0p-53;
  temp[c_idx(g, i, j)] = 0. + amplitude * (2. * u - 1.);
}

void launch_init_temp(hipStream_t s, const Geo &g, double *temp, uint64_t seed,
                      double amplitude) {
  hipLaunchKernelGGL(k_init_temp, dim3((g.itot + 255) / 256, g.jtot), dim3(256), 0, s, g, temp,
                     seed, amplitude);
  launch_bc_temp(s, g, temp);
}

}  // namespace sns
