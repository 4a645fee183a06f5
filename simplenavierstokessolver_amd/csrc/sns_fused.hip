// sns_fused.hip -- fused substage kernels for the time-step hot path (gfx950).
//
// One RK substage (src/main.c:81-98) becomes five kernels:
//   pred : T, ux and uy predictors in one pass over the substage-start fields
//          (buoyancy from T in registers, RK history read/written in place,
//          Crank-Nicolson in x solved on chip for a whole row), writing
//          T' (ping-pong buffer), ux*, uy* with their boundary values;
//   poisson_fwd / poisson_x / poisson_inv : the potential (sns_solve.hip);
//   corr : velocity correction and pressure update in one pass, plus on
//          the last substage the advective time-step bound for the next
//          step (a min-reduction over the corrected velocities).
// The explicit and implicit diffusive parts, the RK increments and the
// boundary values follow the reference expression for expression (see
// sns_fast.h for the division scheme); only the line solves round
// differently.
#include "sns_launch.h"
#include "sns_fast.h"
#include "sns_fused.h"
#include "sns_tridiag.h"

namespace sns {

// ---------------------------------------------------------------- stores

__device__ __forceinline__ void put_ux(const Geo &g, double *f, int i, int j, double v) {
  f[ux_idx(g, i, j)] = v;
  if (j == 1) f[ux_idx(g, i, g.jtot + 1)] = v;
  if (j == g.jtot) f[ux_idx(g, i, 0)] = v;
  if (i == 2 || i == g.itot) {
    const int iw = (i == 2) ? 1 : g.itot + 1;
    f[ux_idx(g, iw, j)] = 0.;
    if (j == 1) f[ux_idx(g, iw, g.jtot + 1)] = 0.;
    if (j == g.jtot) f[ux_idx(g, iw, 0)] = 0.;
  }
}

// WALL: 0 zero walls (uy), 1 Neumann copies (p), 2 temperature walls (rows 1..jtot)
template <int WALL>
__device__ __forceinline__ void put_c(const Geo &g, double *f, int i, int j, double v) {
  f[c_idx(g, i, j)] = v;
  if (j == 1) f[c_idx(g, i, g.jtot + 1)] = v;
  if (j == g.jtot) f[c_idx(g, i, 0)] = v;
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const bool hit = side == 0 ? (i == 1) : (i == g.itot);
    if (!hit) continue;
    const int iw = side == 0 ? 0 : g.itot + 1;
    const double w = WALL == 0 ? 0. : (WALL == 1 ? v : (side == 0 ? +0.5 : -0.5));
    f[c_idx(g, iw, j)] = w;
    if (WALL != 2) {
      if (j == 1) f[c_idx(g, iw, g.jtot + 1)] = w;
      if (j == g.jtot) f[c_idx(g, iw, 0)] = w;
    }
  }
}

// ------------------------------------------------------------- predictor

// Increments of the three fields at (i, j) from the substage-start state;
// RK history (explicit terms) is read (beta != 0) and rewritten in place.
struct Incr {
  double t, u, v;
  bool has_t, has_u, has_v;
};

__device__ __forceinline__ Incr pred_point(const PredArgs &a, int i, int j, double dt) {
  const Geo &g = a.g;
  Incr r;
  r.has_t = a.with_temp && i <= g.itot;
  r.has_u = i >= 2 && i <= g.itot;
  r.has_v = i <= g.itot;
  r.t = r.u = r.v = 0.;
  if (r.has_t) {
    double A, G;
    temp_src_t(g, a.tb, a.ux, a.uy, a.temp, i, j, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x,
               a.tr.w_im_y, A, G);
    const size_t s = sy_idx(g, i, j);
    const double B = a.read_hist ? a.hist_t[s] : 0.;
    if (a.write_hist) a.hist_t[s] = A;
    r.t = rk_incr(a.alpha, a.beta, a.gamma, dt, A, B, G);
  }
  if (r.has_u) {
    double tf = 0.;
    if (a.forcing) tf = 0.5 * a.temp[c_idx(g, i - 1, j)] + 0.5 * a.temp[c_idx(g, i, j)];
    double A, G;
    ux_src_t(g, a.tb, a.ux, a.uy, a.p, tf, i, j, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x,
             a.tr.w_im_y, A, G);
    const size_t s = sx_idx(g, i, j);
    const double B = a.read_hist ? a.hist_ux[s] : 0.;
    if (a.write_hist) a.hist_ux[s] = A;
    r.u = rk_incr(a.alpha, a.beta, a.gamma, dt, A, B, G);
  }
  if (r.has_v) {
    double A, G;
    uy_src_t(g, a.tb, a.ux, a.uy, a.p, i, j, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x, a.tr.w_im_y,
             A, G);
    const size_t s = sy_idx(g, i, j);
    const double B = a.read_hist ? a.hist_uy[s] : 0.;
    if (a.write_hist) a.hist_uy[s] = A;
    r.v = rk_incr(a.alpha, a.beta, a.gamma, dt, A, B, G);
  }
  return r;
}

// Final value of each field at (i, j), or the x-aligned increment when a
// y solve follows.
__device__ __forceinline__ void pred_emit(const PredArgs &a, int i, int j, const Incr &r) {
  const Geo &g = a.g;
  if (a.q_t) {
    if (r.has_t) a.q_t[sy_idx(g, i, j)] = r.t;
    if (r.has_u) a.q_u[sx_idx(g, i, j)] = r.u;
    if (r.has_v) a.q_v[sy_idx(g, i, j)] = r.v;
    return;
  }
  if (r.has_t) put_c<2>(g, a.out_t, i, j, a.temp[c_idx(g, i, j)] + r.t);
  if (r.has_u) put_ux(g, a.out_u, i, j, a.ux[ux_idx(g, i, j)] + r.u);
  if (r.has_v) put_c<0>(g, a.out_v, i, j, a.uy[c_idx(g, i, j)] + r.v);
}

// explicit diffusion in x: point-wise
__global__ void __launch_bounds__(256) k_pred_exp(PredArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = xcd_remap(blockIdx.y, gridDim.y) + 1;
  if (i > a.g.itot) return;
  const double dt = *a.dtp;
  const Incr r = pred_point(a, i, j, dt);
  pred_emit(a, i, j, r);
}

// Crank-Nicolson in x: one workgroup per row; increments staged in LDS
// (lane order for the stencils, chunk order for the solver).
template <int NT, int M>
struct PredLds {
  double d[3][NT * (M + 1)];  // padded: chunk t starts at t*(M+1)
  TridiagLds<NT, 1> s;
};

template <int M>
__device__ __forceinline__ int pad_idx(int r) {
  return r + r / M;
}

template <int KIND, int NT, int M>
__device__ __forceinline__ void cnx_solve_row(const PredArgs &a, double pf, double *sd,
                                              TridiagLds<NT, 1> &s) {
  const Geo &g = a.g;
  const int n = KIND == KIND_UX ? g.itot - 1 : g.itot;
  const int i0 = KIND == KIND_UX ? 2 : 1;
  const double *cxm = KIND == KIND_UX ? a.tb.nU_xm : a.tb.nT_xm;
  const double *cxp = KIND == KIND_UX ? a.tb.nU_xp : a.tb.nT_xp;
  double va[M], vb[M], vc[M], vd[1][M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r < n) {
      // src/temperature/update.c:141-149, src/fluid/update_velocity.c:253-261,398-406
      const double coef_xm = cxm[r + i0 - 1];
      const double coef_xp = cxp[r + i0 - 1];
      const double coef_x0 = -coef_xm - coef_xp;
      va[k] = (r == 0) ? 0. : -pf * coef_xm;
      vc[k] = (r == n - 1) ? 0. : -pf * coef_xp;
      vb[k] = 1. - pf * coef_x0;
      vd[0][k] = sd[pad_idx<M>(r)];
    } else {
      va[k] = 0.; vb[k] = 1.; vc[k] = 0.; vd[0][k] = 0.;
    }
  }
  block_tridiag<NT, M, 1>(va, vb, vc, vd, s);
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r < n) sd[pad_idx<M>(r)] = vd[0][k];
  }
}

template <int NT, int M>
__global__ void __launch_bounds__(NT) k_pred_impx(PredArgs a) {
  extern __shared__ double lds_raw[];
  PredLds<NT, M> &L = *reinterpret_cast<PredLds<NT, M> *>(lds_raw);
  const Geo &g = a.g;
  const int j = xcd_remap(blockIdx.x, gridDim.x) + 1;
  const double dt = *a.dtp;
  const int nk = (g.itot + NT - 1) / NT;
  for (int k = 0; k < nk; k++) {
    const int i = 1 + threadIdx.x + k * NT;
    if (i <= g.itot) {
      const Incr r = pred_point(a, i, j, dt);
      if (r.has_t) L.d[0][pad_idx<M>(i - 1)] = r.t;
      if (r.has_u) L.d[1][pad_idx<M>(i - 2)] = r.u;
      if (r.has_v) L.d[2][pad_idx<M>(i - 1)] = r.v;
    }
  }
  __syncthreads();
  const double gdt = a.gamma * dt;
  if (a.with_temp) cnx_solve_row<KIND_TEMP, NT, M>(a, gdt / g.two_sqrtPr_sqrtRa, L.d[0], L.s);
  const double pfu = (gdt * g.sqrtPr) / g.two_sqrtRa;
  cnx_solve_row<KIND_UX, NT, M>(a, pfu, L.d[1], L.s);
  cnx_solve_row<KIND_UY, NT, M>(a, pfu, L.d[2], L.s);
  __syncthreads();
  for (int k = 0; k < nk; k++) {
    const int i = 1 + threadIdx.x + k * NT;
    if (i <= g.itot) {
      Incr r;
      r.has_t = a.with_temp;
      r.has_u = i >= 2;
      r.has_v = true;
      r.t = r.has_t ? L.d[0][pad_idx<M>(i - 1)] : 0.;
      r.u = r.has_u ? L.d[1][pad_idx<M>(i - 2)] : 0.;
      r.v = L.d[2][pad_idx<M>(i - 1)];
      pred_emit(a, i, j, r);
    }
  }
}

template <int NT, int M>
static void pred_impx_launch(hipStream_t s, const PredArgs &a) {
  const size_t lds = sizeof(PredLds<NT, M>);
  static bool attr = false;
  if (!attr) {
    hipFuncSetAttribute((const void *)k_pred_impx<NT, M>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr = true;
  }
  hipLaunchKernelGGL((k_pred_impx<NT, M>), dim3(a.g.jtot), dim3(NT), lds, s, a);
}

bool pred_impx_supported(int itot) { return itot <= 4096; }

void launch_pred(hipStream_t s, const PredArgs &a, int implicitx) {
  const int itot = a.g.itot;
  if (!implicitx) {
    hipLaunchKernelGGL(k_pred_exp, dim3((itot + 255) / 256, a.g.jtot), dim3(256), 0, s, a);
    return;
  }
  if (itot <= 64 * 4) pred_impx_launch<64, 4>(s, a);
  else if (itot <= 256 * 4) pred_impx_launch<256, 4>(s, a);
  else if (itot <= 512 * 4) pred_impx_launch<512, 4>(s, a);
  else pred_impx_launch<512, 8>(s, a);
}

// ----------------------------------------------------- y-solve completion

// out = in + q (x-aligned increment after the y solve), with boundary values
template <int KIND>
__global__ void k_add_out(Geo g, const double *__restrict__ in, const double *__restrict__ q,
                          double *__restrict__ out) {
  const int n = KIND == KIND_UX ? g.itot - 1 : g.itot;
  const int i0 = KIND == KIND_UX ? 2 : 1;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y + 1;
  if (ii >= n) return;
  const int i = ii + i0;
  const double d = q[(size_t)(j - 1) * n + ii];
  if (KIND == KIND_UX) put_ux(g, out, i, j, in[ux_idx(g, i, j)] + d);
  else if (KIND == KIND_UY) put_c<0>(g, out, i, j, in[c_idx(g, i, j)] + d);
  else put_c<2>(g, out, i, j, in[c_idx(g, i, j)] + d);
}

void launch_add_out(hipStream_t s, const Geo &g, int kind, const double *in, const double *q,
                    double *out) {
  const int n = kind == KIND_UX ? g.itot - 1 : g.itot;
  const dim3 gr((n + 255) / 256, g.jtot);
  if (kind == KIND_UX) hipLaunchKernelGGL(k_add_out<KIND_UX>, gr, dim3(256), 0, s, g, in, q, out);
  else if (kind == KIND_UY) hipLaunchKernelGGL(k_add_out<KIND_UY>, gr, dim3(256), 0, s, g, in, q, out);
  else hipLaunchKernelGGL(k_add_out<KIND_TEMP>, gr, dim3(256), 0, s, g, in, q, out);
}

// ------------------------------------------------- correction + pressure

// src/fluid/correct_velocity.c:6-42 and src/fluid/update_pressure.c:7-60 in
// one pass: reads ux*, uy*, p and psi, writes ux, uy, p with boundary
// values.  With red != nullptr it also folds the advective bound of
// src/param/decide_dt.c:25-45 over the corrected velocities into *red.
__global__ void __launch_bounds__(256) k_corr(CorrArgs a) {
  const Geo &g = a.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = xcd_remap(blockIdx.y, gridDim.y) + 1;
  const double dt = *a.dtp;
  const double gdt = a.gamma * dt;
  double m = 5.e-2;
  if (i <= g.itot) {
    const double psi_c = a.psi[c_idx(g, i, j)];
    if (i >= 2) {
      const double dxi = a.tb.rdxc[i - 1];  // 1./DXC(i)
      const double v = a.ux_star[ux_idx(g, i, j)] - gdt * (psi_c - a.psi[c_idx(g, i - 1, j)]) * dxi;
      put_ux(g, a.ux, i, j, v);
      if (a.red) m = fmin(m, a.tb.dxc[i - 1] / fmax(fabs(v), 1.e-8));
    }
    {
      const double v = a.uy_star[c_idx(g, i, j)] -
                       div_cr(gdt * (psi_c - a.psi[c_idx(g, i, j - 1)]), a.tb.dy, a.tb.rdy);
      put_c<0>(g, a.uy, i, j, v);
      if (a.red) m = fmin(m, a.tb.dy / fmax(fabs(v), 1.e-8));
    }
    double pv = a.p[c_idx(g, i, j)] + psi_c;
    const double fac = -gdt / g.two_Re;
    if (a.tr.implicitx) {
      const double dm = div_cr(-a.psi[c_idx(g, i - 1, j)] + psi_c, a.tb.dxc[i - 1], a.tb.rdxc[i - 1]);
      const double dp = div_cr(-psi_c + a.psi[c_idx(g, i + 1, j)], a.tb.dxc[i], a.tb.rdxc[i]);
      pv += fac * (+div_cr(dp - dm, a.tb.dxf[i - 1], a.tb.rdxf[i - 1]));
    }
    if (a.tr.implicity) {
      const double dm = div_cr(-a.psi[c_idx(g, i, j - 1)] + psi_c, a.tb.dy, a.tb.rdy);
      const double dp = div_cr(-psi_c + a.psi[c_idx(g, i, j + 1)], a.tb.dy, a.tb.rdy);
      pv += fac * (+div_cr(dp - dm, a.tb.dy, a.tb.rdy));
    }
    put_c<1>(g, a.p, i, j, pv);
  }
  if (a.red) {
    __shared__ double red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(a.red, (unsigned long long)__double_as_longlong(red[0]));
  }
}

void launch_corr(hipStream_t s, const CorrArgs &a) {
  hipLaunchKernelGGL(k_corr, dim3((a.g.itot + 255) / 256, a.g.jtot), dim3(256), 0, s, a);
}

// ------------------------------------------------ division self-check

__global__ void k_check_div(int n, const double *__restrict__ a, const double *__restrict__ b,
                            const double *__restrict__ r, double *__restrict__ q_cr,
                            double *__restrict__ q_ieee) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  q_cr[t] = div_cr(a[t], b[t], r[t]);
  q_ieee[t] = a[t] / b[t];
}

void launch_check_div(hipStream_t s, int n, const double *a, const double *b, const double *r,
                      double *q_cr, double *q_ieee) {
  hipLaunchKernelGGL(k_check_div, dim3((n + 255) / 256), dim3(256), 0, s, n, a, b, r, q_cr, q_ieee);
}

// ------------------------------------------------------------- time step

// src/param/decide_dt.c:45-97 on the device for a fixed treatment: the
// advective bound times 0.7, the diffusive bounds (fixed per run, computed on
// the host) and the treatment's dt; the run time advances by dt
// (src/main.c:34-35).  The reduction slot is reset for the next step.
__global__ void k_dt(unsigned long long *red, double *dt, double *time, double dt_dif_x,
                     double dt_dif_y, int implicitx, int implicity) {
  const double dt_adv = __longlong_as_double((long long)*red) * 0.7;
  const double d = implicitx ? implicity ? dt_adv : fmin(dt_adv, dt_dif_y)
                             : implicity ? fmin(dt_adv, dt_dif_x)
                                         : fmin(dt_adv, fmin(dt_dif_x, dt_dif_y));
  *dt = d;
  *time += d;
  *red = 0x7ff0000000000000ull;
}

void launch_dt(hipStream_t s, unsigned long long *red, double *dt, double *time, double dt_dif_x,
               double dt_dif_y, int implicitx, int implicity) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, s, red, dt, time, dt_dif_x, dt_dif_y, implicitx,
                     implicity);
}

}  // namespace sns
