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
#include "sns_lu.h"

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

// ------------------------------------------ shared LU factorisation of CN-x

// Coefficients of row r (0-based unknown index) of the CN-x operator of
// field f (0: T, 1: ux, 2: uy): src/temperature/update.c:141-149,
// src/fluid/update_velocity.c:253-261,398-406.
__device__ __forceinline__ void cnx_row(const Geo &g, const Tables &tb, int f, double pf, int r,
                                        double &l, double &c, double &u) {
  const int i = r + (f == 1 ? 2 : 1);
  const double coef_xm = f == 1 ? tb.nU_xm[i - 1] : tb.nT_xm[i - 1];
  const double coef_xp = f == 1 ? tb.nU_xp[i - 1] : tb.nT_xp[i - 1];
  const double coef_x0 = -coef_xm - coef_xp;
  l = -pf * coef_xm;
  u = -pf * coef_xp;
  c = 1. - pf * coef_x0;
}

__device__ __forceinline__ double cn_pf(const Geo &g, int f, double gdt) {
  return f == 0 ? gdt / g.two_sqrtPr_sqrtRa : (gdt * g.sqrtPr) / g.two_sqrtRa;
}

struct M22 {
  double a, b, c, d;
};
__device__ __forceinline__ M22 mmul(const M22 &x, const M22 &y) {  // x*y
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c,
          x.c * y.b + x.d * y.d};
}
// exact power-of-two rescaling (ratios of entries are unchanged)
__device__ __forceinline__ M22 mnorm(M22 m) {
  const double mx = fmax(fmax(fabs(m.a), fabs(m.b)), fmax(fabs(m.c), fabs(m.d)));
  if (mx == 0. || !isfinite(mx)) return m;
  const int e = ilogb(mx);
  return {ldexp(m.a, -e), ldexp(m.b, -e), ldexp(m.c, -e), ldexp(m.d, -e)};
}
__device__ __forceinline__ M22 mshfl_up(const M22 &m, int off) {
  return {__shfl_up(m.a, off), __shfl_up(m.b, off), __shfl_up(m.c, off), __shfl_up(m.d, off)};
}

// Thomas factorisation of the CN-x operator of field blockIdx.x for the
// substage blockIdx.y, written as the interleaved tables of sns_lu.h.  The
// pivot recurrence u'_i = u_i/(c_i - l_i u'_{i-1}) is the Moebius map of the
// vector v_i = [[0, u_i], [-l_i, c_i]] v_{i-1}, v_0 = (u_0, c_0), with
// u'_i = p_i/q_i and 1/pivot_i = q_{i-1}/q_i; the products of these 2x2
// matrices are scanned across the workgroup.
template <int NT, int M>
__global__ void __launch_bounds__(NT) k_lu_factor(LuFactorArgs a) {
  constexpr int NW = NT / 64;
  __shared__ double wtot[NW][4];
  const Geo &g = a.g;
  const int f = blockIdx.x, s = blockIdx.y;
  const int n = f == 1 ? g.itot - 1 : g.itot;
  const double pf = cn_pf(g, f, a.gamma[s] * (*a.dtp));
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // chunk product of the row matrices (rows >= 1)
  M22 T = {1., 0., 0., 1.};
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = t * M + k;
    if (r >= 1 && r < n) {
      double l, c, u;
      cnx_row(g, a.tb, f, pf, r, l, c, u);
      T = mnorm(mmul(M22{0., u, -l, c}, T));
    }
  }
  // inclusive scan within the wave (later chunks multiply from the left)
  M22 X = T;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const M22 o = mshfl_up(X, off);
    if (lane >= off) X = mnorm(mmul(X, o));
  }
  if (lane == 63) {
    wtot[wave][0] = X.a; wtot[wave][1] = X.b; wtot[wave][2] = X.c; wtot[wave][3] = X.d;
  }
  __syncthreads();
  M22 cin = {1., 0., 0., 1.};
  for (int w = 0; w < wave; w++)
    cin = mnorm(mmul(M22{wtot[w][0], wtot[w][1], wtot[w][2], wtot[w][3]}, cin));
  const M22 incl = mnorm(mmul(X, cin));
  M22 excl = mshfl_up(incl, 1);
  if (lane == 0) excl = cin;
  // v before this chunk's first row
  double l0, c0, u0;
  cnx_row(g, a.tb, f, pf, 0, l0, c0, u0);
  double p = excl.a * u0 + excl.b * c0, q = excl.c * u0 + excl.d * c0;
  double rr[M], al[M], be[M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = t * M + k;
    rr[k] = al[k] = be[k] = 0.;
    if (r >= n) continue;
    double l, c, u;
    cnx_row(g, a.tb, f, pf, r, l, c, u);
    double up;
    if (r == 0) {
      rr[k] = 1. / c;
      up = u / c;
      p = u0;
      q = c0;
    } else {
      const double pn = u * q, qn = -l * p + c * q;
      rr[k] = q / qn;
      up = pn / qn;
      al[k] = -rr[k] * l;
      const double mx = fmax(fabs(pn), fabs(qn));
      const int e = ilogb(mx);
      p = ldexp(pn, -e);
      q = ldexp(qn, -e);
    }
    be[k] = (r == n - 1) ? 0. : -up;
  }
  double *base = a.out + (size_t)((s * 3 + f) * 5) * a.ntm;
  double P = 1., Q = 1.;
#pragma unroll
  for (int k = 0; k < M; k++) {
    P *= al[k];
    base[0 * a.ntm + k * NT + t] = rr[k];
    base[1 * a.ntm + k * NT + t] = al[k];
    base[2 * a.ntm + k * NT + t] = be[k];
    base[3 * a.ntm + k * NT + t] = P;
  }
#pragma unroll
  for (int k = M - 1; k >= 0; k--) {
    Q *= be[k];
    base[4 * a.ntm + k * NT + t] = Q;
  }
}

// ----------------------------------------------- row-wise fused predictor

// Crank-Nicolson in x: one workgroup per row; increments staged in LDS (lane
// order for the stencils, chunk order for the line solves).
template <int NT, int M>
struct PredRowLds {
  double d[3][NT * (M + 1)];  // padded: chunk t starts at t*(M+1)
  double scr[2 * 3 * (NT / 64)];
};

template <int M>
__device__ __forceinline__ int pad_idx(int r) {
  return r + r / M;
}

template <int NT, int M>
__global__ void __launch_bounds__(NT) k_pred_row(PredArgs a, const double *lu, int ntm) {
  extern __shared__ double lds_raw[];
  PredRowLds<NT, M> &L = *reinterpret_cast<PredRowLds<NT, M> *>(lds_raw);
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
  const int nf[3] = {g.itot, g.itot - 1, g.itot};
  double dd[3][M];
#pragma unroll
  for (int f = 0; f < 3; f++)
#pragma unroll
    for (int k = 0; k < M; k++) {
      const int r = threadIdx.x * M + k;
      dd[f][k] = r < nf[f] ? L.d[f][pad_idx<M>(r)] : 0.;
    }
  LuTab tb[3];
#pragma unroll
  for (int f = 0; f < 3; f++) {
    const double *b = lu + (size_t)(f * 5) * ntm;
    tb[f] = {b, b + ntm, b + 2 * ntm, b + 3 * ntm, b + 4 * ntm};
  }
  if (!a.with_temp) tb[0].r = nullptr;
  lu_solve<NT, M, 3>(dd, tb, L.scr);
#pragma unroll
  for (int f = 0; f < 3; f++)
#pragma unroll
    for (int k = 0; k < M; k++) {
      const int r = threadIdx.x * M + k;
      if (r < nf[f]) L.d[f][pad_idx<M>(r)] = dd[f][k];
    }
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
static void pred_row_launch(hipStream_t s, const PredArgs &a, const double *lu, int ntm) {
  const size_t lds = sizeof(PredRowLds<NT, M>);
  static bool attr = false;
  if (!attr) {
    hipFuncSetAttribute((const void *)k_pred_row<NT, M>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr = true;
  }
  hipLaunchKernelGGL((k_pred_row<NT, M>), dim3(a.g.jtot), dim3(NT), lds, s, a, lu, ntm);
}

template <int NT, int M>
static void lu_factor_launch(hipStream_t s, const LuFactorArgs &a) {
  hipLaunchKernelGGL((k_lu_factor<NT, M>), dim3(3, 3), dim3(NT), 0, s, a);
}

// (NT, M) of the row kernels for a given itot; ntm = NT*M
static int row_config(int itot, int &NT, int &M) {
  M = 4;
  if (itot <= 64 * 4) NT = 64;
  else if (itot <= 256 * 4) NT = 256;
  else if (itot <= 512 * 4) NT = 512;
  else if (itot <= 1024 * 4) NT = 1024;
  else return 0;
  return NT * M;
}

bool pred_impx_supported(int itot) {
  int NT, M;
  return row_config(itot, NT, M) > 0;
}

int lu_table_entries(int itot) {
  int NT, M;
  return row_config(itot, NT, M);
}

void launch_lu_factor(hipStream_t s, const LuFactorArgs &a) {
  int NT, M;
  row_config(a.g.itot, NT, M);
  if (NT == 64) lu_factor_launch<64, 4>(s, a);
  else if (NT == 256) lu_factor_launch<256, 4>(s, a);
  else if (NT == 512) lu_factor_launch<512, 4>(s, a);
  else lu_factor_launch<1024, 4>(s, a);
}

void launch_pred(hipStream_t s, const PredArgs &a, int implicitx, const double *lu, int ntm) {
  const int itot = a.g.itot;
  if (!implicitx) {
    hipLaunchKernelGGL(k_pred_exp, dim3((itot + 255) / 256, a.g.jtot), dim3(256), 0, s, a);
    return;
  }
  int NT, M;
  row_config(itot, NT, M);
  if (NT == 64) pred_row_launch<64, 4>(s, a, lu, ntm);
  else if (NT == 256) pred_row_launch<256, 4>(s, a, lu, ntm);
  else if (NT == 512) pred_row_launch<512, 4>(s, a, lu, ntm);
  else pred_row_launch<1024, 4>(s, a, lu, ntm);
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
