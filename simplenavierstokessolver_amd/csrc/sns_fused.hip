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

// Increment of field f (0: T, 1: ux, 2: uy) at (i, j); the RK history of the
// explicit terms is read (beta != 0) and rewritten in place.
__device__ __forceinline__ double pred_field(const PredArgs &a, int f, int i, int j, double dt) {
  const Geo &g = a.g;
  double A, G, B = 0.;
  if (f == 0) {
    temp_src_t(g, a.tb, a.ux, a.uy, a.temp, i, j, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x,
               a.tr.w_im_y, A, G);
    const size_t s = sy_idx(g, i, j);
    if (a.read_hist) B = a.hist_t[s];
    if (a.write_hist) a.hist_t[s] = A;
  } else if (f == 1) {
    double tf = 0.;
    if (a.forcing) tf = 0.5 * a.temp[c_idx(g, i - 1, j)] + 0.5 * a.temp[c_idx(g, i, j)];
    ux_src_t(g, a.tb, a.ux, a.uy, a.p, tf, i, j, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x,
             a.tr.w_im_y, A, G);
    const size_t s = sx_idx(g, i, j);
    if (a.read_hist) B = a.hist_ux[s];
    if (a.write_hist) a.hist_ux[s] = A;
  } else {
    uy_src_t(g, a.tb, a.ux, a.uy, a.p, i, j, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x, a.tr.w_im_y,
             A, G);
    const size_t s = sy_idx(g, i, j);
    if (a.read_hist) B = a.hist_uy[s];
    if (a.write_hist) a.hist_uy[s] = A;
  }
  return rk_incr(a.alpha, a.beta, a.gamma, dt, A, B, G);
}

// Final value of field f at (i, j), or its x-aligned increment when a y
// solve follows.
__device__ __forceinline__ void emit_field(const PredArgs &a, int f, int i, int j, double d) {
  const Geo &g = a.g;
  if (f == 0) {
    if (a.q_t) a.q_t[sy_idx(g, i, j)] = d;
    else put_c<2>(g, a.out_t, i, j, a.temp[c_idx(g, i, j)] + d);
  } else if (f == 1) {
    if (a.q_u) a.q_u[sx_idx(g, i, j)] = d;
    else put_ux(g, a.out_u, i, j, a.ux[ux_idx(g, i, j)] + d);
  } else {
    if (a.q_v) a.q_v[sy_idx(g, i, j)] = d;
    else put_c<0>(g, a.out_v, i, j, a.uy[c_idx(g, i, j)] + d);
  }
}

// Crank-Nicolson in x: one workgroup per (row, field).  The increments are
// staged in LDS: lane order for the stencils, chunk order (M unknowns per
// thread) for the line solve.
template <int NT, int M>
struct PredRowLds {
  double d[NT * (M + 1)];  // padded: chunk t starts at t*(M+1)
  double scr[2 * (NT / 64)];
};

template <int M>
__device__ __forceinline__ int pad_idx(int r) {
  return r + r / M;
}

// Workgroup w covers row w / nfld + 1 and field w % nfld (+1 without the
// temperature equation); consecutive w share an XCD, so the fields of a row
// and neighbouring rows reuse one L2.
template <int NT, int M>
__global__ void __launch_bounds__(NT) k_pred_row(PredArgs a, const double *lu, int ntm, int nfld) {
  extern __shared__ double lds_raw[];
  PredRowLds<NT, M> &L = *reinterpret_cast<PredRowLds<NT, M> *>(lds_raw);
  const Geo &g = a.g;
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int j = w / nfld + 1;
  const int f = w % nfld + (nfld == 3 ? 0 : 1);
  const int n = f == 1 ? g.itot - 1 : g.itot;
  const int i0 = f == 1 ? 2 : 1;
  const double dt = *a.dtp;
  for (int ii = threadIdx.x; ii < n; ii += NT) L.d[pad_idx<M>(ii)] = pred_field(a, f, ii + i0, j, dt);
  __syncthreads();
  double dd[1][M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    dd[0][k] = r < n ? L.d[pad_idx<M>(r)] : 0.;
  }
  const double *b = lu + (size_t)(f * 5) * ntm;
  const LuTab tb[1] = {{b, b + ntm, b + 2 * ntm, b + 3 * ntm, b + 4 * ntm}};
  lu_solve<NT, M, 1>(dd, tb, L.scr);
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r < n) L.d[pad_idx<M>(r)] = dd[0][k];
  }
  __syncthreads();
  for (int ii = threadIdx.x; ii < n; ii += NT) emit_field(a, f, ii + i0, j, L.d[pad_idx<M>(ii)]);
}

template <int NT, int M>
static void pred_row_launch(hipStream_t s, const PredArgs &a, const double *lu, int ntm) {
  const size_t lds = sizeof(PredRowLds<NT, M>);
  const int nfld = a.with_temp ? 3 : 2;
  hipLaunchKernelGGL((k_pred_row<NT, M>), dim3(a.g.jtot * nfld), dim3(NT), lds, s, a, lu, ntm,
                     nfld);
}

template <int NT, int M>
static void lu_factor_launch(hipStream_t s, const LuFactorArgs &a) {
  hipLaunchKernelGGL((k_lu_factor<NT, M>), dim3(3, 3), dim3(NT), 0, s, a);
}

// (NT, M) of the row kernels for a given itot; ntm = NT*M >= itot
static constexpr int kRowM = 8;
static int row_config(int itot, int &NT, int &M) {
  M = kRowM;
  if (itot <= 64 * kRowM) NT = 64;
  else if (itot <= 128 * kRowM) NT = 128;
  else if (itot <= 256 * kRowM) NT = 256;
  else if (itot <= 512 * kRowM) NT = 512;
  else if (itot <= 1024 * kRowM) NT = 1024;
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
  switch (NT) {
    case 64: lu_factor_launch<64, kRowM>(s, a); break;
    case 128: lu_factor_launch<128, kRowM>(s, a); break;
    case 256: lu_factor_launch<256, kRowM>(s, a); break;
    case 512: lu_factor_launch<512, kRowM>(s, a); break;
    default: lu_factor_launch<1024, kRowM>(s, a); break;
  }
}

void launch_pred(hipStream_t s, const PredArgs &a, int implicitx, const double *lu, int ntm) {
  const int itot = a.g.itot;
  if (!implicitx) {
    hipLaunchKernelGGL(k_pred_exp, dim3((itot + 255) / 256, a.g.jtot), dim3(256), 0, s, a);
    return;
  }
  int NT, M;
  row_config(itot, NT, M);
  switch (NT) {
    case 64: pred_row_launch<64, kRowM>(s, a, lu, ntm); break;
    case 128: pred_row_launch<128, kRowM>(s, a, lu, ntm); break;
    case 256: pred_row_launch<256, kRowM>(s, a, lu, ntm); break;
    case 512: pred_row_launch<512, kRowM>(s, a, lu, ntm); break;
    default: pred_row_launch<1024, kRowM>(s, a, lu, ntm); break;
  }
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

// rows per workgroup of the correction kernel
constexpr int kCorrRows = 4;

// src/fluid/correct_velocity.c:6-42 and src/fluid/update_pressure.c:7-60 in
// one pass: reads ux*, uy*, p and psi, writes ux, uy, p with boundary
// values.  With red != nullptr it also folds the advective bound of
// src/param/decide_dt.c:25-45 over the corrected velocities: each workgroup
// stores the minimum over its cells in red[workgroup] and k_dt reduces them
// (min is exact, so the result does not depend on the order).
__global__ void __launch_bounds__(256) k_corr(CorrArgs a) {
  const Geo &g = a.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j0 = blockIdx.y * kCorrRows + 1;
  const double dt = *a.dtp;
  const double gdt = a.gamma * dt;
  const double fac = -gdt / g.two_Re;
  double m = 5.e-2;
  if (i <= g.itot) {
    double psi_s = a.psi[c_idx(g, i, j0 - 1)];  // psi(i, j-1)
    double psi_c = a.psi[c_idx(g, i, j0)];
#pragma unroll
    for (int r = 0; r < kCorrRows; r++) {
      const int j = j0 + r;
      if (j > g.jtot) break;
      const double psi_n = a.psi[c_idx(g, i, j + 1)];  // psi(i, j+1)
      if (i >= 2) {
        const double dxi = a.tb.rdxc[i - 1];  // 1./DXC(i)
        const double v = a.ux_star[ux_idx(g, i, j)] - gdt * (psi_c - a.psi[c_idx(g, i - 1, j)]) * dxi;
        put_ux(g, a.ux, i, j, v);
        if (a.red) m = fmin(m, a.tb.dxc[i - 1] / fmax(fabs(v), 1.e-8));
      }
      {
        const double v = a.uy_star[c_idx(g, i, j)] - div_cr(gdt * (psi_c - psi_s), a.tb.dy, a.tb.rdy);
        put_c<0>(g, a.uy, i, j, v);
        if (a.red) m = fmin(m, a.tb.dy / fmax(fabs(v), 1.e-8));
      }
      double pv = a.p[c_idx(g, i, j)] + psi_c;
      if (a.tr.implicitx) {
        const double dm = div_cr(-a.psi[c_idx(g, i - 1, j)] + psi_c, a.tb.dxc[i - 1], a.tb.rdxc[i - 1]);
        const double dp = div_cr(-psi_c + a.psi[c_idx(g, i + 1, j)], a.tb.dxc[i], a.tb.rdxc[i]);
        pv += fac * (+div_cr(dp - dm, a.tb.dxf[i - 1], a.tb.rdxf[i - 1]));
      }
      if (a.tr.implicity) {
        const double dm = div_cr(-psi_s + psi_c, a.tb.dy, a.tb.rdy);
        const double dp = div_cr(-psi_c + psi_n, a.tb.dy, a.tb.rdy);
        pv += fac * (+div_cr(dp - dm, a.tb.dy, a.tb.rdy));
      }
      put_c<1>(g, a.p, i, j, pv);
      psi_s = psi_c;
      psi_c = psi_n;
    }
  }
  if (a.red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
    __shared__ double wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
      a.red[blockIdx.y * gridDim.x + blockIdx.x] = fmin(fmin(wm[0], wm[1]), fmin(wm[2], wm[3]));
  }
}

int corr_partials(const Geo &g) {
  return ((g.itot + 255) / 256) * ((g.jtot + kCorrRows - 1) / kCorrRows);
}

void launch_corr(hipStream_t s, const CorrArgs &a) {
  hipLaunchKernelGGL(k_corr, dim3((a.g.itot + 255) / 256, (a.g.jtot + kCorrRows - 1) / kCorrRows),
                     dim3(256), 0, s, a);
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
// advective bound (minimum of the n partial minima in part) times 0.7, the
// diffusive bounds (fixed per run, computed on the host) and the treatment's
// dt; the run time advances by dt (src/main.c:34-35).  One workgroup of
// kDtThreads threads.
constexpr int kDtThreads = 1024;

__global__ void __launch_bounds__(kDtThreads) k_dt(const double *part, int n, double *dt,
                                                   double *time, double dt_dif_x, double dt_dif_y,
                                                   int implicitx, int implicity) {
  double m = part[0];
  for (int k = threadIdx.x; k < n; k += kDtThreads) m = fmin(m, part[k]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
  __shared__ double wm[kDtThreads / 64];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < kDtThreads / 64; w++) m = fmin(m, wm[w]);
  m = fmin(m, wm[0]);
  const double dt_adv = m * 0.7;
  const double d = implicitx ? implicity ? dt_adv : fmin(dt_adv, dt_dif_y)
                             : implicity ? fmin(dt_adv, dt_dif_x)
                                         : fmin(dt_adv, fmin(dt_dif_x, dt_dif_y));
  *dt = d;
  *time += d;
}

void launch_dt(hipStream_t s, const double *part, int n, double *dt, double *time,
               double dt_dif_x, double dt_dif_y, int implicitx, int implicity) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(kDtThreads), 0, s, part, n, dt, time, dt_dif_x, dt_dif_y,
                     implicitx, implicity);
}

}  // namespace sns
