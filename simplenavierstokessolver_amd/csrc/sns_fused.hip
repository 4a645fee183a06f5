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

// ------------------------------------------------------------- predictor

// Row pointers of the substage-start fields around row j (j uniform across
// the workgroup) and of row j of the RK histories.
struct PredRows {
  Rows3 ux, uy, t, p;
  double *hu, *hv, *ht;
};

__device__ __forceinline__ PredRows pred_rows(const PredArgs &a, int j) {
  const Geo &g = a.g;
  PredRows r;
  r.ux = ux_rows(g, a.ux, j);
  r.uy = c_rows(g, a.uy, j);
  r.t = c_rows(g, a.temp, j);
  r.p = c_rows(g, a.p, j);
  r.hu = a.hist_ux + (size_t)(j - 1) * (g.itot - 1);
  r.hv = a.hist_uy + (size_t)(j - 1) * g.itot;
  r.ht = a.hist_t + (size_t)(j - 1) * g.itot;
  return r;
}

// Increment of field f (0: T, 1: ux, 2: uy) at (i, j); the RK history of the
// explicit terms is read (beta != 0) and rewritten in place.
template <bool P2>
__device__ __forceinline__ double pred_field(const PredArgs &a, const PredRows &R, int f, int i,
                                             double dt) {
  const Geo &g = a.g;
  double A, G, B = 0.;
  if (f == 0) {
    temp_src_t<P2>(g, a.tb, R.ux, R.uy, R.t, i, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x, a.tr.w_im_y,
               A, G);
    if (a.read_hist) B = ldr(R.ht, i - 1);
    if (a.write_hist) str(R.ht, i - 1, A);
  } else if (f == 1) {
    double tf = 0.;
    if (a.forcing) tf = 0.5 * ldr(R.t.c, i - 1) + 0.5 * ldr(R.t.c, i);
    ux_src_t<P2>(g, a.tb, R.ux, R.uy, R.p, tf, i, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x,
             a.tr.w_im_y, A, G);
    if (a.read_hist) B = ldr(R.hu, i - 2);
    if (a.write_hist) str(R.hu, i - 2, A);
  } else {
    uy_src_t<P2>(g, a.tb, R.ux, R.uy, R.p, i, a.tr.w_ex_x, a.tr.w_ex_y, a.tr.w_im_x, a.tr.w_im_y, A,
             G);
    if (a.read_hist) B = ldr(R.hv, i - 1);
    if (a.write_hist) str(R.hv, i - 1, A);
  }
  return rk_incr(a.alpha, a.beta, a.gamma, dt, A, B, G);
}

// Final value of field f at (i, j), or its x-aligned increment when a y
// solve follows.
__device__ __forceinline__ void emit_field(const PredArgs &a, const PredRows &R, int f, int i,
                                           int j, double d) {
  const Geo &g = a.g;
  if (f == 0) {
    if (a.q_t) str(a.q_t + (size_t)(j - 1) * g.itot, i - 1, d);
    else put_c<2>(g, a.out_t, i, j, ldr(R.t.c, i) + d);
  } else if (f == 1) {
    if (a.q_u) str(a.q_u + (size_t)(j - 1) * (g.itot - 1), i - 2, d);
    else put_ux(g, a.out_u, i, j, ldr(R.ux.c, i - 1) + d);
  } else {
    if (a.q_v) str(a.q_v + (size_t)(j - 1) * g.itot, i - 1, d);
    else put_c<0>(g, a.out_v, i, j, ldr(R.uy.c, i) + d);
  }
}

// explicit diffusion in x: point-wise, the three fields per thread
template <bool P2, class S>
__global__ void __launch_bounds__(256) k_pred_exp(S src) {
  const PredArgs a = src.get();
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = xcd_remap(blockIdx.y, gridDim.y) + 1;
  if (i > a.g.itot) return;
  const double dt = *a.dtp;
  const PredRows R = pred_rows(a, j);
  if (a.with_temp) emit_field(a, R, 0, i, j, pred_field<P2>(a, R, 0, i, dt));
  if (i >= 2) emit_field(a, R, 1, i, j, pred_field<P2>(a, R, 1, i, dt));
  emit_field(a, R, 2, i, j, pred_field<P2>(a, R, 2, i, dt));
}

// Power-of-two divisors select the variant that divides by multiplying.
template <class S>
static void pred_exp_launch(hipStream_t s, const PredArgs &shape, S src, int nz) {
  const Geo &g = shape.g;
  const dim3 grid((g.itot + 255) / 256, g.jtot, nz);
  if (pow2_division(shape.tb))
    hipLaunchKernelGGL((k_pred_exp<true, S>), grid, dim3(256), 0, s, src);
  else
    hipLaunchKernelGGL((k_pred_exp<false, S>), grid, dim3(256), 0, s, src);
}

// A folded correction (psi set) is only requested with implicit x: the
// explicit-x kernel takes its fields as they are.
void launch_pred(hipStream_t s, const PredArgs &a, int implicitx) {
  if (implicitx) launch_pred_rows(s, a);
  else pred_exp_launch(s, a, ArgOne<PredArgs>{a}, 1);
}

void launch_pred_batch(hipStream_t s, const PredArgs &shape, int implicitx, const PredArgs *dev,
                       int nz) {
  if (implicitx) launch_pred_rows_batch(s, shape, dev, nz);
  else pred_exp_launch(s, shape, arg_batch<PredArgs>(dev), nz);
}

// ------------------------------------------------- correction + pressure

// rows per workgroup of the correction kernel
constexpr int kCorrRows = 2;
// output columns per wave: lanes 1..62; lanes 0 and 63 hold the neighbours
constexpr int kCorrCols = 62;
constexpr int kCorrThreads = 256;

// src/fluid/correct_velocity.c:6-42 and src/fluid/update_pressure.c:7-60 in
// one pass: reads ux*, uy*, p and psi, writes ux, uy, p with boundary
// values.  psi(i-1) and psi(i+1) come from the adjacent lanes; psi of the
// rows above and below is carried in registers; the column metrics are read
// once per thread.  With red != nullptr it also folds the advective bound of
// src/param/decide_dt.c:25-45 over the corrected velocities: each workgroup
// stores the minimum over its cells in red[workgroup] and k_dt reduces them
// (min is exact, so the result does not depend on the order).
template <class S>
__global__ void __launch_bounds__(kCorrThreads) k_corr(S src) {
  const CorrArgs a = src.get();
  const Geo &g = a.g;
  const Tables &tb = a.tb;
  const int itot = g.itot;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = (blockIdx.x * (kCorrThreads / 64) + wave) * kCorrCols + lane;  // lane 0: i-1 halo
  const bool out = lane >= 1 && lane <= kCorrCols && i <= itot;
  const int ci = clampi(i, 0, itot + 1);
  const int ti = clampi(i - 1, 0, itot + 1);  // table entry of column i
  const int j0 = blockIdx.y * kCorrRows + 1;
  const double dt = *a.dtp;
  const double gdt = a.gamma * dt;
  const double fac = -gdt / g.two_Re;
  const size_t wc = (size_t)g.pc, wu = (size_t)g.pu;
  // column metrics; on a uniform grid a wave whose 64 columns all lie in
  // 2..itot-1 takes them as constants (the table entries themselves)
  const int c0 = i - lane;
  const bool interior = tb.uniform && c0 >= 2 && c0 + 63 <= itot - 1;
  double dxc_i, rdxc_i, dxc_p = 0., rdxc_p = 0., dxf_i = 0., rdxf_i = 0.;
  if (interior) {
    dxc_i = dxc_p = tb.u_dxc;
    rdxc_i = rdxc_p = tb.u_rdxc;
    dxf_i = tb.u_dxf;
    rdxf_i = tb.u_rdxf;
  } else {
    dxc_i = ldr(tb.dxc(), ti);  // DXC(i), 1./DXC(i)
    rdxc_i = ldr(tb.rdxc(), ti);
    if (a.tr.implicitx) {
      dxc_p = from_right(dxc_i);  // DXC(i+1)
      rdxc_p = from_right(rdxc_i);
      dxf_i = ldr(tb.dxf(), ti);
      rdxf_i = ldr(tb.rdxf(), ti);
    }
  }
  double m = 5.e-2;
  double psi_s = ldr(a.psi + (size_t)(j0 - 1) * wc, ci);  // psi(i, j-1)
  double psi_c = ldr(a.psi + (size_t)j0 * wc, ci);
#pragma unroll
  for (int r = 0; r < kCorrRows; r++) {
    const int j = j0 + r;
    if (j > g.jtot) break;
    const double psi_n = ldr(a.psi + (size_t)(j + 1) * wc, ci);  // psi(i, j+1)
    const double psi_w = from_left(psi_c);                        // psi(i-1, j)
    const double psi_e = from_right(psi_c);                       // psi(i+1, j)
    const double us = ldr(a.ux_star + (size_t)j * wu, clampi(i - 1, 0, itot));
    const double vs = ldr(a.uy_star + (size_t)j * wc, ci);
    const double pj = ldr(a.p + (size_t)j * wc, ci);
    if (out) {
      if (i >= 2) {
        const double v = proj_ux(us, psi_c, psi_w, gdt, rdxc_i);
        put_ux(g, a.ux, i, j, v);
        if (a.red) m = fmin(m, dxc_i / fmax(fabs(v), 1.e-8));
      }
      {
        const double v = proj_uy(vs, psi_c, psi_s, gdt, tb.dy, tb.rdy);
        put_c<0>(g, a.uy, i, j, v);
        if (a.red) m = fmin(m, tb.dy / fmax(fabs(v), 1.e-8));
      }
      const ProjX x = {dxc_i, rdxc_i, dxc_p, rdxc_p, dxf_i, rdxf_i};
      const double pv = proj_p(pj, psi_c, psi_w, psi_e, psi_s, psi_n, x, tb.dy, tb.rdy, fac,
                               a.tr.implicitx, a.tr.implicity);
      put_c<1>(g, a.p, i, j, pv);
    }
    psi_s = psi_c;
    psi_c = psi_n;
  }
  if (a.red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
    __shared__ double wm[kCorrThreads / 64];
    if (lane == 0) wm[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = wm[0];
      for (int w = 1; w < kCorrThreads / 64; w++) r = fmin(r, wm[w]);
      a.red[blockIdx.y * gridDim.x + blockIdx.x] = r;
    }
  }
}

static dim3 corr_grid(const Geo &g) {
  const int cols = (kCorrThreads / 64) * kCorrCols;
  return dim3((g.itot + cols - 1) / cols, (g.jtot + kCorrRows - 1) / kCorrRows);
}

int corr_partials(const Geo &g) {
  const dim3 gr = corr_grid(g);
  return (int)(gr.x * gr.y);
}

void launch_corr(hipStream_t s, const CorrArgs &a) {
  hipLaunchKernelGGL(k_corr<ArgOne<CorrArgs>>, corr_grid(a.g), dim3(kCorrThreads), 0, s,
                     ArgOne<CorrArgs>{a});
}

static dim3 with_z(dim3 d, int nz) { return dim3(d.x, d.y, nz); }

void launch_corr_batch(hipStream_t s, const CorrArgs &shape, const CorrArgs *dev, int nz) {
  hipLaunchKernelGGL(k_corr<ArgBatch<CorrArgs>>, with_z(corr_grid(shape.g), nz),
                     dim3(kCorrThreads), 0, s, arg_batch<CorrArgs>(dev));
}

// The advective bound of src/param/decide_dt.c:25-45 over ux, uy as they
// are, with the tiling and the expressions of the reduction in k_corr: the
// minimum over the cells of workgroup w goes to red[w], so a time-step
// decision from these partials equals one from the partials the previous
// step's last correction stored.
template <class S>
__global__ void __launch_bounds__(kCorrThreads) k_adv_bound(S src) {
  const CorrArgs a = src.get();
  const Geo &g = a.g;
  const Tables &tb = a.tb;
  const int itot = g.itot;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = (blockIdx.x * (kCorrThreads / 64) + wave) * kCorrCols + lane;
  const bool out = lane >= 1 && lane <= kCorrCols && i <= itot;
  const int ci = clampi(i, 0, itot + 1);
  const int ti = clampi(i - 1, 0, itot + 1);
  const int j0 = blockIdx.y * kCorrRows + 1;
  const size_t wc = (size_t)g.pc, wu = (size_t)g.pu;
  const double dxc_i = ldr(tb.dxc(), ti);
  double m = 5.e-2;
#pragma unroll
  for (int r = 0; r < kCorrRows; r++) {
    const int j = j0 + r;
    if (j > g.jtot) break;
    const double u = ldr(a.ux + (size_t)j * wu, clampi(i - 1, 0, itot));
    const double v = ldr(a.uy + (size_t)j * wc, ci);
    if (out) {
      if (i >= 2) m = fmin(m, dxc_i / fmax(fabs(u), 1.e-8));
      m = fmin(m, tb.dy / fmax(fabs(v), 1.e-8));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
  __shared__ double wm[kCorrThreads / 64];
  if (lane == 0) wm[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = wm[0];
    for (int w = 1; w < kCorrThreads / 64; w++) r = fmin(r, wm[w]);
    a.red[blockIdx.y * gridDim.x + blockIdx.x] = r;
  }
}

void launch_adv_bound(hipStream_t s, const CorrArgs &a) {
  hipLaunchKernelGGL(k_adv_bound<ArgOne<CorrArgs>>, corr_grid(a.g), dim3(kCorrThreads), 0, s,
                     ArgOne<CorrArgs>{a});
}

void launch_adv_bound_batch(hipStream_t s, const CorrArgs &shape, const CorrArgs *dev, int nz) {
  hipLaunchKernelGGL(k_adv_bound<ArgBatch<CorrArgs>>, with_z(corr_grid(shape.g), nz),
                     dim3(kCorrThreads), 0, s, arg_batch<CorrArgs>(dev));
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

template <class S>
__global__ void __launch_bounds__(kDtThreads) k_dt(S src) {
  const DtArgs a = src.get();
  const double *part = a.part;
  // eight independent loads in flight per thread (the minimum is exact, so
  // the grouping does not change it)
  constexpr int U = 8;
  double mu[U];
#pragma unroll
  for (int u = 0; u < U; u++) mu[u] = part[0];
  int k = threadIdx.x;
  for (; k + (U - 1) * kDtThreads < a.n; k += U * kDtThreads) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = part[k + u * kDtThreads];
#pragma unroll
    for (int u = 0; u < U; u++) mu[u] = fmin(mu[u], v[u]);
  }
  for (; k < a.n; k += kDtThreads) mu[0] = fmin(mu[0], part[k]);
  double m = mu[0];
#pragma unroll
  for (int u = 1; u < U; u++) m = fmin(m, mu[u]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
  __shared__ double wm[kDtThreads / 64];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < kDtThreads / 64; w++) m = fmin(m, wm[w]);
  m = fmin(m, wm[0]);
  const double dt_adv = m * 0.7;
  const double dx = a.dt_dif_x, dy = a.dt_dif_y;
  const double d = a.implicitx ? a.implicity ? dt_adv : fmin(dt_adv, dy)
                               : a.implicity ? fmin(dt_adv, dx) : fmin(dt_adv, fmin(dx, dy));
  *a.dt = d;
  *a.time += d;
}

void launch_dt(hipStream_t s, const DtArgs &a) {
  hipLaunchKernelGGL(k_dt<ArgOne<DtArgs>>, dim3(1), dim3(kDtThreads), 0, s, ArgOne<DtArgs>{a});
}

void launch_dt_batch(hipStream_t s, const DtArgs *dev, int nz) {
  hipLaunchKernelGGL(k_dt<ArgBatch<DtArgs>>, dim3(1, 1, nz), dim3(kDtThreads), 0, s,
                     arg_batch<DtArgs>(dev));
}

}  // namespace sns
