// sns_device.h -- shared device-side definitions for the gfx950 kernels.
//
// Layout contract (the reference's include/arrays/{param,fluid,temperature}.h,
// single rank, jsize == jtot): row-major, x (i) fastest, one halo row on each
// side in y.  Row pitches are Geo::pu / Geo::pc: itot+1 / itot+2 for arrays
// in the reference layout (the per-function entry points), padded to a
// multiple of 16 for the fused state (first interior cell of every row on a
// 128-byte boundary).
//   ux            [(J)*pu + (I-1)]   I = 1..itot+1, J = 0..jtot+1
//   uy, p, psi, T [(J)*pc + (I)]     I = 0..itot+1, J = 0..jtot+1
//   srcux*, tempforcex [(J-1)*(itot-1) + (I-2)]   I = 2..itot
//   srcuy*, srctemp*   [(J-1)*itot + (I-1)]       I = 1..itot
//   dxf[I-1] = DXF(I) (I = 1..itot), dxc[I-1] = DXC(I) (I = 1..itot+1)
//
// Arithmetic in every stencil is written operand-for-operand like the
// reference expression it restates (cited at each function) and the library
// is compiled with -ffp-contract=off, so results are bitwise equal to a
// scalar IEEE evaluation of the same expression.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sns {

// Per-instance constants, passed by value to kernels.  Every derived value
// (square roots, products of them) is evaluated on the host with the same
// operations the reference uses.
struct Geo {
  int itot, jtot;
  int pu, pc;         // row pitches (elements) of ux-type and cell-type arrays
  double dy;
  const double *dxf;  // device, DXF(I) = dxf[I-1]
  const double *dxc;  // device, DXC(I) = dxc[I-1]
  double sqrtPr, sqrtRa;
  double kappa_u;     // sqrt(Pr)/sqrt(Ra)        (src/fluid/update_velocity.c:69)
  double kappa_T;     // 1./sqrt(Pr)/sqrt(Ra)     (src/temperature/update.c:63)
  double two_sqrtRa;  // 2.*sqrt(Ra)              (src/fluid/update_velocity.c:254)
  double two_sqrtPr_sqrtRa; // 2.*sqrt(Pr)*sqrt(Ra) (src/temperature/update.c:142)
  double two_Re;      // 2.*(sqrt(Ra)/sqrt(Pr))   (src/fluid/update_pressure.c:20,36)
};

__device__ __forceinline__ double DXF(const Geo &g, int i) { return g.dxf[i - 1]; }
__device__ __forceinline__ double DXC(const Geo &g, int i) { return g.dxc[i - 1]; }

__device__ __forceinline__ size_t ux_idx(const Geo &g, int i, int j) {
  return (size_t)j * g.pu + (i - 1);
}
__device__ __forceinline__ size_t c_idx(const Geo &g, int i, int j) {  // uy, p, psi, temp
  return (size_t)j * g.pc + i;
}
__device__ __forceinline__ size_t sx_idx(const Geo &g, int i, int j) {  // srcux*, tempforcex
  return (size_t)(j - 1) * (g.itot - 1) + (i - 2);
}
__device__ __forceinline__ size_t sy_idx(const Geo &g, int i, int j) {  // srcuy*, srctemp*
  return (size_t)(j - 1) * g.itot + (i - 1);
}

// ---------------------------------------------------------------- stencils

// Temperature source terms at cell (i, j), i = 1..itot
// (src/temperature/update.c:33-84).
__device__ __forceinline__ void temp_src(const Geo &g, const double *ux, const double *uy,
                                         const double *temp, int i, int j, double w_ex_x,
                                         double w_ex_y, double w_im_x, double w_im_y,
                                         double &srca, double &srcg) {
  const double dy = g.dy;
  const double t_c = temp[c_idx(g, i, j)];
  const double dtdx_xm = (-temp[c_idx(g, i - 1, j)] + t_c) / DXC(g, i);
  const double dtdx_xp = (-t_c + temp[c_idx(g, i + 1, j)]) / DXC(g, i + 1);
  const double dtdy_ym = (-temp[c_idx(g, i, j - 1)] + t_c) / dy;
  const double dtdy_yp = (-t_c + temp[c_idx(g, i, j + 1)]) / dy;
  double adv1;
  {
    const double c_xm = DXC(g, i) / (2. * DXF(g, i));
    const double c_xp = DXC(g, i + 1) / (2. * DXF(g, i));
    const double ux_xm = ux[ux_idx(g, i, j)];
    const double ux_xp = ux[ux_idx(g, i + 1, j)];
    adv1 = -c_xm * ux_xm * dtdx_xm - c_xp * ux_xp * dtdx_xp;
  }
  double adv2;
  {
    const double uy_ym = uy[c_idx(g, i, j)];
    const double uy_yp = uy[c_idx(g, i, j + 1)];
    adv2 = -0.5 * uy_ym * dtdy_ym - 0.5 * uy_yp * dtdy_yp;
  }
  const double dif1 = g.kappa_T * (dtdx_xp - dtdx_xm) / DXF(g, i);
  const double dif2 = g.kappa_T * (dtdy_yp - dtdy_ym) / dy;
  srca = +(adv1 + adv2) + w_ex_x * dif1 + w_ex_y * dif2;
  srcg = +w_im_x * dif1 + w_im_y * dif2;
}

// x-momentum source terms at face (i, j), i = 2..itot
// (src/fluid/update_velocity.c:34-101).  tf is TEMPFORCEX(i, j).
__device__ __forceinline__ void ux_src(const Geo &g, const double *ux, const double *uy,
                                       const double *p, double tf, int i, int j, double w_ex_x,
                                       double w_ex_y, double w_im_x, double w_im_y, double &srca,
                                       double &srcg) {
  const double dy = g.dy;
  const double u_c = ux[ux_idx(g, i, j)];
  const double u_xm = ux[ux_idx(g, i - 1, j)];
  const double u_xp = ux[ux_idx(g, i + 1, j)];
  const double duxdx_xm = (-u_xm + u_c) / DXF(g, i - 1);
  const double duxdx_xp = (-u_c + u_xp) / DXF(g, i);
  const double duxdy_ym = (-ux[ux_idx(g, i, j - 1)] + u_c) / dy;
  const double duxdy_yp = (-u_c + ux[ux_idx(g, i, j + 1)]) / dy;
  const double c_xm = DXF(g, i - 1) / (2. * DXC(g, i));
  const double c_xp = DXF(g, i) / (2. * DXC(g, i));
  double adv1;
  {
    const double ux_xm = 0.5 * u_xm + 0.5 * u_c;
    const double ux_xp = 0.5 * u_c + 0.5 * u_xp;
    adv1 = -c_xm * ux_xm * duxdx_xm - c_xp * ux_xp * duxdx_xp;
  }
  double adv2;
  {
    const double uy_ym = c_xm * uy[c_idx(g, i - 1, j)] + c_xp * uy[c_idx(g, i, j)];
    const double uy_yp = c_xm * uy[c_idx(g, i - 1, j + 1)] + c_xp * uy[c_idx(g, i, j + 1)];
    adv2 = -0.5 * uy_ym * duxdy_ym - 0.5 * uy_yp * duxdy_yp;
  }
  const double dif1 = g.kappa_u * (duxdx_xp - duxdx_xm) / DXC(g, i);
  const double dif2 = g.kappa_u * (duxdy_yp - duxdy_ym) / dy;
  double pre;
  {
    const double p_xm = p[c_idx(g, i - 1, j)];
    const double p_xp = p[c_idx(g, i, j)];
    pre = -(p_xp - p_xm) / DXC(g, i);
  }
  srca = +(adv1 + adv2) + w_ex_x * dif1 + w_ex_y * dif2 + tf;
  srcg = +w_im_x * dif1 + w_im_y * dif2 + pre;
}

// y-momentum source terms at face (i, j), i = 1..itot
// (src/fluid/update_velocity.c:128-189).
__device__ __forceinline__ void uy_src(const Geo &g, const double *ux, const double *uy,
                                       const double *p, int i, int j, double w_ex_x,
                                       double w_ex_y, double w_im_x, double w_im_y, double &srca,
                                       double &srcg) {
  const double dy = g.dy;
  const double v_c = uy[c_idx(g, i, j)];
  const double v_ym = uy[c_idx(g, i, j - 1)];
  const double v_yp = uy[c_idx(g, i, j + 1)];
  const double duydx_xm = (-uy[c_idx(g, i - 1, j)] + v_c) / DXC(g, i);
  const double duydx_xp = (-v_c + uy[c_idx(g, i + 1, j)]) / DXC(g, i + 1);
  const double duydy_ym = (-v_ym + v_c) / dy;
  const double duydy_yp = (-v_c + v_yp) / dy;
  double adv1;
  {
    const double c_xm = DXC(g, i) / (2. * DXF(g, i));
    const double c_xp = DXC(g, i + 1) / (2. * DXF(g, i));
    const double ux_xm = 0.5 * ux[ux_idx(g, i, j - 1)] + 0.5 * ux[ux_idx(g, i, j)];
    const double ux_xp = 0.5 * ux[ux_idx(g, i + 1, j - 1)] + 0.5 * ux[ux_idx(g, i + 1, j)];
    adv1 = -c_xm * ux_xm * duydx_xm - c_xp * ux_xp * duydx_xp;
  }
  double adv2;
  {
    const double uy_ym = 0.5 * v_ym + 0.5 * v_c;
    const double uy_yp = 0.5 * v_c + 0.5 * v_yp;
    adv2 = -0.5 * uy_ym * duydy_ym - 0.5 * uy_yp * duydy_yp;
  }
  const double dif1 = g.kappa_u * (duydx_xp - duydx_xm) / DXF(g, i);
  const double dif2 = g.kappa_u * (duydy_yp - duydy_ym) / dy;
  double pre;
  {
    const double p_ym = p[c_idx(g, i, j - 1)];
    const double p_yp = p[c_idx(g, i, j)];
    pre = -(p_yp - p_ym) / dy;
  }
  srca = +(adv1 + adv2) + w_ex_x * dif1 + w_ex_y * dif2;
  srcg = +w_im_x * dif1 + w_im_y * dif2 + pre;
}

// RK increment alpha*dt*A + beta*dt*B + gamma*dt*G
// (src/temperature/update.c:115-122, src/fluid/update_velocity.c:222-227,374-377).
__device__ __forceinline__ double rk_incr(double alpha, double beta, double gamma, double dt,
                                          double a, double b, double gsrc) {
  return +alpha * dt * a + beta * dt * b + gamma * dt * gsrc;
}

// Tri-diagonal coefficients of the Crank-Nicolson operators in x, row i
// (1-based reference index).  Returns (l, c, u).
// T:  src/temperature/update.c:141-149
__device__ __forceinline__ void cnx_coef_temp(const Geo &g, double prefactor, int i, double &l,
                                              double &c, double &u) {
  const double coef_xm = 1. / DXF(g, i) / DXC(g, i);
  const double coef_xp = 1. / DXF(g, i) / DXC(g, i + 1);
  const double coef_x0 = -coef_xm - coef_xp;
  l = -prefactor * coef_xm;
  u = -prefactor * coef_xp;
  c = 1. - prefactor * coef_x0;
}
// ux (interior rows i = 2..itot): src/fluid/update_velocity.c:253-261
__device__ __forceinline__ void cnx_coef_ux(const Geo &g, double prefactor, int i, double &l,
                                            double &c, double &u) {
  const double coef_xm = 1. / DXC(g, i) / DXF(g, i - 1);
  const double coef_xp = 1. / DXC(g, i) / DXF(g, i);
  const double coef_x0 = -coef_xm - coef_xp;
  l = -prefactor * coef_xm;
  u = -prefactor * coef_xp;
  c = 1. - prefactor * coef_x0;
}
// uy: src/fluid/update_velocity.c:398-406 (same form as T)
__device__ __forceinline__ void cnx_coef_uy(const Geo &g, double prefactor, int i, double &l,
                                            double &c, double &u) {
  cnx_coef_temp(g, prefactor, i, l, c, u);
}

// Field kinds shared by the CN solvers and increment kernels.
enum FieldKind { KIND_TEMP = 0, KIND_UX = 1, KIND_UY = 2 };

}  // namespace sns
