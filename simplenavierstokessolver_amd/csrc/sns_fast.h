// sns_fast.h -- table-driven stencils for the fused kernels (gfx950).
//
// The reference divides by grid spacings (DXF, DXC, dy) at every point.
// Here each divisor b has its reciprocal r = RN(1/b) tabulated on the host,
// and a/b is formed as
//   q0 = RN(a*r), q1 = RN(q0 + r*(a - b*q0)), q2 = RN(q1 + r*(a - b*q1)).
// q0 is within two ulps of a/b, q1 within one ulp, and by Markstein's
// theorem (IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, Thm. 4.10) q2 = RN(a/b): the quotient is the
// correctly rounded IEEE one, bit for bit, in five FMA-class instructions
// instead of the generic division sequence.  Zero quotients keep the sign
// of q0, which is the sign of a/b.
//
// All other per-column constants (interpolation weights, Crank-Nicolson
// coefficients) are tabulated on the host with the reference's operation
// order, so every stencil below returns the same bits as its restatement
// in sns_device.h.
#pragma once
#include <hip/hip_runtime.h>
#include "sns_device.h"

namespace sns {

__device__ __forceinline__ double div_cr(double a, double b, double r) {
  const double q0 = a * r;
  const double e0 = __builtin_fma(-b, q0, a);
  const double q1 = __builtin_fma(r, e0, q0);
  const double e1 = __builtin_fma(-b, q1, a);
  const double q2 = __builtin_fma(r, e1, q1);
  return q0 == 0.0 ? q0 : q2;
}

// Per-column tables, all indexed like the reference arrays they derive from
// (entry I-1 holds the value for column I unless noted).
struct Tables {
  const double *dxf, *dxc;     // DXF(I), DXC(I)
  const double *rdxf, *rdxc;   // 1./DXF(I), 1./DXC(I)
  const double *cT_xm, *cT_xp; // DXC(I)/(2.*DXF(I)), DXC(I+1)/(2.*DXF(I))   I = 1..itot
  const double *cU_xm, *cU_xp; // DXF(I-1)/(2.*DXC(I)), DXF(I)/(2.*DXC(I))   I = 2..itot (entry I-1)
  const double *nT_xm, *nT_xp; // 1./DXF(I)/DXC(I), 1./DXF(I)/DXC(I+1)       (CN of T and uy)
  const double *nU_xm, *nU_xp; // 1./DXC(I)/DXF(I-1), 1./DXC(I)/DXF(I)       (CN of ux)
  const double *pl, *pu;       // 1./DXC(I)/DXF(I), 1./DXC(I+1)/DXF(I)       (Poisson)
  double dy, rdy;
};

// Temperature sources at (i, j): src/temperature/update.c:33-84
__device__ __forceinline__ void temp_src_t(const Geo &g, const Tables &tb, const double *ux,
                                           const double *uy, const double *temp, int i, int j,
                                           double wex, double wey, double wix, double wiy,
                                           double &srca, double &srcg) {
  const double dy = tb.dy, rdy = tb.rdy;
  const double dxc_i = tb.dxc[i - 1], dxc_ip = tb.dxc[i], dxf_i = tb.dxf[i - 1];
  const double t_c = temp[c_idx(g, i, j)];
  const double dtdx_xm = div_cr(-temp[c_idx(g, i - 1, j)] + t_c, dxc_i, tb.rdxc[i - 1]);
  const double dtdx_xp = div_cr(-t_c + temp[c_idx(g, i + 1, j)], dxc_ip, tb.rdxc[i]);
  const double dtdy_ym = div_cr(-temp[c_idx(g, i, j - 1)] + t_c, dy, rdy);
  const double dtdy_yp = div_cr(-t_c + temp[c_idx(g, i, j + 1)], dy, rdy);
  const double c_xm = tb.cT_xm[i - 1], c_xp = tb.cT_xp[i - 1];
  const double adv1 = -c_xm * ux[ux_idx(g, i, j)] * dtdx_xm - c_xp * ux[ux_idx(g, i + 1, j)] * dtdx_xp;
  const double adv2 = -0.5 * uy[c_idx(g, i, j)] * dtdy_ym - 0.5 * uy[c_idx(g, i, j + 1)] * dtdy_yp;
  const double dif1 = div_cr(g.kappa_T * (dtdx_xp - dtdx_xm), dxf_i, tb.rdxf[i - 1]);
  const double dif2 = div_cr(g.kappa_T * (dtdy_yp - dtdy_ym), dy, rdy);
  srca = +(adv1 + adv2) + wex * dif1 + wey * dif2;
  srcg = +wix * dif1 + wiy * dif2;
}

// x-momentum sources at (i, j), i = 2..itot: src/fluid/update_velocity.c:34-101
__device__ __forceinline__ void ux_src_t(const Geo &g, const Tables &tb, const double *ux,
                                         const double *uy, const double *p, double tf, int i,
                                         int j, double wex, double wey, double wix, double wiy,
                                         double &srca, double &srcg) {
  const double dy = tb.dy, rdy = tb.rdy;
  const double u_c = ux[ux_idx(g, i, j)];
  const double u_xm = ux[ux_idx(g, i - 1, j)];
  const double u_xp = ux[ux_idx(g, i + 1, j)];
  const double duxdx_xm = div_cr(-u_xm + u_c, tb.dxf[i - 2], tb.rdxf[i - 2]);
  const double duxdx_xp = div_cr(-u_c + u_xp, tb.dxf[i - 1], tb.rdxf[i - 1]);
  const double duxdy_ym = div_cr(-ux[ux_idx(g, i, j - 1)] + u_c, dy, rdy);
  const double duxdy_yp = div_cr(-u_c + ux[ux_idx(g, i, j + 1)], dy, rdy);
  const double c_xm = tb.cU_xm[i - 1], c_xp = tb.cU_xp[i - 1];
  const double ux_xm = 0.5 * u_xm + 0.5 * u_c;
  const double ux_xp = 0.5 * u_c + 0.5 * u_xp;
  const double adv1 = -c_xm * ux_xm * duxdx_xm - c_xp * ux_xp * duxdx_xp;
  const double uy_ym = c_xm * uy[c_idx(g, i - 1, j)] + c_xp * uy[c_idx(g, i, j)];
  const double uy_yp = c_xm * uy[c_idx(g, i - 1, j + 1)] + c_xp * uy[c_idx(g, i, j + 1)];
  const double adv2 = -0.5 * uy_ym * duxdy_ym - 0.5 * uy_yp * duxdy_yp;
  const double dxc_i = tb.dxc[i - 1], rdxc_i = tb.rdxc[i - 1];
  const double dif1 = div_cr(g.kappa_u * (duxdx_xp - duxdx_xm), dxc_i, rdxc_i);
  const double dif2 = div_cr(g.kappa_u * (duxdy_yp - duxdy_ym), dy, rdy);
  const double pre = div_cr(-(p[c_idx(g, i, j)] - p[c_idx(g, i - 1, j)]), dxc_i, rdxc_i);
  srca = +(adv1 + adv2) + wex * dif1 + wey * dif2 + tf;
  srcg = +wix * dif1 + wiy * dif2 + pre;
}

// y-momentum sources at (i, j), i = 1..itot: src/fluid/update_velocity.c:128-189
__device__ __forceinline__ void uy_src_t(const Geo &g, const Tables &tb, const double *ux,
                                         const double *uy, const double *p, int i, int j,
                                         double wex, double wey, double wix, double wiy,
                                         double &srca, double &srcg) {
  const double dy = tb.dy, rdy = tb.rdy;
  const double v_c = uy[c_idx(g, i, j)];
  const double v_ym = uy[c_idx(g, i, j - 1)];
  const double v_yp = uy[c_idx(g, i, j + 1)];
  const double duydx_xm = div_cr(-uy[c_idx(g, i - 1, j)] + v_c, tb.dxc[i - 1], tb.rdxc[i - 1]);
  const double duydx_xp = div_cr(-v_c + uy[c_idx(g, i + 1, j)], tb.dxc[i], tb.rdxc[i]);
  const double duydy_ym = div_cr(-v_ym + v_c, dy, rdy);
  const double duydy_yp = div_cr(-v_c + v_yp, dy, rdy);
  const double c_xm = tb.cT_xm[i - 1], c_xp = tb.cT_xp[i - 1];
  const double ux_xm = 0.5 * ux[ux_idx(g, i, j - 1)] + 0.5 * ux[ux_idx(g, i, j)];
  const double ux_xp = 0.5 * ux[ux_idx(g, i + 1, j - 1)] + 0.5 * ux[ux_idx(g, i + 1, j)];
  const double adv1 = -c_xm * ux_xm * duydx_xm - c_xp * ux_xp * duydx_xp;
  const double uy_ym = 0.5 * v_ym + 0.5 * v_c;
  const double uy_yp = 0.5 * v_c + 0.5 * v_yp;
  const double adv2 = -0.5 * uy_ym * duydy_ym - 0.5 * uy_yp * duydy_yp;
  const double dif1 = div_cr(g.kappa_u * (duydx_xp - duydx_xm), tb.dxf[i - 1], tb.rdxf[i - 1]);
  const double dif2 = div_cr(g.kappa_u * (duydy_yp - duydy_ym), dy, rdy);
  const double pre = div_cr(-(p[c_idx(g, i, j)] - p[c_idx(g, i, j - 1)]), dy, rdy);
  srca = +(adv1 + adv2) + wex * dif1 + wey * dif2;
  srcg = +wix * dif1 + wiy * dif2 + pre;
}

// Block index -> work index so that consecutive work items run on the same
// XCD (blocks are dispatched round-robin over the 8 XCDs, each with its own
// L2); neighbouring rows / strips then share that L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  constexpr int NX = 8;
  const int per = nblocks / NX, rem = nblocks % NX;
  const int xcd = b % NX, slot = b / NX;
  return xcd * per + (xcd < rem ? xcd : rem) + slot;
}

}  // namespace sns
