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
// instead of the generic division sequence.  The result takes the sign of
// q0, which is the sign of a/b (this only matters for zero quotients).
//
// All other per-column constants (interpolation weights, Crank-Nicolson
// coefficients) are tabulated on the host with the reference's operation
// order, so every stencil below returns the same bits as its restatement
// in sns_device.h.
#pragma once
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime.h>
#include "sns_device.h"

namespace sns {

__device__ __forceinline__ double div_cr(double a, double b, double r) {
  const double q0 = a * r;
  const double e0 = __builtin_fma(-b, q0, a);
  const double q1 = __builtin_fma(r, e0, q0);
  const double e1 = __builtin_fma(-b, q1, a);
  const double q2 = __builtin_fma(r, e1, q1);
  // q2 carries the sign of a/b except for a zero quotient, where the
  // refinement may turn -0 into +0; q0 = a*r has the sign of a/b always
  return __builtin_copysign(q2, q0);
}

// a/b for a divisor b with reciprocal r = 1/b.  P2: b is a power of two,
// so r is exact and a*r is the correctly rounded quotient bit for bit
// (subnormal and overflowing results included: both round the same real
// number a*2^-k); otherwise the exact division above.
template <bool P2>
__device__ __forceinline__ double dv(double a, double b, double r) {
  return P2 ? a * r : div_cr(a, b, r);
}

// ---------------------------------------------------------- lane shifts

// Value of v in lane l-1 (lane 0: 0).
__device__ __forceinline__ double from_left(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Value of v in lane l+1 (lane 63: 0).
__device__ __forceinline__ double from_right(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// Per-column tables, all indexed like the reference arrays they derive from
// (entry I-1 holds the value for column I unless noted).  They live in one
// allocation of 14 double tables and one of 4 double2 tables, each table
// `stride` entries long, so an instance's tables cost two pointers.
// number of double and of double2 tables
constexpr int kTables = 14, kTables2 = 4;

struct Tables {
  const double *t;             // 14 tables, in the order of the accessors below
  const double2 *t2;           // 4 paired tables
  long stride;
  double dy, rdy;
  // Set when the grid is uniform bit for bit away from the walls: every
  // column I = 2..itot-1 has the same DXF, DXC and interpolation weights
  // (u_*).  A wave whose columns all lie there uses these values instead of
  // loading the paired tables; the values are the table entries themselves.
  int uniform;
  double u_dxc, u_rdxc, u_dxf, u_rdxf, u_cT, u_cU;
  // Set when every divisor of the stencils (DXF, DXC of all columns and dy)
  // is a power of two: the kernels then divide by multiplying with the
  // tabulated reciprocal (see dv).
  int pow2;
  // Set when the Poisson row coefficients the x substitutions use are each
  // one value bit for bit: pl(I) for rows 1..itot-1 (plc) and pu(I) for rows
  // 0..itot-2 (puc), 0-based; the x solve then takes them as constants.
  int plu_const;
  double plc, puc;
  __host__ __device__ const double *tab(int k) const { return t + k * stride; }
  __host__ __device__ const double *dxf() const { return tab(0); }    // DXF(I)
  __host__ __device__ const double *dxc() const { return tab(1); }    // DXC(I)
  __host__ __device__ const double *rdxf() const { return tab(2); }   // 1./DXF(I)
  __host__ __device__ const double *rdxc() const { return tab(3); }   // 1./DXC(I)
  // DXC(I)/(2.*DXF(I)), DXC(I+1)/(2.*DXF(I))   I = 1..itot
  __host__ __device__ const double *cT_xm() const { return tab(4); }
  __host__ __device__ const double *cT_xp() const { return tab(5); }
  // DXF(I-1)/(2.*DXC(I)), DXF(I)/(2.*DXC(I))   I = 2..itot (entry I-1)
  __host__ __device__ const double *cU_xm() const { return tab(6); }
  __host__ __device__ const double *cU_xp() const { return tab(7); }
  // 1./DXF(I)/DXC(I), 1./DXF(I)/DXC(I+1)       (CN of T and uy)
  __host__ __device__ const double *nT_xm() const { return tab(8); }
  __host__ __device__ const double *nT_xp() const { return tab(9); }
  // 1./DXC(I)/DXF(I-1), 1./DXC(I)/DXF(I)       (CN of ux)
  __host__ __device__ const double *nU_xm() const { return tab(10); }
  __host__ __device__ const double *nU_xp() const { return tab(11); }
  // 1./DXC(I)/DXF(I), 1./DXC(I+1)/DXF(I)       (Poisson)
  __host__ __device__ const double *pl() const { return tab(12); }
  __host__ __device__ const double *pu() const { return tab(13); }
  // the same values in pairs, for one 16-byte load each (entry as above)
  __host__ __device__ const double2 *dxc2() const { return t2; }               // (DXC, 1./DXC)
  __host__ __device__ const double2 *dxf2() const { return t2 + stride; }      // (DXF, 1./DXF)
  __host__ __device__ const double2 *cT2() const { return t2 + 2 * stride; }   // (cT_xm, cT_xp)
  __host__ __device__ const double2 *cU2() const { return t2 + 3 * stride; }   // (cU_xm, cU_xp)
};

// Host: whether the kernels take the power-of-two division path for these
// tables (SNSG_EXACT_DIV=1 keeps the division sequence; read at every
// launch, both paths return the same bits).
inline bool pow2_division(const Tables &tb) {
  const char *ex = getenv("SNSG_EXACT_DIV");
  return tb.pow2 && !(ex && strcmp(ex, "1") == 0);
}

__device__ __forceinline__ double2 ldr2(const double2 *row, int k) {
  return *reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(row) +
                                            ((unsigned)k << 4));
}

// Loads and stores through a row base (uniform across the workgroup, so it
// stays in scalar registers) plus a 32-bit byte offset: the SGPR-base +
// VGPR-offset addressing form, one VGPR per address instead of a 64-bit
// pair.  k >= 0 and k*8 < 2^32.
__device__ __forceinline__ double ldr(const double *row, int k) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(row) +
                                           ((unsigned)k << 3));
}
__device__ __forceinline__ void str(double *row, int k, double v) {
  *reinterpret_cast<double *>(reinterpret_cast<char *>(row) + ((unsigned)k << 3)) = v;
}
// Rows j-1, j, j+1 of a field: pointers to the first stored element of each
// row (I = 0 for uy, p, psi, T; I = 1 for ux).
struct Rows3 {
  const double *m, *c, *p;
};
__device__ __forceinline__ Rows3 ux_rows(const Geo &g, const double *f, int j) {
  const size_t w = (size_t)g.pu;
  return {f + (size_t)(j - 1) * w, f + (size_t)j * w, f + (size_t)(j + 1) * w};
}
__device__ __forceinline__ Rows3 c_rows(const Geo &g, const double *f, int j) {
  const size_t w = (size_t)g.pc;
  return {f + (size_t)(j - 1) * w, f + (size_t)j * w, f + (size_t)(j + 1) * w};
}

// ------------------------------------------------------------ source terms
//
// Each source term is written once, on the values around the point; the
// loaders below (pointer form) and the row kernels (values shared across
// lanes and rows) feed it.  Operation order follows the reference
// expression by expression.

// Temperature at (i, j) and its neighbours, with the velocities it is
// advected by: src/temperature/update.c:33-84
struct TempPt {
  double t_c, t_xm, t_xp, t_ym, t_yp;  // T(i,j), T(i-1,j), T(i+1,j), T(i,j-1), T(i,j+1)
  double ux_c, ux_p;                   // UX(i,j), UX(i+1,j)
  double uy_c, uy_p;                   // UY(i,j), UY(i,j+1)
};
// Column metrics of cell i: DXC(i), DXC(i+1), DXF(i) with reciprocals and the
// interpolation weights DXC(i)/(2DXF(i)), DXC(i+1)/(2DXF(i))
struct CellX {
  double dxc, rdxc, dxc_p, rdxc_p, dxf, rdxf, c_xm, c_xp;
};
// LANE: the lanes hold consecutive columns, so the gradient at the left face
// of column i is the one the lane of column i-1 forms at its right face,
// bit for bit (same operands, same operations); it is taken from there.
// The x-gradients at the faces of cell i, dtdx_xm = (T(i)-T(i-1))/DXC(i) and
// dtdx_xp = (T(i+1)-T(i))/DXC(i+1), are given (the t_xm, t_xp entries of v
// are not read): a kernel holding consecutive columns forms each face
// gradient once and hands it to both cells.
template <bool P2 = false>
__device__ __forceinline__ void temp_src_g(const Geo &g, const Tables &tb, const CellX &x,
                                           const TempPt &v, double dtdx_xm, double dtdx_xp,
                                           double wex, double wey, double wix, double wiy,
                                           double &srca, double &srcg) {
  const double dy = tb.dy, rdy = tb.rdy;
  const double dtdy_ym = dv<P2>(-v.t_ym + v.t_c, dy, rdy);
  const double dtdy_yp = dv<P2>(-v.t_c + v.t_yp, dy, rdy);
  const double adv1 = -x.c_xm * v.ux_c * dtdx_xm - x.c_xp * v.ux_p * dtdx_xp;
  const double adv2 = -0.5 * v.uy_c * dtdy_ym - 0.5 * v.uy_p * dtdy_yp;
  const double dif1 = dv<P2>(g.kappa_T * (dtdx_xp - dtdx_xm), x.dxf, x.rdxf);
  const double dif2 = dv<P2>(g.kappa_T * (dtdy_yp - dtdy_ym), dy, rdy);
  srca = +(adv1 + adv2) + wex * dif1 + wey * dif2;
  srcg = +wix * dif1 + wiy * dif2;
}

template <bool LANE = false, bool P2 = false>
__device__ __forceinline__ void temp_src_v(const Geo &g, const Tables &tb, const CellX &x,
                                           const TempPt &v, double wex, double wey, double wix,
                                           double wiy, double &srca, double &srcg) {
  const double dtdx_xp = dv<P2>(-v.t_c + v.t_xp, x.dxc_p, x.rdxc_p);
  const double dtdx_xm = LANE ? from_left(dtdx_xp) : dv<P2>(-v.t_xm + v.t_c, x.dxc, x.rdxc);
  temp_src_g<P2>(g, tb, x, v, dtdx_xm, dtdx_xp, wex, wey, wix, wiy, srca, srcg);
}

// x-momentum at face (i, j), i = 2..itot: src/fluid/update_velocity.c:34-101
struct UxPt {
  double u_c, u_xm, u_xp, u_ym, u_yp;  // UX(i,j), UX(i-1,j), UX(i+1,j), UX(i,j-1), UX(i,j+1)
  double uy_cm, uy_cc, uy_pm, uy_pc;   // UY(i-1,j), UY(i,j), UY(i-1,j+1), UY(i,j+1)
  double p_c, p_m;                     // P(i,j), P(i-1,j)
  double tf;                           // buoyancy at the face
};
// Face metrics of face i: DXF(i-1), DXF(i) with reciprocals, the weights
// DXF(i-1)/(2DXC(i)), DXF(i)/(2DXC(i)) and DXC(i) with its reciprocal
struct FaceX {
  double dxf_m, rdxf_m, dxf, rdxf, c_xm, c_xp, dxc, rdxc;
};
// The x-gradients (UX(i)-UX(i-1))/DXF(i-1) and (UX(i+1)-UX(i))/DXF(i) are
// given, as for temp_src_g (u_xm and u_xp still enter the interpolations).
template <bool P2 = false>
__device__ __forceinline__ void ux_src_g(const Geo &g, const Tables &tb, const FaceX &x,
                                         const UxPt &v, double duxdx_xm, double duxdx_xp,
                                         double wex, double wey, double wix, double wiy,
                                         double &srca, double &srcg) {
  const double dy = tb.dy, rdy = tb.rdy;
  const double duxdy_ym = dv<P2>(-v.u_ym + v.u_c, dy, rdy);
  const double duxdy_yp = dv<P2>(-v.u_c + v.u_yp, dy, rdy);
  const double ux_xm = 0.5 * v.u_xm + 0.5 * v.u_c;
  const double ux_xp = 0.5 * v.u_c + 0.5 * v.u_xp;
  const double adv1 = -x.c_xm * ux_xm * duxdx_xm - x.c_xp * ux_xp * duxdx_xp;
  const double uy_ym = x.c_xm * v.uy_cm + x.c_xp * v.uy_cc;
  const double uy_yp = x.c_xm * v.uy_pm + x.c_xp * v.uy_pc;
  const double adv2 = -0.5 * uy_ym * duxdy_ym - 0.5 * uy_yp * duxdy_yp;
  const double dif1 = dv<P2>(g.kappa_u * (duxdx_xp - duxdx_xm), x.dxc, x.rdxc);
  const double dif2 = dv<P2>(g.kappa_u * (duxdy_yp - duxdy_ym), dy, rdy);
  const double pre = dv<P2>(-(v.p_c - v.p_m), x.dxc, x.rdxc);
  srca = +(adv1 + adv2) + wex * dif1 + wey * dif2 + v.tf;
  srcg = +wix * dif1 + wiy * dif2 + pre;
}

template <bool LANE = false, bool P2 = false>
__device__ __forceinline__ void ux_src_v(const Geo &g, const Tables &tb, const FaceX &x,
                                         const UxPt &v, double wex, double wey, double wix,
                                         double wiy, double &srca, double &srcg) {
  const double duxdx_xp = dv<P2>(-v.u_c + v.u_xp, x.dxf, x.rdxf);
  const double duxdx_xm =
      LANE ? from_left(duxdx_xp) : dv<P2>(-v.u_xm + v.u_c, x.dxf_m, x.rdxf_m);
  ux_src_g<P2>(g, tb, x, v, duxdx_xm, duxdx_xp, wex, wey, wix, wiy, srca, srcg);
}

// y-momentum at (i, j), i = 1..itot: src/fluid/update_velocity.c:128-189
struct UyPt {
  double v_c, v_xm, v_xp, v_ym, v_yp;  // UY(i,j), UY(i-1,j), UY(i+1,j), UY(i,j-1), UY(i,j+1)
  double ux_mc, ux_cc, ux_mp, ux_cp;   // UX(i,j-1), UX(i,j), UX(i+1,j-1), UX(i+1,j)
  double p_c, p_ym;                    // P(i,j), P(i,j-1)
};
// The x-gradients (UY(i)-UY(i-1))/DXC(i) and (UY(i+1)-UY(i))/DXC(i+1) are
// given, as for temp_src_g.
template <bool P2 = false>
__device__ __forceinline__ void uy_src_g(const Geo &g, const Tables &tb, const CellX &x,
                                         const UyPt &v, double duydx_xm, double duydx_xp,
                                         double wex, double wey, double wix, double wiy,
                                         double &srca, double &srcg) {
  const double dy = tb.dy, rdy = tb.rdy;
  const double duydy_ym = dv<P2>(-v.v_ym + v.v_c, dy, rdy);
  const double duydy_yp = dv<P2>(-v.v_c + v.v_yp, dy, rdy);
  const double ux_xm = 0.5 * v.ux_mc + 0.5 * v.ux_cc;
  const double ux_xp = 0.5 * v.ux_mp + 0.5 * v.ux_cp;
  const double adv1 = -x.c_xm * ux_xm * duydx_xm - x.c_xp * ux_xp * duydx_xp;
  const double uy_ym = 0.5 * v.v_ym + 0.5 * v.v_c;
  const double uy_yp = 0.5 * v.v_c + 0.5 * v.v_yp;
  const double adv2 = -0.5 * uy_ym * duydy_ym - 0.5 * uy_yp * duydy_yp;
  const double dif1 = dv<P2>(g.kappa_u * (duydx_xp - duydx_xm), x.dxf, x.rdxf);
  const double dif2 = dv<P2>(g.kappa_u * (duydy_yp - duydy_ym), dy, rdy);
  const double pre = dv<P2>(-(v.p_c - v.p_ym), dy, rdy);
  srca = +(adv1 + adv2) + wex * dif1 + wey * dif2;
  srcg = +wix * dif1 + wiy * dif2 + pre;
}

template <bool LANE = false, bool P2 = false>
__device__ __forceinline__ void uy_src_v(const Geo &g, const Tables &tb, const CellX &x,
                                         const UyPt &v, double wex, double wey, double wix,
                                         double wiy, double &srca, double &srcg) {
  const double duydx_xp = dv<P2>(-v.v_c + v.v_xp, x.dxc_p, x.rdxc_p);
  const double duydx_xm = LANE ? from_left(duydx_xp) : dv<P2>(-v.v_xm + v.v_c, x.dxc, x.rdxc);
  uy_src_g<P2>(g, tb, x, v, duydx_xm, duydx_xp, wex, wey, wix, wiy, srca, srcg);
}

// ------------------------------------------------ loaders (pointer form)

__device__ __forceinline__ CellX cell_x(const Tables &tb, int i) {
  return {ldr(tb.dxc(), i - 1), ldr(tb.rdxc(), i - 1), ldr(tb.dxc(), i),   ldr(tb.rdxc(), i),
          ldr(tb.dxf(), i - 1), ldr(tb.rdxf(), i - 1), ldr(tb.cT_xm(), i - 1), ldr(tb.cT_xp(), i - 1)};
}
__device__ __forceinline__ FaceX face_x(const Tables &tb, int i) {
  return {ldr(tb.dxf(), i - 2),   ldr(tb.rdxf(), i - 2),  ldr(tb.dxf(), i - 1), ldr(tb.rdxf(), i - 1),
          ldr(tb.cU_xm(), i - 1), ldr(tb.cU_xp(), i - 1), ldr(tb.dxc(), i - 1), ldr(tb.rdxc(), i - 1)};
}

template <bool P2 = false>
__device__ __forceinline__ void temp_src_t(const Geo &g, const Tables &tb, const Rows3 &ux,
                                           const Rows3 &uy, const Rows3 &temp, int i, double wex,
                                           double wey, double wix, double wiy, double &srca,
                                           double &srcg) {
  const TempPt v = {ldr(temp.c, i), ldr(temp.c, i - 1), ldr(temp.c, i + 1), ldr(temp.m, i),
                    ldr(temp.p, i), ldr(ux.c, i - 1),   ldr(ux.c, i),       ldr(uy.c, i),
                    ldr(uy.p, i)};
  temp_src_v<false, P2>(g, tb, cell_x(tb, i), v, wex, wey, wix, wiy, srca, srcg);
}

template <bool P2 = false>
__device__ __forceinline__ void ux_src_t(const Geo &g, const Tables &tb, const Rows3 &ux,
                                         const Rows3 &uy, const Rows3 &p, double tf, int i,
                                         double wex, double wey, double wix, double wiy,
                                         double &srca, double &srcg) {
  const UxPt v = {ldr(ux.c, i - 1), ldr(ux.c, i - 2), ldr(ux.c, i), ldr(ux.m, i - 1),
                  ldr(ux.p, i - 1), ldr(uy.c, i - 1), ldr(uy.c, i), ldr(uy.p, i - 1),
                  ldr(uy.p, i),     ldr(p.c, i),      ldr(p.c, i - 1), tf};
  ux_src_v<false, P2>(g, tb, face_x(tb, i), v, wex, wey, wix, wiy, srca, srcg);
}

template <bool P2 = false>
__device__ __forceinline__ void uy_src_t(const Geo &g, const Tables &tb, const Rows3 &ux,
                                         const Rows3 &uy, const Rows3 &p, int i, double wex,
                                         double wey, double wix, double wiy, double &srca,
                                         double &srcg) {
  const UyPt v = {ldr(uy.c, i),     ldr(uy.c, i - 1), ldr(uy.c, i + 1), ldr(uy.m, i),
                  ldr(uy.p, i),     ldr(ux.m, i - 1), ldr(ux.c, i - 1), ldr(ux.m, i),
                  ldr(ux.c, i),     ldr(p.c, i),      ldr(p.m, i)};
  uy_src_v<false, P2>(g, tb, cell_x(tb, i), v, wex, wey, wix, wiy, srca, srcg);
}

// ---------------------------------------------------------------- stores

// Value of ux at face (i, j) (j uniform across the workgroup) with its
// periodic copy and the wall values of the row.
__device__ __forceinline__ void put_ux(const Geo &g, double *f, int i, int j, double v) {
  const size_t w = (size_t)g.pu;
  double *row = f + (size_t)j * w;
  double *halo = j == 1 ? f + (size_t)(g.jtot + 1) * w : f;  // used when j is 1 or jtot
  const bool copy = j == 1 || j == g.jtot;
  str(row, i - 1, v);
  if (copy) str(halo, i - 1, v);
  if (i == 2 || i == g.itot) {
    const int kw = (i == 2) ? 0 : g.itot;
    str(row, kw, 0.);
    if (copy) str(halo, kw, 0.);
  }
}

// Value at cell (i, j) of uy / p / T (j uniform across the workgroup), with
// its periodic copy and wall values.  WALL: 0 zero walls (uy), 1 Neumann
// copies (p), 2 temperature walls (rows 1..jtot only).
template <int WALL>
__device__ __forceinline__ void put_c(const Geo &g, double *f, int i, int j, double v) {
  const size_t w = (size_t)g.pc;
  double *row = f + (size_t)j * w;
  double *halo = j == 1 ? f + (size_t)(g.jtot + 1) * w : f;
  const bool copy = j == 1 || j == g.jtot;
  str(row, i, v);
  if (copy) str(halo, i, v);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const bool hit = side == 0 ? (i == 1) : (i == g.itot);
    if (!hit) continue;
    const int iw = side == 0 ? 0 : g.itot + 1;
    const double wv = WALL == 0 ? 0. : (WALL == 1 ? v : (side == 0 ? +0.5 : -0.5));
    str(row, iw, wv);
    if (WALL != 2 && copy) str(halo, iw, wv);
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------- projection of one point

// src/fluid/correct_velocity.c:6-42 and src/fluid/update_pressure.c:7-60
// at one point.  k_corr and the predictors that apply the previous
// substage's correction to the fields they load both use these, so both
// produce the same bits.  gdt is gamma*dt of the substage that solved for psi.

// ux(i,j) = ux*(i,j) - gdt (psi(i,j) - psi(i-1,j)) / DXC(i), the quotient
// as a product with the tabulated 1./DXC(i)
__device__ __forceinline__ double proj_ux(double us, double psi_c, double psi_w, double gdt,
                                          double rdxc) {
  return us - gdt * (psi_c - psi_w) * rdxc;
}

// uy(i,j) = uy*(i,j) - gdt (psi(i,j) - psi(i,j-1)) / dy.  P2 (dy a power
// of two): the product with 1/dy, the same bits as div_cr (see dv).
template <bool P2 = false>
__device__ __forceinline__ double proj_uy(double vs, double psi_c, double psi_s, double gdt,
                                          double dy, double rdy) {
  return vs - dv<P2>(gdt * (psi_c - psi_s), dy, rdy);
}

// Metrics of cell i: DXC(i), DXC(i+1), DXF(i) and their reciprocals
struct ProjX {
  double dxc_i, rdxc_i, dxc_p, rdxc_p, dxf_i, rdxf_i;
};

// p(i,j) + psi(i,j) + fac * (the implicitly treated parts of the Laplacian
// of psi), fac = -gdt / (2 Re); psi_w, psi_e, psi_s, psi_n are the values at
// (i-1,j), (i+1,j), (i,j-1), (i,j+1).  P2 as for proj_uy.
template <bool P2 = false>
__device__ __forceinline__ double proj_p(double pj, double psi_c, double psi_w, double psi_e,
                                         double psi_s, double psi_n, const ProjX &x, double dy,
                                         double rdy, double fac, int implicitx, int implicity) {
  double pv = pj + psi_c;
  if (implicitx) {
    const double dm = dv<P2>(-psi_w + psi_c, x.dxc_i, x.rdxc_i);
    const double dp = dv<P2>(-psi_c + psi_e, x.dxc_p, x.rdxc_p);
    pv += fac * (+dv<P2>(dp - dm, x.dxf_i, x.rdxf_i));
  }
  if (implicity) {
    const double dm = dv<P2>(-psi_s + psi_c, dy, rdy);
    const double dp = dv<P2>(-psi_c + psi_n, dy, rdy);
    pv += fac * (+dv<P2>(dp - dm, dy, rdy));
  }
  return pv;
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
