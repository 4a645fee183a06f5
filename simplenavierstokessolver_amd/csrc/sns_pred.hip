// sns_pred.hip -- implicit-x predictor: shared LU factorisation of the
// Crank-Nicolson x operators and the row kernel (gfx950).
//
// k_pred_rows evaluates, for one field and two adjacent rows per workgroup,
// the source terms and RK increments of src/temperature/update.c:11-122 and
// src/fluid/update_velocity.c:11-229, then solves the Crank-Nicolson system
// in x of both rows (src/temperature/update.c:124-156,
// src/fluid/update_velocity.c:231-270, 381-414) with the shared LU tables.
//
// Memory traffic is set by the number of load instructions, so neighbours
// are not loaded twice:
//  * x neighbours come from the adjacent lane (DPP wave shifts); each wave
//    covers 62 columns and its first and last lanes only hold the halo
//    values of the columns on either side;
//  * the y neighbours of one row are the centre values of the other, and the
//    per-column metrics and the LU tables serve both rows.
#include "sns_launch.h"
#include "sns_fast.h"
#include "sns_fused.h"
#include "sns_lu.h"

namespace sns {

// ------------------------------------------ shared LU factorisation of CN-x

// Coefficients of row r (0-based unknown index) of the CN-x operator of
// field f (0: T, 1: ux, 2: uy): src/temperature/update.c:141-149,
// src/fluid/update_velocity.c:253-261,398-406.
__device__ __forceinline__ void cnx_row(const Geo &g, const Tables &tb, int f, double pf, int r,
                                        double &l, double &c, double &u) {
  const int i = r + (f == 1 ? 2 : 1);
  const double coef_xm = f == 1 ? tb.nU_xm()[i - 1] : tb.nT_xm()[i - 1];
  const double coef_xp = f == 1 ? tb.nU_xp()[i - 1] : tb.nT_xp()[i - 1];
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
template <int NT, int M, class S>
__global__ void __launch_bounds__(NT) k_lu_factor(S src) {
  const LuFactorArgs a = src.get();
  constexpr int NW = NT / 64;
  __shared__ double wtot[NW][4];
  const Geo &g = a.g;
  const int f = blockIdx.x, s = blockIdx.y;
  const int n = f == 1 ? g.itot - 1 : g.itot;
  // substage gamma selected without indexing (keeps the argument block in registers)
  const double gam = s == 0 ? a.gamma[0] : s == 1 ? a.gamma[1] : a.gamma[2];
  const double pf = cn_pf(g, f, gam * (*a.dtp));
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
  double *base = a.out + (size_t)((s * 3 + f) * 3) * a.ntm;
#pragma unroll
  for (int k = 0; k < M; k++) {
    base[0 * a.ntm + k * NT + t] = rr[k];
    base[1 * a.ntm + k * NT + t] = al[k];
    base[2 * a.ntm + k * NT + t] = be[k];
  }
  // Header: P = the last row in 1..n-2 whose (r, al, be) differs bit for bit
  // from the row before (0 if none), so that rows P..n-2 all hold the
  // triple of row n-2; on a grid whose interior rows are alike the pivot
  // recurrence reaches that fixed point after a few rows.  Small systems
  // (n < 4) keep every row in the table (P = n).
  __shared__ double wlast[NW][3];
  __shared__ int wmax[NW];
  auto bits = [](double v) { return __double_as_longlong(v); };
  if (lane == 63) {
    wlast[wave][0] = rr[M - 1];
    wlast[wave][1] = al[M - 1];
    wlast[wave][2] = be[M - 1];
  }
  __syncthreads();
  double pr = __shfl_up(rr[M - 1], 1), pa = __shfl_up(al[M - 1], 1), pb = __shfl_up(be[M - 1], 1);
  if (lane == 0 && wave > 0) {
    pr = wlast[wave - 1][0];
    pa = wlast[wave - 1][1];
    pb = wlast[wave - 1][2];
  }
  int pt = 0;
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = t * M + k;
    const double qr = k ? rr[k - 1] : pr, qa = k ? al[k - 1] : pa, qb = k ? be[k - 1] : pb;
    if (r >= 1 && r <= n - 2 &&
        (bits(rr[k]) != bits(qr) || bits(al[k]) != bits(qa) || bits(be[k]) != bits(qb)))
      pt = r;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) pt = max(pt, __shfl_xor(pt, off));
  if (lane == 0) wmax[wave] = pt;
  double *hdr = a.out + 27 * (size_t)a.ntm + (s * 3 + f) * kLuHdr;
#pragma unroll
  for (int k = 0; k < M; k++) {
    if (t * M + k == n - 2) {
      hdr[1] = rr[k];
      hdr[2] = al[k];
      hdr[3] = be[k];
    }
  }
  __syncthreads();
  if (t == 0) {
    int P = 0;
    for (int w = 0; w < NW; w++) P = max(P, wmax[w]);
    hdr[0] = (double)(n < 4 ? n : P);
    if (n < 2) hdr[1] = hdr[2] = hdr[3] = 0.;
  }
}

// -------------------------------------------------------- row predictor

template <int NT, int M>
struct PredRowsLds {
  double d[2][NT * (M + 1)];  // padded: chunk t starts at t*(M+1)
  double scr[4 * 2 * (NT / 64)];
};

template <int M>
__device__ __forceinline__ int pad_idx(int r) {
  return r + r / M;
}

constexpr int kWaveCols = 62;  // output columns per wave (lanes 1..62)

// Lane position of stencil block k: unknown index ii (lanes 1..62 of each
// wave own a column, lanes 0 and 63 hold the neighbours), the column i and
// the clamped element indices of the cell arrays (ci), ux (ui), the column
// tables (ti) and the RK history (hi).
struct LanePos {
  int ii, i, ci, ui, ti, hi;
  bool out;
  bool interior;  // all 64 lanes of the wave on columns 2..itot-1 (wave-uniform)
};

__device__ __forceinline__ LanePos lane_pos(int k, int nw, int i0, int n, int itot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  LanePos p;
  p.ii = (k * nw + wave) * kWaveCols + lane - 1;
  p.i = p.ii + i0;
  p.out = lane >= 1 && lane <= kWaveCols && p.ii < n;
  p.ci = clampi(p.i, 0, itot + 1);
  p.ui = clampi(p.i - 1, 0, itot);
  p.ti = clampi(p.i - 1, 0, itot + 1);
  p.hi = clampi(p.ii, 0, n - 1);
  const int c0 = (k * nw + wave) * kWaveCols - 1 + i0;  // column of lane 0
  p.interior = c0 >= 2 && c0 + 63 <= itot - 1;
  return p;
}

// The values one lane loads for the two rows of field F (0: T, 1: ux,
// 2: uy): the substage-start fields around rows ja, jb, the column tables
// and the RK history.
template <int F>
struct Raw;
template <>
struct Raw<0> {
  double t_am, t_a, t_b, t_bp, u_a, u_b, v_a, v_b, v_bp, ba, bb;
  double2 c2, f2, w2;
};
template <>
struct Raw<1> {
  double u_am, u_a, u_b, u_bp, v_a, v_b, v_bp, p_a, p_b, t_a, t_b, ba, bb;
  double2 c2, f2, w2;
};
template <>
struct Raw<2> {
  double v_am, v_a, v_b, v_bp, u_am, u_a, u_b, p_am, p_a, p_b, ba, bb;
  double2 c2, f2, w2;
};

template <int F>
__device__ __forceinline__ Raw<F> raw_load(const PredArgs &a, const LanePos &p, int ja, int jb,
                                           int n) {
  const Geo &g = a.g;
  const Tables &tb = a.tb;
  const size_t wc = (size_t)g.pc, wu = (size_t)g.pu;
  const double *T = a.temp, *U = a.ux, *V = a.uy, *P = a.p;
  Raw<F> r;
  const double *H = F == 0 ? a.hist_t : (F == 1 ? a.hist_ux : a.hist_uy);
  r.ba = a.read_hist ? ldr(H + (size_t)(ja - 1) * n, p.hi) : 0.;
  r.bb = a.read_hist ? ldr(H + (size_t)(jb - 1) * n, p.hi) : 0.;
  if (p.interior && tb.uniform) {
    r.c2 = make_double2(tb.u_dxc, tb.u_rdxc);
    r.f2 = make_double2(tb.u_dxf, tb.u_rdxf);
    const double w = F == 1 ? tb.u_cU : tb.u_cT;
    r.w2 = make_double2(w, w);
  } else {
    r.c2 = ldr2(tb.dxc2(), p.ti);
    r.f2 = ldr2(tb.dxf2(), p.ti);
    r.w2 = ldr2(F == 1 ? tb.cU2() : tb.cT2(), p.ti);
  }
  if constexpr (F == 0) {
    r.t_am = ldr(T + (ja - 1) * wc, p.ci);
    r.t_a = ldr(T + ja * wc, p.ci);
    r.t_b = ldr(T + jb * wc, p.ci);
    r.t_bp = ldr(T + (jb + 1) * wc, p.ci);
    r.u_a = ldr(U + ja * wu, p.ui);
    r.u_b = ldr(U + jb * wu, p.ui);
    r.v_a = ldr(V + ja * wc, p.ci);
    r.v_b = ldr(V + jb * wc, p.ci);
    r.v_bp = ldr(V + (jb + 1) * wc, p.ci);
  } else if constexpr (F == 1) {
    r.u_am = ldr(U + (ja - 1) * wu, p.ui);
    r.u_a = ldr(U + ja * wu, p.ui);
    r.u_b = ldr(U + jb * wu, p.ui);
    r.u_bp = ldr(U + (jb + 1) * wu, p.ui);
    r.v_a = ldr(V + ja * wc, p.ci);
    r.v_b = ldr(V + jb * wc, p.ci);
    r.v_bp = ldr(V + (jb + 1) * wc, p.ci);
    r.p_a = ldr(P + ja * wc, p.ci);
    r.p_b = ldr(P + jb * wc, p.ci);
    r.t_a = a.forcing ? ldr(T + ja * wc, p.ci) : 0.;
    r.t_b = a.forcing ? ldr(T + jb * wc, p.ci) : 0.;
  } else {
    r.v_am = ldr(V + (ja - 1) * wc, p.ci);
    r.v_a = ldr(V + ja * wc, p.ci);
    r.v_b = ldr(V + jb * wc, p.ci);
    r.v_bp = ldr(V + (jb + 1) * wc, p.ci);
    r.u_am = ldr(U + (ja - 1) * wu, p.ui);
    r.u_a = ldr(U + ja * wu, p.ui);
    r.u_b = ldr(U + jb * wu, p.ui);
    r.p_am = ldr(P + (ja - 1) * wc, p.ci);
    r.p_a = ldr(P + ja * wc, p.ci);
    r.p_b = ldr(P + jb * wc, p.ci);
  }
  return r;
}

// The previous substage's correction applied to the values raw_load took
// from its ux*, uy* and pressure, with the operations of k_corr (proj_* of
// sns_fast.h; P2 gives the same bits as its divisions): psi of the rows the
// corrected values need is loaded here, its x neighbours come from the
// adjacent lanes.  Lane 0 of a ux workgroup loads psi(i-1) itself, since
// the face and the pressure it corrects feed lane 1; in the other
// workgroups lane 0 only feeds uy and T, which need no x neighbour of psi.
//
// Faces 2..itot and cells 1..itot are corrected, the rest keep what the
// predictor stored, which is what k_corr leaves there:
//  * faces: psi's wall columns are exact copies of its columns 1 and itot,
//    so psi(i) - psi(i-1) is +0 at the wall faces 1 and itot+1 (and at the
//    clamped lanes past them) and the correction leaves those faces as they
//    are, with no select;
//  * uy: the wall columns 0 and itot+1 are kept by a select in the uy
//    workgroup, the only one whose outputs read them;
//  * p: values outside cells 1..itot feed no output.
// Row ja-2 of a uy workgroup wraps to jtot-1 for ja = 1 (periodic in y), and
// its pressure of row 0 is taken from row jtot: ux*, uy* and psi carry
// periodic copies in their halo rows, the pressure only once k_corr has
// written it.  The uy workgroup stores the corrected pressure of its rows in
// p_out, the stores of k_corr for those cells.
template <int F, bool P2>
__device__ __forceinline__ void raw_fold(const PredArgs &a, const LanePos &p, int ja, int jb,
                                         bool write_a, double dt, Raw<F> &r) {
  const Geo &g = a.g;
  const Tables &tb = a.tb;
  const int itot = g.itot;
  const size_t wc = (size_t)g.pc;
  const double *S = a.psi;
  const double gdt = a.corr_gamma * dt;
  const double fac = -gdt / g.two_Re;
  const double rdxc = r.c2.y;  // 1./DXC(i)
  const double dy = tb.dy, rdy = tb.rdy;
  const int imx = a.tr.implicitx, imy = a.tr.implicity;
  const double s_am = ldr(S + (size_t)(ja - 1) * wc, p.ci);
  const double s_a = ldr(S + (size_t)ja * wc, p.ci);
  const double s_b = ldr(S + (size_t)jb * wc, p.ci);
  const double s_bp = ldr(S + (size_t)(jb + 1) * wc, p.ci);
  // Lane shifts run on every lane, never under a divergent condition (a
  // shift there would read the inactive neighbour lanes as 0).
  if constexpr (F == 0) {
    r.u_a = proj_ux(r.u_a, s_a, from_left(s_a), gdt, rdxc);
    r.u_b = proj_ux(r.u_b, s_b, from_left(s_b), gdt, rdxc);
    r.v_a = proj_uy<P2>(r.v_a, s_a, s_am, gdt, dy, rdy);
    r.v_b = proj_uy<P2>(r.v_b, s_b, s_a, gdt, dy, rdy);
    r.v_bp = proj_uy<P2>(r.v_bp, s_bp, s_b, gdt, dy, rdy);
  } else {
    const ProjX x = {r.c2.x, r.c2.y, from_right(r.c2.x), from_right(r.c2.y), r.f2.x, r.f2.y};
    if constexpr (F == 1) {
      const int cw = clampi(p.i - 1, 0, itot + 1);
      double w_am = from_left(s_am), w_a = from_left(s_a), w_b = from_left(s_b),
             w_bp = from_left(s_bp);
      if ((threadIdx.x & 63) == 0) {
        w_am = ldr(S + (size_t)(ja - 1) * wc, cw);
        w_a = ldr(S + (size_t)ja * wc, cw);
        w_b = ldr(S + (size_t)jb * wc, cw);
        w_bp = ldr(S + (size_t)(jb + 1) * wc, cw);
      }
      const double e_a = from_right(s_a), e_b = from_right(s_b);
      r.u_am = proj_ux(r.u_am, s_am, w_am, gdt, rdxc);
      r.u_a = proj_ux(r.u_a, s_a, w_a, gdt, rdxc);
      r.u_b = proj_ux(r.u_b, s_b, w_b, gdt, rdxc);
      r.u_bp = proj_ux(r.u_bp, s_bp, w_bp, gdt, rdxc);
      r.v_a = proj_uy<P2>(r.v_a, s_a, s_am, gdt, dy, rdy);
      r.v_b = proj_uy<P2>(r.v_b, s_b, s_a, gdt, dy, rdy);
      r.v_bp = proj_uy<P2>(r.v_bp, s_bp, s_b, gdt, dy, rdy);
      r.p_a = proj_p<P2>(r.p_a, s_a, w_a, e_a, s_am, s_b, x, dy, rdy, fac, imx, imy);
      r.p_b = proj_p<P2>(r.p_b, s_b, w_b, e_b, s_a, s_bp, x, dy, rdy, fac, imx, imy);
    } else {
      const bool cc = p.i >= 1 && p.i <= itot;  // uy cells the correction moves
      const int jmm = ja >= 2 ? ja - 2 : g.jtot - 1;
      const double s_amm = ldr(S + (size_t)jmm * wc, p.ci);
      if (ja == 1) r.p_am = ldr(a.p + (size_t)g.jtot * wc, p.ci);
      const double w_am = from_left(s_am), w_a = from_left(s_a), w_b = from_left(s_b);
      const double e_am = from_right(s_am), e_a = from_right(s_a), e_b = from_right(s_b);
      const double v_am = proj_uy<P2>(r.v_am, s_am, s_amm, gdt, dy, rdy);
      const double v_a = proj_uy<P2>(r.v_a, s_a, s_am, gdt, dy, rdy);
      const double v_b = proj_uy<P2>(r.v_b, s_b, s_a, gdt, dy, rdy);
      const double v_bp = proj_uy<P2>(r.v_bp, s_bp, s_b, gdt, dy, rdy);
      r.v_am = cc ? v_am : r.v_am;
      r.v_a = cc ? v_a : r.v_a;
      r.v_b = cc ? v_b : r.v_b;
      r.v_bp = cc ? v_bp : r.v_bp;
      r.u_am = proj_ux(r.u_am, s_am, w_am, gdt, rdxc);
      r.u_a = proj_ux(r.u_a, s_a, w_a, gdt, rdxc);
      r.u_b = proj_ux(r.u_b, s_b, w_b, gdt, rdxc);
      r.p_am = proj_p<P2>(r.p_am, s_am, w_am, e_am, s_amm, s_a, x, dy, rdy, fac, imx, imy);
      r.p_a = proj_p<P2>(r.p_a, s_a, w_a, e_a, s_am, s_b, x, dy, rdy, fac, imx, imy);
      r.p_b = proj_p<P2>(r.p_b, s_b, w_b, e_b, s_a, s_bp, x, dy, rdy, fac, imx, imy);
      if (p.out) {
        if (write_a) put_c<1>(g, a.p_out, p.i, ja, r.p_a);
        put_c<1>(g, a.p_out, p.i, jb, r.p_b);
      }
    }
  }
}

// Source terms and RK increments of both rows from the loaded values (x
// neighbours from the adjacent lanes); the new RK history is stored.
template <int F, bool P2>
__device__ __forceinline__ void raw_eval(const PredArgs &a, const Raw<F> &r, const LanePos &p,
                                         int ja, int jb, int n, bool write_a, double dt,
                                         double &da, double &db) {
  const Geo &g = a.g;
  const Tables &tb = a.tb;
  const double wex = a.tr.w_ex_x, wey = a.tr.w_ex_y, wix = a.tr.w_im_x, wiy = a.tr.w_im_y;
  double *H = F == 0 ? a.hist_t : (F == 1 ? a.hist_ux : a.hist_uy);
  double *HA = H + (size_t)(ja - 1) * n, *HB = H + (size_t)(jb - 1) * n;
  double A, G;
  if constexpr (F == 0) {
    CellX x;
    x.dxc = r.c2.x;
    x.rdxc = r.c2.y;
    x.dxf = r.f2.x;
    x.rdxf = r.f2.y;
    x.c_xm = r.w2.x;
    x.c_xp = r.w2.y;
    x.dxc_p = from_right(x.dxc);
    x.rdxc_p = from_right(x.rdxc);
    const TempPt pa = {r.t_a, from_left(r.t_a), from_right(r.t_a), r.t_am, r.t_b,
                       r.u_a, from_right(r.u_a), r.v_a, r.v_b};
    const TempPt pb = {r.t_b, from_left(r.t_b), from_right(r.t_b), r.t_a, r.t_bp,
                       r.u_b, from_right(r.u_b), r.v_b, r.v_bp};
    temp_src_v<true, P2>(g, tb, x, pa, wex, wey, wix, wiy, A, G);
    if (a.write_hist && p.out && write_a) str(HA, p.ii, A);
    da = rk_incr(a.alpha, a.beta, a.gamma, dt, A, r.ba, G);
    temp_src_v<true, P2>(g, tb, x, pb, wex, wey, wix, wiy, A, G);
    if (a.write_hist && p.out) str(HB, p.ii, A);
    db = rk_incr(a.alpha, a.beta, a.gamma, dt, A, r.bb, G);
  } else if constexpr (F == 1) {
    FaceX x;
    x.dxf = r.f2.x;
    x.rdxf = r.f2.y;
    x.c_xm = r.w2.x;
    x.c_xp = r.w2.y;
    x.dxc = r.c2.x;
    x.rdxc = r.c2.y;
    x.dxf_m = from_left(x.dxf);
    x.rdxf_m = from_left(x.rdxf);
    double tf_a = 0., tf_b = 0.;
    if (a.forcing) {
      tf_a = 0.5 * from_left(r.t_a) + 0.5 * r.t_a;
      tf_b = 0.5 * from_left(r.t_b) + 0.5 * r.t_b;
    }
    const double v_bm = from_left(r.v_b);
    const UxPt pa = {r.u_a, from_left(r.u_a), from_right(r.u_a), r.u_am, r.u_b,
                     from_left(r.v_a), r.v_a, v_bm, r.v_b, r.p_a, from_left(r.p_a), tf_a};
    const UxPt pb = {r.u_b, from_left(r.u_b), from_right(r.u_b), r.u_a, r.u_bp,
                     v_bm, r.v_b, from_left(r.v_bp), r.v_bp, r.p_b, from_left(r.p_b), tf_b};
    ux_src_v<true, P2>(g, tb, x, pa, wex, wey, wix, wiy, A, G);
    if (a.write_hist && p.out && write_a) str(HA, p.ii, A);
    da = rk_incr(a.alpha, a.beta, a.gamma, dt, A, r.ba, G);
    ux_src_v<true, P2>(g, tb, x, pb, wex, wey, wix, wiy, A, G);
    if (a.write_hist && p.out) str(HB, p.ii, A);
    db = rk_incr(a.alpha, a.beta, a.gamma, dt, A, r.bb, G);
  } else {
    CellX x;
    x.dxc = r.c2.x;
    x.rdxc = r.c2.y;
    x.dxf = r.f2.x;
    x.rdxf = r.f2.y;
    x.c_xm = r.w2.x;
    x.c_xp = r.w2.y;
    x.dxc_p = from_right(x.dxc);
    x.rdxc_p = from_right(x.rdxc);
    const double u_ap = from_right(r.u_a);
    const UyPt pa = {r.v_a, from_left(r.v_a), from_right(r.v_a), r.v_am, r.v_b,
                     r.u_am, r.u_a, from_right(r.u_am), u_ap, r.p_a, r.p_am};
    const UyPt pb = {r.v_b, from_left(r.v_b), from_right(r.v_b), r.v_a, r.v_bp,
                     r.u_a, r.u_b, u_ap, from_right(r.u_b), r.p_b, r.p_a};
    uy_src_v<true, P2>(g, tb, x, pa, wex, wey, wix, wiy, A, G);
    if (a.write_hist && p.out && write_a) str(HA, p.ii, A);
    da = rk_incr(a.alpha, a.beta, a.gamma, dt, A, r.ba, G);
    uy_src_v<true, P2>(g, tb, x, pb, wex, wey, wix, wiy, A, G);
    if (a.write_hist && p.out) str(HB, p.ii, A);
    db = rk_incr(a.alpha, a.beta, a.gamma, dt, A, r.bb, G);
  }
}

// Blocks of kWaveCols columns per wave of the stencil phase: at most
// ceil(ceil(NT*M / kWaveCols) / (NT/64)), 9 for M = 8.
template <int NT, int M>
constexpr int stencil_blocks() {
  return ((NT * M + kWaveCols - 1) / kWaveCols + NT / 64 - 1) / (NT / 64);
}

// The workgroup's own field at rows ja, jb in Raw (the values the increments
// are added to).
template <int F>
__device__ __forceinline__ double2 raw_own(const Raw<F> &r) {
  if constexpr (F == 0) return make_double2(r.t_a, r.t_b);
  else if constexpr (F == 1) return make_double2(r.u_a, r.u_b);
  else return make_double2(r.v_a, r.v_b);
}

// Stencil phase of one workgroup: blocks K.. of the nk blocks per wave
// (compile-time block index).  The workgroup's own field at rows ja, jb is
// kept per block in own[] for the final values, so those rows are not read
// a second time after the line solve.
template <int K, int F, int NT, int M, bool P2, bool FOLD>
__device__ __forceinline__ void pred_stencils(const PredArgs &a, PredRowsLds<NT, M> &L, int ja,
                                              int jb, bool write_a, int n, int i0, double dt,
                                              int nk, double2 (&own)[stencil_blocks<NT, M>()]) {
  if constexpr (K < stencil_blocks<NT, M>()) {
    if (K >= nk) return;
    const LanePos p = lane_pos(K, NT / 64, i0, n, a.g.itot);
    Raw<F> r = raw_load<F>(a, p, ja, jb, n);
    if constexpr (FOLD) raw_fold<F, P2>(a, p, ja, jb, write_a, dt, r);
    double da, db;
    raw_eval<F, P2>(a, r, p, ja, jb, n, write_a, dt, da, db);
    own[K] = raw_own<F>(r);
    if (p.out) {
      L.d[0][pad_idx<M>(p.ii)] = da;
      L.d[1][pad_idx<M>(p.ii)] = db;
    }
    pred_stencils<K + 1, F, NT, M, P2, FOLD>(a, L, ja, jb, write_a, n, i0, dt, nk, own);
  }
}

// Final values of blocks K.. of the stencil phase (compile-time block
// index, so own[] stays in registers): own + increment with boundary values,
// or the increment itself when a y solve follows.
template <int K, int F, int NT, int M>
__device__ __forceinline__ void pred_emit(const PredArgs &a, const PredRowsLds<NT, M> &L,
                                          const double2 (&own)[stencil_blocks<NT, M>()], int ja,
                                          int jb, bool write_a, int n, int i0, int nk) {
  if constexpr (K < stencil_blocks<NT, M>()) {
    if (K >= nk) return;
    const LanePos p = lane_pos(K, NT / 64, i0, n, a.g.itot);
    double *q = F == 0 ? a.q_t : (F == 1 ? a.q_u : a.q_v);
    if (p.out) {
#pragma unroll
      for (int s = 0; s < 2; s++) {
        if (s == 0 && !write_a) continue;
        const int j = s == 0 ? ja : jb;
        const double d = L.d[s][pad_idx<M>(p.ii)];
        const double o = s == 0 ? own[K].x : own[K].y;
        if (q) str(q + (size_t)(j - 1) * n, p.ii, d);
        else if (F == 0) put_c<2>(a.g, a.out_t, p.i, j, o + d);
        else if (F == 1) put_ux(a.g, a.out_u, p.i, j, o + d);
        else put_c<0>(a.g, a.out_v, p.i, j, o + d);
      }
    }
    pred_emit<K + 1, F, NT, M>(a, L, own, ja, jb, write_a, n, i0, nk);
  }
}

// One workgroup's field F for rows ja and jb (write_a: row ja is written):
// stencils into LDS, the CN-x solve of both rows, final values.
template <int F, int NT, int M, bool P2, bool FOLD>
__device__ __forceinline__ void pred_rows_body(const PredArgs &a, PredRowsLds<NT, M> &L, int ja,
                                               int jb, bool write_a) {
  constexpr int NW = NT / 64;
  constexpr int KB = stencil_blocks<NT, M>();
  const Geo &g = a.g;
  const int itot = g.itot;
  const int n = F == 1 ? itot - 1 : itot;
  const int i0 = F == 1 ? 2 : 1;
  const double dt = *a.dtp;
  const int nk = ((n + kWaveCols - 1) / kWaveCols + NW - 1) / NW;

  // ---- stencils and RK increments of both rows into LDS (lane order)
  double2 own[KB];
  pred_stencils<0, F, NT, M, P2, FOLD>(a, L, ja, jb, write_a, n, i0, dt, nk, own);
  __syncthreads();

  // ---- Crank-Nicolson in x of both rows (chunk order)
  double dd[2][M];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int k = 0; k < M; k++) {
      const int r = threadIdx.x * M + k;
      dd[s][k] = r < n ? L.d[s][pad_idx<M>(r)] : 0.;
    }
  const double *b = a.lu + (size_t)(F * 3) * a.ntm;
  const double *h = a.lu_hdr + F * kLuHdr;
  const LuTabC<M> tab = {b, b + a.ntm, b + 2 * a.ntm, (int)h[0], n, h[1], h[2], h[3]};
  lu_solve_shared<NT, M, 2>(dd, tab, L.scr);
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int k = 0; k < M; k++) {
      const int r = threadIdx.x * M + k;
      if (r < n) L.d[s][pad_idx<M>(r)] = dd[s][k];
    }
  __syncthreads();

  // ---- final values with boundary values, or the increments for a y
  // solve, in the lane order of the stencil phase
  pred_emit<0, F, NT, M>(a, L, own, ja, jb, write_a, n, i0, nk);
}

// Workgroup w: field w % nfld (+1 without the temperature equation) of rows
// a = 2q+1 and b = a+1, q = w / nfld.  For odd jtot the last pair is shifted
// down to (jtot-1, jtot) and writes row b only.  Consecutive w share an XCD.
// FOLD: the fields are corrected as they are loaded (raw_fold).
template <int NT, int M, bool P2, bool FOLD, class S>
__global__ void __launch_bounds__(NT) k_pred_rows(S src, int nfld) {
  const PredArgs a = src.get();
  extern __shared__ double lds_raw[];
  PredRowsLds<NT, M> &L = *reinterpret_cast<PredRowsLds<NT, M> *>(lds_raw);
  const int w = xcd_remap(blockIdx.x, gridDim.x);
  const int q = w / nfld;
  const int f = w % nfld + (nfld == 3 ? 0 : 1);
  const int ja = min(2 * q + 1, a.g.jtot - 1), jb = ja + 1;
  const bool write_a = ja == 2 * q + 1;
  if (f == 0) pred_rows_body<0, NT, M, P2, FOLD>(a, L, ja, jb, write_a);
  else if (f == 1) pred_rows_body<1, NT, M, P2, FOLD>(a, L, ja, jb, write_a);
  else pred_rows_body<2, NT, M, P2, FOLD>(a, L, ja, jb, write_a);
}

// ---------------------------------------------------------------- launch

template <int NT, int M, bool P2, bool FOLD, class S>
static void pred_rows_launch(hipStream_t s, const PredArgs &shape, S src, int nz) {
  const size_t lds = sizeof(PredRowsLds<NT, M>);
  static std::atomic<unsigned> attr{0};
  lds_limit_once((const void *)k_pred_rows<NT, M, P2, FOLD, S>, lds, attr);
  const int nfld = shape.with_temp ? 3 : 2;
  const int npair = (shape.g.jtot + 1) / 2;
  hipLaunchKernelGGL((k_pred_rows<NT, M, P2, FOLD, S>), dim3(npair * nfld, 1, nz), dim3(NT), lds,
                     s, src, nfld);
}

template <int NT, int M, class S>
static void lu_factor_launch(hipStream_t s, S src, int nz) {
  hipLaunchKernelGGL((k_lu_factor<NT, M, S>), dim3(3, 3, nz), dim3(NT), 0, s, src);
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

int lu_row_threads(int itot) {
  int NT, M;
  return row_config(itot, NT, M) ? NT : 0;
}

template <class S>
static void lu_factor_dispatch(hipStream_t s, int itot, S src, int nz) {
  int NT, M;
  row_config(itot, NT, M);
  switch (NT) {
    case 64: lu_factor_launch<64, kRowM>(s, src, nz); break;
    case 128: lu_factor_launch<128, kRowM>(s, src, nz); break;
    case 256: lu_factor_launch<256, kRowM>(s, src, nz); break;
    case 512: lu_factor_launch<512, kRowM>(s, src, nz); break;
    default: lu_factor_launch<1024, kRowM>(s, src, nz); break;
  }
}

void launch_lu_factor(hipStream_t s, const LuFactorArgs &a) {
  lu_factor_dispatch(s, a.g.itot, ArgOne<LuFactorArgs>{a}, 1);
}

void launch_lu_factor_batch(hipStream_t s, const LuFactorArgs &shape, const LuFactorArgs *dev,
                            int nz) {
  lu_factor_dispatch(s, shape.g.itot, arg_batch<LuFactorArgs>(dev), nz);
}

template <bool P2, bool FOLD, class S>
static void pred_rows_select(hipStream_t s, const PredArgs &shape, S src, int nz, int NT) {
  switch (NT) {
    case 64: pred_rows_launch<64, kRowM, P2, FOLD>(s, shape, src, nz); break;
    case 128: pred_rows_launch<128, kRowM, P2, FOLD>(s, shape, src, nz); break;
    case 256: pred_rows_launch<256, kRowM, P2, FOLD>(s, shape, src, nz); break;
    case 512: pred_rows_launch<512, kRowM, P2, FOLD>(s, shape, src, nz); break;
    default: pred_rows_launch<1024, kRowM, P2, FOLD>(s, shape, src, nz); break;
  }
}

// Power-of-two divisors select the kernel that divides by multiplying; a
// psi argument the one that applies the previous substage's correction.
template <class S>
static void pred_rows_dispatch(hipStream_t s, const PredArgs &shape, S src, int nz) {
  int NT, M;
  row_config(shape.g.itot, NT, M);
  const bool p2 = pow2_division(shape.tb), fold = shape.psi != nullptr;
  if (p2 && fold) pred_rows_select<true, true>(s, shape, src, nz, NT);
  else if (p2) pred_rows_select<true, false>(s, shape, src, nz, NT);
  else if (fold) pred_rows_select<false, true>(s, shape, src, nz, NT);
  else pred_rows_select<false, false>(s, shape, src, nz, NT);
}

void launch_pred_rows(hipStream_t s, const PredArgs &a) {
  pred_rows_dispatch(s, a, ArgOne<PredArgs>{a}, 1);
}

void launch_pred_rows_batch(hipStream_t s, const PredArgs &shape, const PredArgs *dev, int nz) {
  pred_rows_dispatch(s, shape, arg_batch<PredArgs>(dev), nz);
}

}  // namespace sns
