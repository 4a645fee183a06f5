// sns_ysolve.hip -- Crank-Nicolson in y on the row-major increments (gfx950).
//
// The implicit-y operator of one field (src/temperature/update.c:184-203,
// src/fluid/update_velocity.c:302-316, 428-468) is, for every column, the
// same periodic tridiagonal matrix with constant coefficients: l = u =
// lambda = -pf/dy^2 off the diagonal, c = 1 + 2 pf/dy^2 on it.  The jtot rows
// are cut into S chunks of L rows: an interface row sL and L-1 interior rows
// sL+1 .. sL+L-1.
//   * With the interface values p_s known, the interior rows of chunk s
//     solve T y = q (T: the (L-1)-row Dirichlet block, the same for every
//     chunk and column) corrected by the spikes of the two couplings:
//       x = y - lambda p_s v - lambda p_{s+1} w,   T v = e_0, w = reverse(v)
//     (T is symmetric and persymmetric, so T w = e_{L-2}).
//   * Substituting the interior rows into the interface equations leaves a
//     cyclic tridiagonal system in p with constant coefficients
//       O p_{s-1} + D p_s + O p_{s+1} = q_sL - lambda (y_{s-1,L-2} + y_{s,0}),
//       D = c - 2 lambda^2 v_0,  O = -lambda^2 v_{L-2},
//     solved per column as the reference solves periodic systems
//     (src/linalg.c:52-108): the (S-1)-row block with the right-hand side
//     and with [-O, 0, .., 0, -O], then the last unknown, then the rest.
// Three kernels: k_ylocal (the interior solves; the first and last interior
// values of each chunk go to qz rows S..3S-1), k_yreduced (the interface
// system, one wave per column, into qz rows 0..S-1), k_yapply (the
// corrections and out = in + x with boundary values).  When there are
// enough chunk columns to fill the GPU, k_yapply solves the interiors again,
// bit for bit the same, instead of reading them back: a few flops per cell
// for 32 instead of 40 bytes per cell.  Otherwise k_ylocal also stores the
// interior solution and k_yapply_rows applies it row by row (one thread per
// cell, which keeps small instances busy).  The factors of T and of the reduced system depend on
// dt; k_yfactor computes them once per step for the three substages and
// fields.  Rounding differs from the reference's sequential sweep at the
// ulp level.
#include "sns_launch.h"
#include "sns_fast.h"
#include "sns_fused.h"
#include "sns_lu.h"

namespace sns {

constexpr int kYMaxL = 32;  // chunk length bound
constexpr int kYRedNT = 64;  // threads of the reduced solve (one wave)
constexpr int kYRedM = 8;    // reduced unknowns per thread: S-1 <= 512
constexpr int kYHead = 8;

// Chunk length: the longest that still leaves at least 32 chunks (the
// interior solves are sequential within a chunk, so short columns need more,
// shorter chunks to fill the GPU), else the longest with at least 4.
int ysolve_chunk(int jtot) {
  for (int minS : {32, 4})
    for (int L = kYMaxL; L >= 4; L--)
      if (jtot % L == 0 && jtot / L >= minS && jtot / L - 1 <= kYRedNT * kYRedM) return L;
  return 0;
}

// A chunk length the kernels take: 4..32 rows (the interior solves hold a
// chunk in registers), dividing jtot into 4..513 chunks (the reduced
// system's tables hold up to 512 rows).
bool ysolve_chunk_ok(int jtot, int L) {
  return L >= 4 && L <= kYMaxL && jtot % L == 0 && jtot / L >= 4 &&
         jtot / L - 1 <= kYRedNT * kYRedM;
}

size_t ysolve_table_doubles(int L) {
  return kYHead + 3 * (size_t)(L - 1) + 4 * (size_t)(kYRedNT * kYRedM);
}

// Table of one (substage, field): header [lambda, c, D, O, den, 0, 0, 0],
// the interior factors r, beta and the spike v (L-1 entries each), then the
// reduced (S-1)-row block's interleaved LU tables r, al, be (sns_lu.h) and
// its matrix-only solution q1 in the same layout.
struct YTab {
  const double *p;
  int L1;
  __device__ __forceinline__ double lambda() const { return p[0]; }
  __device__ __forceinline__ double O() const { return p[3]; }
  __device__ __forceinline__ double den() const { return p[4]; }
  __device__ __forceinline__ const double *r() const { return p + kYHead; }
  __device__ __forceinline__ const double *beta() const { return p + kYHead + L1; }
  __device__ __forceinline__ const double *v() const { return p + kYHead + 2 * L1; }
  __device__ __forceinline__ const double *red(int k) const {
    return p + kYHead + 3 * L1 + (size_t)k * kYRedNT * kYRedM;
  }
};

template <int KIND>
__device__ __forceinline__ double cny_pf(const Geo &g, double gamma, double dt) {
  // T:     (gamma*dt)/(2.*sqrt(Pr)*sqrt(Ra))   src/temperature/update.c:142
  // ux,uy: (gamma*dt*sqrt(Pr))/(2.*sqrt(Ra))   src/fluid/update_velocity.c:254,399
  return KIND == KIND_TEMP ? (gamma * dt) / g.two_sqrtPr_sqrtRa
                           : (gamma * dt * g.sqrtPr) / g.two_sqrtRa;
}

// ------------------------------------------------------------- factors

// Workgroup (substage blockIdx.x, field blockIdx.y); thread 0 runs the
// sequential recurrences in LDS (each step reads what the previous one
// wrote: an LDS round trip instead of an HBM one), then all threads copy
// the tables out, zeros past the reduced block's last row included.
template <class S>
__global__ void __launch_bounds__(64) k_yfactor(S src) {
  const YFactorArgs a = src.get();
  constexpr int NR = kYRedNT * kYRedM;
  __shared__ double sr[kYMaxL], sbeta[kYMaxL], sv[kYMaxL];
  __shared__ double srr[NR], sal[NR], sbe[NR], sq1[NR];
  __shared__ double shead[kYHead];
  const int s = blockIdx.x, f = blockIdx.y;
  const int L1 = a.L - 1, SR = a.S - 1;  // interior rows; rows of the reduced block
  double *t = a.out + (size_t)(s * 3 + f) * a.stride;
  for (int k = threadIdx.x; k < NR; k += 64) srr[k] = sal[k] = sbe[k] = sq1[k] = 0.;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double gam = s == 0 ? a.gamma[0] : s == 1 ? a.gamma[1] : a.gamma[2];
    const double dt = *a.dtp;
    const double pf = f == 0 ? cny_pf<KIND_TEMP>(a.g, gam, dt) : cny_pf<KIND_UX>(a.g, gam, dt);
    // coefficients as src/temperature/update.c:188-197
    const double coef_ym = 1. / a.g.dy / a.g.dy, coef_yp = 1. / a.g.dy / a.g.dy;
    const double coef_y0 = -coef_ym - coef_yp;
    const double lam = -pf * coef_ym;
    const double c = 1. - pf * coef_y0;
    // interior Dirichlet block: r_m = 1/pivot, beta_m = lambda r_m (u')
    double prev_beta = 0.;
    for (int m = 0; m < L1; m++) {
      const double piv = m == 0 ? c : c - lam * prev_beta;
      sr[m] = 1. / piv;
      sbeta[m] = lam * sr[m];
      prev_beta = sbeta[m];
    }
    // spike v = T^{-1} e_0
    double y = 0.;
    for (int m = 0; m < L1; m++) {
      y = ((m == 0 ? 1. : 0.) - lam * y) * sr[m];
      sv[m] = y;
    }
    for (int m = L1 - 2; m >= 0; m--) sv[m] = sv[m] - sbeta[m] * sv[m + 1];
    const double lam2 = lam * lam;
    const double D = c - lam2 * sv[0] - lam2 * sv[0];
    const double O = -(lam2 * sv[L1 - 1]);
    // reduced (S-1)-row block: interleaved factors (row k at lu_slot), and
    // q1, the block's solution for the right-hand side [-O, 0, .., 0, -O]:
    // its forward sweep in the same loop, both carried in registers
    auto slot = [](int k) { return (k % kYRedM) * kYRedNT + k / kYRedM; };
    double rp = 0., bp = 0., yq = 0.;
    for (int k = 0; k < SR; k++) {
      const double piv = k == 0 ? D : D - O * bp;
      rp = 1. / piv;
      bp = O * rp;
      const double alk = k == 0 ? 0. : -(rp * O);
      srr[slot(k)] = rp;
      sal[slot(k)] = alk;
      sbe[slot(k)] = k == SR - 1 ? 0. : -bp;
      const double rhs = (k == 0 ? -O : 0.) + (k == SR - 1 ? -O : 0.);
      yq = rp * rhs + alk * yq;
      sq1[slot(k)] = yq;
    }
    double xq = sq1[slot(SR - 1)];
    for (int k = SR - 2; k >= 0; k--) {
      xq = sq1[slot(k)] + sbe[slot(k)] * xq;
      sq1[slot(k)] = xq;
    }
    shead[0] = lam;
    shead[1] = c;
    shead[2] = D;
    shead[3] = O;
    shead[4] = D + O * sq1[slot(0)] + O * sq1[slot(SR - 1)];
    shead[5] = shead[6] = shead[7] = 0.;
  }
  __syncthreads();
  double *r = t + kYHead, *beta = r + L1, *v = beta + L1;
  double *rr = t + kYHead + 3 * L1, *al = rr + NR, *be = al + NR, *q1 = be + NR;
  if (threadIdx.x < kYHead) t[threadIdx.x] = shead[threadIdx.x];
  for (int m = threadIdx.x; m < L1; m += 64) {
    r[m] = sr[m];
    beta[m] = sbeta[m];
    v[m] = sv[m];
  }
  for (int k = threadIdx.x; k < NR; k += 64) {
    rr[k] = srr[k];
    al[k] = sal[k];
    be[k] = sbe[k];
    q1[k] = sq1[k];
  }
}

// ------------------------------------------------------ interior solves

// Interior solve of chunk s, column col: the L-1 rows in x[OFF..] (Dirichlet
// block, shared factors), from the increments qf.
template <int OFF, int N>
__device__ __forceinline__ void ychunk_solve(const YArgs &a, const YTab &tb, int s, int col,
                                             double (&x)[N]) {
  static_assert(N >= OFF + kYMaxL - 1, "room for the interior rows");
  const int L1 = a.L - 1;
  const double lam = tb.lambda();
  const double *q = a.qf + ((size_t)s * a.L + 1) * a.n + col;
#pragma unroll
  for (int m = 0; m < kYMaxL - 1; m++)
    if (m < L1) x[OFF + m] = q[(size_t)m * a.n];
  double y = 0.;
#pragma unroll
  for (int m = 0; m < kYMaxL - 1; m++)
    if (m < L1) {
      y = (x[OFF + m] - lam * y) * tb.r()[m];
      x[OFF + m] = y;
    }
#pragma unroll
  for (int m = kYMaxL - 3; m >= 0; m--)
    if (m < L1 - 1) x[OFF + m] = x[OFF + m] - tb.beta()[m] * x[OFF + m + 1];
}

// Chunk and column of a thread of the per-chunk kernels: the four waves of a
// workgroup cover CW adjacent runs of 64 columns of 4/CW chunks, so each row
// access of the workgroup is CW*512 contiguous bytes.
template <int CW>
__device__ __forceinline__ void ychunk_of(int &s, int &col) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  s = blockIdx.y * (4 / CW) + wave / CW;
  col = (blockIdx.x * CW + wave % CW) * 64 + lane;
}

// Thread of chunk s, column col (ychunk_of).  The chunk's first and last
// interior values go to qz rows S+s and 2S+s; STORE: the interior solution
// also replaces the increments.
template <bool STORE, int CW, class S>
__global__ void __launch_bounds__(256) k_ylocal(S src) {
  const YArgs a = src.get();
  int s, col;
  ychunk_of<CW>(s, col);
  if (s >= a.S || col >= a.n) return;
  const int L1 = a.L - 1;
  const YTab tb{a.yfac, L1};
  double x[kYMaxL - 1];
  ychunk_solve<0>(a, tb, s, col, x);
  double last = x[0];
#pragma unroll
  for (int m = 1; m < kYMaxL - 1; m++)
    if (m == L1 - 1) last = x[m];
  a.qz[(size_t)(a.S + s) * a.n + col] = x[0];
  a.qz[(size_t)(2 * a.S + s) * a.n + col] = last;
  if (STORE) {
    double *q = a.qf + ((size_t)s * a.L + 1) * a.n + col;
#pragma unroll
    for (int m = 0; m < kYMaxL - 1; m++)
      if (m < L1) q[(size_t)m * a.n] = x[m];
  }
}

// -------------------------------------------------- interface system

// One wave per column blockIdx.x: the S interface values into qz (row k at
// qz[k*n + col]).
template <class S>
__global__ void __launch_bounds__(kYRedNT) k_yreduced(S src) {
  const YArgs a = src.get();
  __shared__ double scr[4 * 1 * (kYRedNT / 64)];
  __shared__ double ends[3];
  const int col = blockIdx.x;
  if (col >= a.n) return;  // the launch covers the widest field of a group
  const int n = a.n, L = a.L, SR = a.S - 1;
  const YTab tb{a.yfac, L - 1};
  const double lam = tb.lambda(), O = tb.O();
  const double *q = a.qf + col;
  const double *yfirst = a.qz + (size_t)a.S * n + col, *ylast = a.qz + (size_t)2 * a.S * n + col;
  // right-hand sides of the interface rows (chunk k: row kL; its neighbours
  // are the last interior row of chunk k-1 and the first of chunk k)
  auto rhs = [&](int k) {
    const int kl = k == 0 ? a.S - 1 : k - 1;
    return q[(size_t)k * L * n] - lam * (ylast[(size_t)kl * n] + yfirst[(size_t)k * n]);
  };
  double d[1][kYRedM];
#pragma unroll
  for (int kk = 0; kk < kYRedM; kk++) {
    const int k = threadIdx.x * kYRedM + kk;
    d[0][kk] = k < SR ? rhs(k) : 0.;
  }
  const LuTab lt = {tb.red(0), tb.red(1), tb.red(2)};
  lu_solve_shared<kYRedNT, kYRedM, 1>(d, lt, scr);
  // the last unknown from the first and last rows of the block
#pragma unroll
  for (int kk = 0; kk < kYRedM; kk++) {
    const int k = threadIdx.x * kYRedM + kk;
    if (k == 0) ends[0] = d[0][kk];
    if (k == SR - 1) ends[1] = d[0][kk];
  }
  __syncthreads();
  if (threadIdx.x == 0) ends[2] = (rhs(SR) - O * ends[0] - O * ends[1]) / tb.den();
  __syncthreads();
  const double zl = ends[2];
  double *z = a.qz + col;
  const double *q1 = tb.red(3);
#pragma unroll
  for (int kk = 0; kk < kYRedM; kk++) {
    const int k = threadIdx.x * kYRedM + kk;
    if (k < SR) z[(size_t)k * n] = d[0][kk] + zl * q1[kk * kYRedNT + threadIdx.x];
  }
  if (threadIdx.x == 0) z[(size_t)SR * n] = zl;
}

// The same systems for CW columns per workgroup with the lanes over
// columns, so loads and stores are coalesced in runs of CW: lane l of wave w
// holds column l % CW and chunk c = w*(64/CW) + l/CW of the (S-1)-row block,
// rows c*M .. c*M+M-1.  Each chunk runs the two substitutions of
// lu_solve_shared with zero carries, and the carries are composed across the
// chunks in LDS (the chunking differs from k_yreduced, so the rounding
// differs at the ulp level).  Rows past S-2 meet the zero padding of the
// factor tables.  Fewer columns per workgroup make more, shorter chunks,
// at the price of a longer carry scan over the chunks.
constexpr int kYColWaves = 16;

template <int CW, int M, class S>
__global__ void __launch_bounds__(64 * kYColWaves) k_yreduced_c(S src) {
  const YArgs a = src.get();
  constexpr int SUB = 64 / CW, NCH = kYColWaves * SUB;
  __shared__ double sA[NCH][CW], sB[NCH][CW];
  __shared__ double sends[2][CW];
  // the wave index is uniform: with one chunk per wave the factor-table
  // reads become scalar loads
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cl = lane % CW, c = w * SUB + lane / CW;  // column in the group; chunk
  const int n = a.n, L = a.L, SR = a.S - 1;
  const int col = blockIdx.x * CW + cl;
  const int cc = col < n ? col : n - 1;  // lanes past the last column repeat it
  const YTab tb{a.yfac, L - 1};
  const double lam = tb.lambda(), O = tb.O();
  const double *q = a.qf + cc;
  const double *yfirst = a.qz + (size_t)a.S * n + cc, *ylast = a.qz + (size_t)2 * a.S * n + cc;
  auto rhs = [&](int k) {
    const int kl = k == 0 ? a.S - 1 : k - 1;
    return q[(size_t)k * L * n] - lam * (ylast[(size_t)kl * n] + yfirst[(size_t)k * n]);
  };
  auto slot = [](int k) { return (k % kYRedM) * kYRedNT + k / kYRedM; };
  const double *rr = tb.red(0), *al = tb.red(1), *be = tb.red(2), *q1 = tb.red(3);
  // the running products of al and be are formed again where they are used
  // (same factors, same order) instead of being held per row
  double d[M];
  // forward substitution of the chunk with a zero carry
  double y = 0., pr = 1.;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int k = c * M + m;
    const double r = rr[slot(k)], l = al[slot(k)];
    pr *= l;
    const double dk = k < SR ? rhs(k) : 0.;
    y = r * dk + l * y;
    d[m] = y;
  }
  sA[c][cl] = pr;
  sB[c][cl] = y;
  __syncthreads();
  double cin = 0.;
  for (int v = 0; v < c; v++) cin = sA[v][cl] * cin + sB[v][cl];
  pr = 1.;
#pragma unroll
  for (int m = 0; m < M; m++) {
    pr *= al[slot(c * M + m)];
    d[m] += pr * cin;
  }
  __syncthreads();
  // back substitution with a zero carry from the right
  double x = 0.;
  pr = 1.;
#pragma unroll
  for (int m = M - 1; m >= 0; m--) {
    const double b = be[slot(c * M + m)];
    pr *= b;
    x = d[m] + b * x;
    d[m] = x;
  }
  sA[c][cl] = pr;
  sB[c][cl] = x;
  __syncthreads();
  cin = 0.;
  for (int v = NCH - 1; v > c; v--) cin = sA[v][cl] * cin + sB[v][cl];
  pr = 1.;
#pragma unroll
  for (int m = M - 1; m >= 0; m--) {
    pr *= be[slot(c * M + m)];
    d[m] += pr * cin;
  }
  // the last unknown from the first and last rows of the block
  if (c == 0) sends[0][cl] = d[0];
#pragma unroll
  for (int m = 0; m < M; m++)
    if (c * M + m == SR - 1) sends[1][cl] = d[m];
  __syncthreads();
  const double zl = (rhs(SR) - O * sends[0][cl] - O * sends[1][cl]) / tb.den();
  if (col >= n) return;
  double *z = a.qz + col;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int k = c * M + m;
    if (k < SR) z[(size_t)k * n] = d[m] + zl * q1[slot(k)];
  }
  if (c == 0) z[(size_t)SR * n] = zl;
}

// Launch of k_yreduced_c for blocks of up to 256 rows: the chunk length M
// from the number of chunks a workgroup holds
template <int CW, class S>
static void yreduced_c_launch(hipStream_t s, int n, int SR, S src, int nz) {
  constexpr int NCH = kYColWaves * (64 / CW);
  const dim3 g((n + CW - 1) / CW, 1, nz), b(64 * kYColWaves);
  if (SR <= NCH) hipLaunchKernelGGL((k_yreduced_c<CW, 1, S>), g, b, 0, s, src);
  else if (SR <= 2 * NCH) hipLaunchKernelGGL((k_yreduced_c<CW, 2, S>), g, b, 0, s, src);
  else if (SR <= 4 * NCH) hipLaunchKernelGGL((k_yreduced_c<CW, 4, S>), g, b, 0, s, src);
  else if constexpr (CW >= 32) {
    if (SR <= 8 * NCH) hipLaunchKernelGGL((k_yreduced_c<CW, 8, S>), g, b, 0, s, src);
    else if constexpr (CW == 64) hipLaunchKernelGGL((k_yreduced_c<CW, 16, S>), g, b, 0, s, src);
  }
}

// ------------------------------------------------------------- apply

// The substage-start value of unknown ii of row `row` (cell row j = row+1)
__device__ __forceinline__ double yin(const YArgs &a, int ii, int row) {
  const Geo &g = a.g;
  return a.kind == KIND_UX ? a.in[ux_idx(g, ii + 2, row + 1)] : a.in[c_idx(g, ii + 1, row + 1)];
}

// out at unknown ii of row `row` with boundary values
__device__ __forceinline__ void yset(const YArgs &a, int ii, int row, double v) {
  const Geo &g = a.g;
  const int j = row + 1;
  if (a.kind == KIND_UX) put_ux(g, a.out, ii + 2, j, v);
  else if (a.kind == KIND_UY) put_c<0>(g, a.out, ii + 1, j, v);
  else put_c<2>(g, a.out, ii + 1, j, v);
}

// out = in + x with boundary values
__device__ __forceinline__ void yput(const YArgs &a, int ii, int row, double x) {
  yset(a, ii, row, yin(a, ii, row) + x);
}

// Row blockIdx.y (j = row+1) after k_ylocal<true>: x = p_s on an interface
// row, else the stored interior solution with the spikes of both
// interfaces.
template <class S>
__global__ void __launch_bounds__(256) k_yapply_rows(S src) {
  const YArgs a = src.get();
  const int n = a.n;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (ii >= n) return;
  const int s = row / a.L, m = row - s * a.L;
  const int s1 = s + 1 == a.S ? 0 : s + 1;
  const YTab tb{a.yfac, a.L - 1};
  const double zs = a.qz[(size_t)s * n + ii];
  double x;
  if (m == 0) {
    x = zs;
  } else {
    const double lam = tb.lambda();
    const double zn = a.qz[(size_t)s1 * n + ii];
    const double y = a.qf[(size_t)row * n + ii];
    x = y - lam * zs * tb.v()[m - 1] - lam * zn * tb.v()[a.L - 1 - m];
  }
  yput(a, ii, row, x);
}

// Chunk s, column ii (ychunk_of): x = p_s on the interface row, the
// interior solution (solved again) with the spikes of both interfaces on the
// others; out = in + x with boundary values.  The substage-start values are
// loaded kYApplyG rows at a time ahead of those rows' stores: out may be in,
// so the compiler does not move a load past a store by itself.
constexpr int kYApplyG = 8;

template <int CW, class S>
__global__ void __launch_bounds__(256) k_yapply(S src) {
  const YArgs a = src.get();
  int s, ii;
  ychunk_of<CW>(s, ii);
  if (s >= a.S || ii >= a.n) return;
  const int n = a.n, L1 = a.L - 1, r0 = s * a.L;
  const int s1 = s + 1 == a.S ? 0 : s + 1;
  const YTab tb{a.yfac, L1};
  const double lam = tb.lambda();
  const double zs = a.qz[(size_t)s * n + ii], zn = a.qz[(size_t)s1 * n + ii];
  double x[kYMaxL];
  ychunk_solve<1>(a, tb, s, ii, x);
  x[0] = zs;
#pragma unroll
  for (int m = 1; m < kYMaxL; m++)
    if (m <= L1) x[m] = x[m] - lam * zs * tb.v()[m - 1] - lam * zn * tb.v()[a.L - 1 - m];
#pragma unroll
  for (int m0 = 0; m0 < kYMaxL; m0 += kYApplyG) {
    double v[kYApplyG];
#pragma unroll
    for (int k = 0; k < kYApplyG; k++)
      if (m0 + k <= L1) v[k] = yin(a, ii, r0 + m0 + k);
#pragma unroll
    for (int k = 0; k < kYApplyG; k++)
      if (m0 + k <= L1) yset(a, ii, r0 + m0 + k, v[k] + x[m0 + k]);
  }
}

// ------------------------------------------------------------- launch

// Chunk columns (threads of the per-chunk kernels) from which the interiors
// are solved twice rather than stored: 4096 waves, unless the instance's
// launch shapes say (1 recompute, 2 store).  SNSG_YSOLVE_RECOMPUTE=1 or =0
// forces either way (both return the same bits).
constexpr size_t kYRecompute = (size_t)4096 * 64;

static bool y_recompute(size_t columns, int chosen) {
  const char *e = getenv("SNSG_YSOLVE_RECOMPUTE");
  if (e && (e[0] == '0' || e[0] == '1') && e[1] == 0) return e[0] == '1';
  if (chosen == 1 || chosen == 2) return chosen == 1;
  return columns >= kYRecompute;
}

// Columns per workgroup of the reduced solve: the instance's launch shape,
// else the widest whose chunks stay at 4 rows or fewer (measured at 4096
// columns: 17 us with 16 columns and chunks of 4 rows against 24 us with 64
// and chunks of 16; at 1024 columns and 63 rows: 7.5 us with 64 against
// 9.3 us with 16, although 16 columns give four times the workgroups).
// SNSG_YSOLVE_COLS=16, 32 or 64 forces a width (the chunking, and so the
// rounding, differs with it).
int ysolve_cols_rule(int SR) {
  return SR <= 4 * kYColWaves ? 64 : SR <= 8 * kYColWaves ? 32 : 16;
}

static int y_cols(int SR, int chosen) {
  const char *e = getenv("SNSG_YSOLVE_COLS");
  if (e) {
    const int v = atoi(e);
    if (v == 16 || v == 32 || v == 64) return v;
  }
  if (chosen == 16 || chosen == 32 || chosen == 64) return chosen;
  return ysolve_cols_rule(SR);
}

// The partitioned solve of nz argument blocks (instances, or fields of one
// or more instances; each kernel checks its block's n): n the widest n
// among them, nrule the instances the recompute rule counts, per field.
template <int CW, class S>
static void ylocal_launch(hipStream_t s, int n, int S_, bool recompute, S src, int nz) {
  const dim3 gc((n + 64 * CW - 1) / (64 * CW), (S_ + 4 / CW - 1) / (4 / CW), nz);
  if (recompute) hipLaunchKernelGGL((k_ylocal<false, CW, S>), gc, dim3(256), 0, s, src);
  else hipLaunchKernelGGL((k_ylocal<true, CW, S>), gc, dim3(256), 0, s, src);
}

template <int CW, class S>
static void yapply_launch(hipStream_t s, int n, int S_, S src, int nz) {
  const dim3 gc((n + 64 * CW - 1) / (64 * CW), (S_ + 4 / CW - 1) / (4 / CW), nz);
  hipLaunchKernelGGL((k_yapply<CW, S>), gc, dim3(256), 0, s, src);
}

// Runs of 64 columns per workgroup of the per-chunk kernels: 4 from 2048
// columns (IMP-IMP steps at 4096x8192: 8.30 against 8.20 ms with 4 runs,
// 2048x4096: 2.23 against 2.20 ms; at 1024x2048 one run is as fast).
// SNSG_YSOLVE_CW=1, 2 or 4 forces a width (all return the same bits).
static int y_col_waves(int n) {
  const char *e = getenv("SNSG_YSOLVE_CW");
  if (e && (e[0] == '1' || e[0] == '2' || e[0] == '4') && e[1] == 0) return e[0] - '0';
  return n >= 2048 ? 4 : 1;
}

template <class S>
static void ypart(hipStream_t s, const YArgs &shape, int n, S src, int nz, int nrule) {
  const bool recompute = y_recompute((size_t)n * shape.S * nrule, shape.recompute);
  const int cw = y_col_waves(n);
  if (cw == 4) ylocal_launch<4>(s, n, shape.S, recompute, src, nz);
  else if (cw == 2) ylocal_launch<2>(s, n, shape.S, recompute, src, nz);
  else ylocal_launch<1>(s, n, shape.S, recompute, src, nz);
  const int SR = shape.S - 1;
  const int cols = y_cols(SR, shape.cols);
  if (SR > 16 * kYColWaves)
    hipLaunchKernelGGL(k_yreduced<S>, dim3(n, 1, nz), dim3(kYRedNT), 0, s, src);
  else if (cols == 64) yreduced_c_launch<64>(s, n, SR, src, nz);
  else if (cols == 32) yreduced_c_launch<32>(s, n, SR, src, nz);
  else yreduced_c_launch<16>(s, n, SR, src, nz);
  if (recompute) {
    if (cw == 4) yapply_launch<4>(s, n, shape.S, src, nz);
    else if (cw == 2) yapply_launch<2>(s, n, shape.S, src, nz);
    else yapply_launch<1>(s, n, shape.S, src, nz);
  } else {
    const dim3 gr((n + 255) / 256, shape.g.jtot, nz);
    hipLaunchKernelGGL((k_yapply_rows<S>), gr, dim3(256), 0, s, src);
  }
}

// Up to three fields' arguments by value, field blockIdx.z
struct YFields {
  YArgs f[3];
  __device__ __forceinline__ YArgs get() const {
    return blockIdx.z == 0 ? f[0] : (blockIdx.z == 1 ? f[1] : f[2]);
  }
};

static int widest(const YArgs *a, int nf) {
  int n = 0;
  for (int f = 0; f < nf; f++) n = a[f].n > n ? a[f].n : n;
  return n;
}

void launch_ypart_fields(hipStream_t s, const YArgs *a, int nf) {
  YFields src{};
  for (int f = 0; f < nf; f++) src.f[f] = a[f];
  ypart(s, a[0], widest(a, nf), src, nf, 1);
}

void launch_ypart_fields_batch(hipStream_t s, const YArgs &shape, const YArgs *dev, int nz,
                               int nf) {
  // every field's n is itot or itot-1, so itot covers them
  ypart(s, shape, shape.g.itot, arg_batch<YArgs>(dev), nf * nz, nz);
}

void launch_yfactor(hipStream_t s, const YFactorArgs &a) {
  hipLaunchKernelGGL(k_yfactor<ArgOne<YFactorArgs>>, dim3(3, 3), dim3(64), 0, s,
                     ArgOne<YFactorArgs>{a});
}

void launch_yfactor_batch(hipStream_t s, const YFactorArgs *dev, int nz) {
  hipLaunchKernelGGL(k_yfactor<ArgBatch<YFactorArgs>>, dim3(3, 3, nz), dim3(64), 0, s,
                     arg_batch<YFactorArgs>(dev));
}

}  // namespace sns
