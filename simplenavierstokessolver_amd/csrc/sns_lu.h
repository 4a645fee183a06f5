// sns_lu.h -- Crank-Nicolson line solves with a shared LU factorisation.
//
// The x-direction Crank-Nicolson operators (1 - pf*Dxx) of T, ux and uy
// (src/temperature/update.c:141-154, src/fluid/update_velocity.c:248-267,
// 398-411) depend on the column index and on gamma*dt only, so every row of a
// substage solves the same matrix.  The Thomas factorisation of that matrix
// (the reference's u'_i and 1/pivot, src/linalg.c:9-30) is computed once per
// substage by k_lu_factor; each row then needs only the two first-order
// recurrences
//   forward  y_i = r_i d_i + al_i y_{i-1},     al_i = -r_i l_i
//   backward x_i = y_i + be_i x_{i+1},         be_i = -u'_i
// evaluated in parallel: every thread runs them over its chunk of M rows with
// zero carries, the chunk-end values are combined by an affine scan across
// the workgroup (wave shuffles, then one value per wave), and each element
// is corrected with the running products of al (P) and be (Q) across its
// chunk.  Rounding differs from the sequential sweep at the ulp level.
//
// Tables are stored interleaved, entry (t, k) of row i = t*M + k at
// [k*NT + t], so that the loads of a wave are contiguous.
#pragma once
#include <hip/hip_runtime.h>

namespace sns {

struct LuTab {
  const double *r, *al, *be;  // each NT*M entries, interleaved
  // entry k of this thread's chunk (interleaved index k*NT + t)
  __device__ __forceinline__ double get_r(int k, int nt) const { return r[k * nt + threadIdx.x]; }
  __device__ __forceinline__ double get_al(int k, int nt) const { return al[k * nt + threadIdx.x]; }
  __device__ __forceinline__ double get_be(int k, int nt) const { return be[k * nt + threadIdx.x]; }
};

// LuTab with the repeated rows folded: rows P..n-2 all hold the triple
// (rc, alc, bec) bit for bit (k_lu_factor's header), so only rows below P,
// the last row and the padding rows (zeros) are read from the tables; a
// thread whose chunk lies in P..n-2 reads nothing.
template <int M>
struct LuTabC {
  const double *r, *al, *be;  // each NT*M entries, interleaved
  int P, n;
  double rc, alc, bec;
  __device__ __forceinline__ bool in_table(int k) const {
    const int row = threadIdx.x * M + k;
    return row < P || row >= n - 1;
  }
  __device__ __forceinline__ double get_r(int k, int nt) const {
    return in_table(k) ? r[k * nt + threadIdx.x] : rc;
  }
  __device__ __forceinline__ double get_al(int k, int nt) const {
    return in_table(k) ? al[k * nt + threadIdx.x] : alc;
  }
  __device__ __forceinline__ double get_be(int k, int nt) const {
    return in_table(k) ? be[k * nt + threadIdx.x] : bec;
  }
};

// The same factors when al_i = -(r_i*l_i) and be_i = -(r_i*u_i) hold for
// 0 < i < n-1 with row coefficients l, u that do not depend on the system
// (only r does): l and u are interleaved tables like r.  Row 0 has al = 0 and
// be = -up0 (the first pivot's u'_0 = u_0/c_0), the last row be = 0.
struct LuTabLU {
  const double *r, *l, *u;
  int n, m;    // unknowns; entries per thread
  double up0;  // u'_0 of this system
  __device__ __forceinline__ double get_r(int k, int nt) const { return r[k * nt + threadIdx.x]; }
  __device__ __forceinline__ double get_al(int k, int nt) const {
    const int row = threadIdx.x * m + k;
    return row == 0 ? 0. : -(r[k * nt + threadIdx.x] * l[k * nt + threadIdx.x]);
  }
  __device__ __forceinline__ double get_be(int k, int nt) const {
    const int row = threadIdx.x * m + k;
    return row == n - 1 ? 0.
                        : (row == 0 ? -up0 : -(r[k * nt + threadIdx.x] * u[k * nt + threadIdx.x]));
  }
};

// LuTabLU with this thread's M reciprocal pivots loaded once into registers.
// Rows pfx+1 .. n-2 all have the pivot rinf (the factorisation found every
// pivot from row pfx on equal, bit for bit, up to row n-2), so only rows up
// to pfx and the last row are read from the table: a thread whose chunk
// holds none of them reads nothing.
// With cst set, l of rows 1..n-1 and u of rows 0..n-2 are the constants lc,
// uc (the table entries themselves), so the solve reads no coefficient table.
template <int M>
struct LuTabLUr {
  double rr[M];
  const double *l, *u;
  int n;
  double up0;
  bool cst;
  double lc, uc;
  __device__ __forceinline__ LuTabLUr(const double *r, const double *l_, const double *u_,
                                      int n_, double up0_, int pfx, double rinf, int nt,
                                      bool cst_ = false, double lc_ = 0., double uc_ = 0.)
      : l(l_), u(u_), n(n_), up0(up0_), cst(cst_), lc(lc_), uc(uc_) {
    const int row0 = threadIdx.x * M;
    if (row0 <= pfx || row0 + M - 1 >= n - 1) {
#pragma unroll
      for (int k = 0; k < M; k++) {
        const int row = row0 + k;
        rr[k] = (row <= pfx || row == n - 1) ? r[k * nt + threadIdx.x] : rinf;
      }
    } else {
#pragma unroll
      for (int k = 0; k < M; k++) rr[k] = rinf;
    }
  }
  __device__ __forceinline__ double get_r(int k, int) const { return rr[k]; }
  __device__ __forceinline__ double get_al(int k, int nt) const {
    const int row = threadIdx.x * M + k;
    return row == 0 ? 0. : -(rr[k] * (cst ? lc : l[k * nt + threadIdx.x]));
  }
  __device__ __forceinline__ double get_be(int k, int nt) const {
    const int row = threadIdx.x * M + k;
    return row == n - 1 ? 0.
                        : (row == 0 ? -up0 : -(rr[k] * (cst ? uc : u[k * nt + threadIdx.x])));
  }
};

// composition of affine maps x -> A x + B: (A2,B2) after (A1,B1)
__device__ __forceinline__ void aff_after(double &A, double &B, double A1, double B1) {
  B = A * B1 + B;
  A = A * A1;
}

// Exclusive affine scan over the threads of a workgroup (forward: thread t
// receives the composition of maps 0..t-1 applied to 0; reverse: of maps
// NT-1..t+1).  NF independent scans share the barriers.  scr holds
// 2*NF*(NT/64) doubles; the caller orders any later write of scr after the
// reads here (another scan uses other scratch, and the callers' next LDS
// phase starts behind a barrier of their own).
template <int NT, int NF, bool REVERSE>
__device__ __forceinline__ void affine_scan_excl(double (&A)[NF], double (&B)[NF], double (&C)[NF],
                                                 double *scr) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // inclusive scan inside the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const double Ao = REVERSE ? __shfl_down(A[f], off) : __shfl_up(A[f], off);
      const double Bo = REVERSE ? __shfl_down(B[f], off) : __shfl_up(B[f], off);
      const bool ok = REVERSE ? (lane + off < 64) : (lane >= off);
      if (ok) aff_after(A[f], B[f], Ao, Bo);
    }
  }
  // wave totals
  const int last = REVERSE ? 0 : 63;
  if (lane == last) {
#pragma unroll
    for (int f = 0; f < NF; f++) {
      scr[(f * NW + wave) * 2 + 0] = A[f];
      scr[(f * NW + wave) * 2 + 1] = B[f];
    }
  }
  __syncthreads();
#pragma unroll
  for (int f = 0; f < NF; f++) {
    double cin = 0.;
    if (!REVERSE) {
      for (int w = 0; w < wave; w++) cin = scr[(f * NW + w) * 2] * cin + scr[(f * NW + w) * 2 + 1];
    } else {
      for (int w = NW - 1; w > wave; w--) cin = scr[(f * NW + w) * 2] * cin + scr[(f * NW + w) * 2 + 1];
    }
    const double incl = A[f] * cin + B[f];  // value after this thread's map
    const double prev = REVERSE ? __shfl_down(incl, 1) : __shfl_up(incl, 1);
    const bool edge = REVERSE ? (lane == 63) : (lane == 0);
    C[f] = edge ? cin : prev;
  }
}

// Solve NS systems that share one matrix (tab) for this thread's chunk
// d[s][0..M-1] (unknowns t*M .. t*M+M-1).  scr holds 4*NS*(NT/64) doubles
// (the two scans' scratch apart); the caller places a barrier between this
// call and its next write of scr.
template <int NT, int M, int NS, class TAB = LuTab>
__device__ __forceinline__ void lu_solve_shared(double (&d)[NS][M], const TAB &tab,
                                                double *scr) {
  double A[NS], B[NS], C[NS];
  double P[M];
  {
    double y[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) y[s] = 0.;
    double pr = 1.;
#pragma unroll
    for (int k = 0; k < M; k++) {
      const double r = tab.get_r(k, NT), al = tab.get_al(k, NT);
      pr *= al;
      P[k] = pr;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        y[s] = r * d[s][k] + al * y[s];
        d[s][k] = y[s];
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      A[s] = P[M - 1];
      B[s] = y[s];
    }
  }
  affine_scan_excl<NT, NS, false>(A, B, C, scr);
  double Q[M];
  {
    double x[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) x[s] = 0.;
    double pr = 1.;
#pragma unroll
    for (int k = M - 1; k >= 0; k--) {
      const double be = tab.get_be(k, NT);
      pr *= be;
      Q[k] = pr;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        d[s][k] += P[k] * C[s];
        x[s] = d[s][k] + be * x[s];
        d[s][k] = x[s];
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      A[s] = Q[0];
      B[s] = x[s];
    }
  }
  affine_scan_excl<NT, NS, true>(A, B, C, scr + 2 * NS * (NT / 64));
#pragma unroll
  for (int s = 0; s < NS; s++)
#pragma unroll
    for (int k = 0; k < M; k++) d[s][k] += Q[k] * C[s];
}

}  // namespace sns
