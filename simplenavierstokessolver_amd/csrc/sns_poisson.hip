// sns_poisson.hip -- the pressure-potential Poisson solve (gfx950).
//
// Reference: src/fluid/compute_potential.c:15-168: RHS = div(u*)/(gamma dt),
// real FFT along y (FFTW r2c, unnormalised), one tri-diagonal system along x
// per wavenumber, inverse c2r FFT, psi = result / jtot with Neumann walls.
//
// The y transforms run as four-step FFTs, n = n1*n2, j = n2*j1 + j2,
// k = k1 + n1*k2:
//   forward  pass A : Y[k1*n2+j2] = W_n^(j2 k1) * DFT_n1 over j1 of x[n2 j1 + j2]
//            pass B : X[k1 + n1 k2] = DFT_n2 over j2 of Y[k1*n2+j2]
//   inverse  pass B': Y[k1*n2+j2] = W_n^(-j2 k1) * IDFT_n2 over k2 of Z[k1 + n1 k2]
//            pass A': x[n2 j1 + j2] = IDFT_n1 over k1 of Y[k1*n2+j2]
// Every pass works on a tile of 16 real columns (8 complex pairs = 128-byte
// row segments), so all global accesses are whole cache lines; two real
// columns are packed as one complex sequence, and the spectra of the pair are
// separated (forward) or combined (inverse) in pass B / B', which therefore
// handles the wavenumber rows k1 and n1-k1 together.
#include <float.h>
#include "sns_launch.h"
#include "sns_fast.h"
#include "sns_fft.h"
#include "sns_poisson.h"
#include "sns_lu.h"

namespace sns {

static constexpr int TILE = 8;   // complex pairs per tile (16 real columns)
static constexpr int NTF = 512;  // threads of the FFT passes
static constexpr int EMAXF = 4;  // complex values per thread and pass

// P2: every divisor a power of two (the quotients are products, see dv)
template <bool P2>
__device__ __forceinline__ double poisson_rhs2(const Geo &g, const Tables &tb, const double *ux,
                                               const double *uy, double inv_gdt, int i, int j) {
  // 1./(gamma*dt)*((ux_xp-ux_xm)/DXF(i) + (uy_yp-uy_ym)/dy)
  const double ax = dv<P2>(ux[ux_idx(g, i + 1, j)] - ux[ux_idx(g, i, j)], tb.dxf()[i - 1],
                           tb.rdxf()[i - 1]);
  const double ay = dv<P2>(uy[c_idx(g, i, j + 1)] - uy[c_idx(g, i, j)], tb.dy, tb.rdy);
  return inv_gdt * (+ax + ay);
}

__device__ __forceinline__ void put_psi(const Geo &g, double *f, int i, int j, double v) {
  // Neumann walls and periodic halo rows (src/fluid/boundary_conditions.c:39-55)
  f[c_idx(g, i, j)] = v;
  if (j == 1) f[c_idx(g, i, g.jtot + 1)] = v;
  if (j == g.jtot) f[c_idx(g, i, 0)] = v;
  if (i == 1 || i == g.itot) {
    const int iw = i == 1 ? 0 : g.itot + 1;
    f[c_idx(g, iw, j)] = v;
    if (j == 1) f[c_idx(g, iw, g.jtot + 1)] = v;
    if (j == g.jtot) f[c_idx(g, iw, 0)] = v;
  }
}

// ------------------------------------------------------------ forward A

// RHS of the column pair (ia, ia+1), ia odd, at 0-based row j in 16-byte
// loads: ux faces ia, ia+1 (one pair), ia+2; uy of rows j+1 and j+2 (one
// pair each).  Needs the padded state layout (ux rows and uy's column 1 on
// 16-byte boundaries, even pitches); columns past itot give 0.
template <bool P2>
__device__ __forceinline__ double2 poisson_rhs_pair(const Geo &g, const Tables &tb,
                                                    const double *ux, const double *uy,
                                                    double inv_gdt, int ia, int j) {
  const int itot = g.itot;
  const int ic = ia <= itot ? ia : 1;  // lanes past the last column read column 1
  const double *urow = ux + (size_t)(j + 1) * g.pu;
  const double2 u01 = ldr2(reinterpret_cast<const double2 *>(urow + ic - 1), 0);
  const double u2 = ldr(urow, min(ic + 1, itot));  // face ic+2 (only used when ic+1 <= itot)
  const double2 v0 = ldr2(reinterpret_cast<const double2 *>(uy + (size_t)(j + 1) * g.pc + ic), 0);
  const double2 v1 = ldr2(reinterpret_cast<const double2 *>(uy + (size_t)(j + 2) * g.pc + ic), 0);
  double dxa, rdxa, dxb, rdxb;
  if (tb.uniform) {
    dxa = dxb = tb.u_dxf;
    rdxa = rdxb = tb.u_rdxf;
  } else {
    dxa = tb.dxf()[ic - 1];
    rdxa = tb.rdxf()[ic - 1];
    dxb = tb.dxf()[min(ic, itot - 1)];
    rdxb = tb.rdxf()[min(ic, itot - 1)];
  }
  // the expressions of poisson_rhs2, column by column
  const double axa = dv<P2>(u01.y - u01.x, dxa, rdxa);
  const double aya = dv<P2>(v1.x - v0.x, tb.dy, tb.rdy);
  const double axb = dv<P2>(u2 - u01.y, dxb, rdxb);
  const double ayb = dv<P2>(v1.y - v0.y, tb.dy, tb.rdy);
  double2 z;
  z.x = ia <= itot ? inv_gdt * (+axa + aya) : 0.;
  z.y = ia + 1 <= itot ? inv_gdt * (+axb + ayb) : 0.;
  return z;
}

template <int L, bool P2, bool AL>
__device__ __forceinline__ void pfwd_a(const Geo &g, const Tables &tb, const Fft2Plan &pl,
                                       double gamma, const double *__restrict__ dtp,
                                       const double *__restrict__ ux,
                                       const double *__restrict__ uy,
                                       const double2 *__restrict__ tw, double2 *__restrict__ Y) {
  __shared__ double2 buf[EMAXF * NTF];
  // consecutive tiles of one row group on the same XCD: a tile's last ux
  // face is the first of the next tile's 128-byte line, which then stays in
  // that XCD's L2
  const int nb = gridDim.x * gridDim.y;
  const int w = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, nb);
  const int tile = w % gridDim.x, j20 = (w / gridDim.x) * pl.G;
  const int S = pl.G * TILE, n1 = pl.n1, n2 = pl.n2;
  const FastDiv fS(S);
  const double inv_gdt = 1. / (gamma * (*dtp));
  for (int e = threadIdx.x; e < S * n1; e += NTF) {
    const int j1 = fS.div(e), s = e - j1 * S;
    const int gg = s / TILE, q = s - gg * TILE;
    const int j = n2 * j1 + j20 + gg;  // 0-based row
    const int ia = 2 * (tile * TILE + q) + 1;
    double2 z;
    if constexpr (AL) {
      z = poisson_rhs_pair<P2>(g, tb, ux, uy, inv_gdt, ia, j);
    } else {
      z.x = ia <= g.itot ? poisson_rhs2<P2>(g, tb, ux, uy, inv_gdt, ia, j + 1) : 0.;
      z.y = ia + 1 <= g.itot ? poisson_rhs2<P2>(g, tb, ux, uy, inv_gdt, ia + 1, j + 1) : 0.;
    }
    buf[e] = z;
  }
  __syncthreads();
  if constexpr (L > 0) fft_p2<NTF, EMAXF, L, -1>(buf, __builtin_ctz(S), tw, pl.n / L);
  else fft_il<NTF, EMAXF, -1>(buf, S, n1, pl.np1, pl.code1, tw, pl.n / n1);
  for (int e = threadIdx.x; e < S * n1; e += NTF) {
    const int k1 = fS.div(e), s = e - k1 * S;
    const int gg = s / TILE, q = s - gg * TILE;
    const int j2 = j20 + gg;
    const int pair = tile * TILE + q;
    if (pair < pl.npair) {
      double2 z = buf[e];
      if (k1 && j2) z = cmul(z, tw[j2 * k1]);
      Y[(size_t)(k1 * n2 + j2) * pl.npair + pair] = z;
    }
  }
}

// ------------------------------------------------------------ forward B

template <int L>
__device__ __forceinline__ void pfwd_b(const Geo &g, const Fft2Plan &pl,
                                       const double2 *__restrict__ tw,
                                       const double2 *__restrict__ Y,
                                       double2 *__restrict__ psihat) {
  __shared__ double2 buf[EMAXF * NTF];
  const int tile = blockIdx.x;
  const int n1 = pl.n1, n2 = pl.n2, n = pl.n;
  const int k1a = blockIdx.y, k1b = (n1 - k1a) % n1;
  const int nh = (k1a == k1b) ? 1 : 2;
  const int S = nh * TILE;
  const FastDiv fS(S);
  for (int e = threadIdx.x; e < S * n2; e += NTF) {
    const int j2 = fS.div(e), s = e - j2 * S;
    const int h = s / TILE, q = s - h * TILE;
    const int pair = tile * TILE + q;
    const int k1 = h ? k1b : k1a;
    buf[e] = pair < pl.npair ? Y[(size_t)(k1 * n2 + j2) * pl.npair + pair] : make_double2(0., 0.);
  }
  __syncthreads();
  if constexpr (L > 0) fft_p2<NTF, EMAXF, L, -1>(buf, __builtin_ctz(S), tw, n / L);
  else fft_il<NTF, EMAXF, -1>(buf, S, n2, pl.np2, pl.code2, tw, n / n2);
  // separate the two real columns of each pair: X_a = (Z_k + conj Z_{n-k})/2,
  // X_b = (Z_k - conj Z_{n-k})/(2i), for the wavenumbers k <= n/2 of this group
  for (int e = threadIdx.x; e < S * n2; e += NTF) {
    const int k2 = fS.div(e), s = e - k2 * S;
    const int h = s / TILE, q = s - h * TILE;
    const int k1 = h ? k1b : k1a;
    const int k = k1 + n1 * k2;
    if (2 * k > n) continue;
    // partner n-k = k1' + n1*k2'
    int k2p, hp;
    if (k1 == 0) {
      k2p = (n2 - k2) % n2;
      hp = h;
    } else {
      k2p = n2 - 1 - k2;
      hp = nh == 1 ? h : 1 - h;
    }
    const double2 zk = buf[e];
    const double2 zn = buf[k2p * S + hp * TILE + q];
    const int pair = tile * TILE + q;
    const int ia = 2 * pair + 1;
    double2 xa, xb;
    xa.x = 0.5 * (zk.x + zn.x);
    xa.y = 0.5 * (zk.y - zn.y);
    xb.x = 0.5 * (zk.y + zn.y);
    xb.y = 0.5 * (zn.x - zk.x);
    if (ia <= g.itot) psihat[(size_t)k * g.itot + (ia - 1)] = xa;
    if (ia + 1 <= g.itot) psihat[(size_t)k * g.itot + ia] = xb;
  }
}

// ------------------------------------------------------------ inverse B'

template <int L>
__device__ __forceinline__ void pinv_b(const Geo &g, const Fft2Plan &pl,
                                       const double2 *__restrict__ tw,
                                       const double2 *__restrict__ psihat,
                                       double2 *__restrict__ Y) {
  __shared__ double2 buf[EMAXF * NTF];
  const int tile = blockIdx.x;
  const int n1 = pl.n1, n2 = pl.n2, n = pl.n;
  const int k1a = blockIdx.y, k1b = (n1 - k1a) % n1;
  const int nh = (k1a == k1b) ? 1 : 2;
  const int S = nh * TILE;
  const FastDiv fS(S);
  // Hermitian extension of the two real columns' spectra packed as
  // Z_k = X_a,k + i X_b,k (src/fluid/compute_potential.c:147-149: c2r input).
  // Each stored wavenumber m <= n/2 is read once and fills the slots of k = m
  // and of its mirror n-m, which this workgroup holds too (row k1 pairs with
  // row n1-k1; the partner index is the one pass B separates with).
  for (int e = threadIdx.x; e < S * n2; e += NTF) {
    const int k2 = fS.div(e), s = e - k2 * S;
    const int h = s / TILE, q = s - h * TILE;
    const int k1 = h ? k1b : k1a;
    const int k = k1 + n1 * k2;
    if (2 * k > n) continue;  // filled from its mirror
    const int pair = tile * TILE + q;
    const int ia = 2 * pair + 1;
    double2 xa = make_double2(0., 0.), xb = make_double2(0., 0.);
    if (ia <= g.itot) xa = psihat[(size_t)k * g.itot + (ia - 1)];
    if (ia + 1 <= g.itot) xb = psihat[(size_t)k * g.itot + ia];
    if (k == 0 || 2 * k == n) { xa.y = 0.; xb.y = 0.; }
    buf[e] = make_double2(xa.x - xb.y, xa.y + xb.x);
    if (k != 0 && 2 * k != n) {
      int k2p, hp;
      if (k1 == 0) {
        k2p = (n2 - k2) % n2;
        hp = h;
      } else {
        k2p = n2 - 1 - k2;
        hp = nh == 1 ? h : 1 - h;
      }
      buf[k2p * S + hp * TILE + q] = make_double2(xa.x + xb.y, -xa.y + xb.x);
    }
  }
  __syncthreads();
  if constexpr (L > 0) fft_p2<NTF, EMAXF, L, +1>(buf, __builtin_ctz(S), tw, n / L);
  else fft_il<NTF, EMAXF, +1>(buf, S, n2, pl.np2, pl.code2, tw, n / n2);
  for (int e = threadIdx.x; e < S * n2; e += NTF) {
    const int j2 = fS.div(e), s = e - j2 * S;
    const int h = s / TILE, q = s - h * TILE;
    const int k1 = h ? k1b : k1a;
    const int pair = tile * TILE + q;
    if (pair < pl.npair) {
      double2 z = buf[e];
      if (k1 && j2) {
        double2 w = tw[j2 * k1];
        w.y = -w.y;
        z = cmul(z, w);
      }
      Y[(size_t)(k1 * n2 + j2) * pl.npair + pair] = z;
    }
  }
}

// ------------------------------------------------------------ inverse A'

template <int L, bool AL>
__device__ __forceinline__ void pinv_a(const Geo &g, const Fft2Plan &pl,
                                       const double2 *__restrict__ tw,
                                       const double2 *__restrict__ Y, double *__restrict__ psi) {
  __shared__ double2 buf[EMAXF * NTF];
  const int tile = blockIdx.x, j20 = blockIdx.y * pl.G;
  const int S = pl.G * TILE, n1 = pl.n1, n2 = pl.n2;
  const FastDiv fS(S);
  for (int e = threadIdx.x; e < S * n1; e += NTF) {
    const int k1 = fS.div(e), s = e - k1 * S;
    const int gg = s / TILE, q = s - gg * TILE;
    const int pair = tile * TILE + q;
    buf[e] = pair < pl.npair ? Y[(size_t)(k1 * n2 + j20 + gg) * pl.npair + pair]
                             : make_double2(0., 0.);
  }
  __syncthreads();
  if constexpr (L > 0) fft_p2<NTF, EMAXF, L, +1>(buf, __builtin_ctz(S), tw, pl.n / L);
  else fft_il<NTF, EMAXF, +1>(buf, S, n1, pl.np1, pl.code1, tw, pl.n / n1);
  // psi = q/jtot: the exact division of div_cr, or (jtot a power of two) the
  // product with the exact reciprocal; both equal IEEE q/jtot bit for bit
  const double njtot = 1. * g.jtot, rnj = 1. / njtot;
  const bool jp2 = (g.jtot & (g.jtot - 1)) == 0;
  for (int e = threadIdx.x; e < S * n1; e += NTF) {
    const int j1 = fS.div(e), s = e - j1 * S;
    const int gg = s / TILE, q = s - gg * TILE;
    const int j = n2 * j1 + j20 + gg + 1;  // 1-based row
    const int ia = 2 * (tile * TILE + q) + 1;
    const double2 z = buf[e];
    const double va = jp2 ? z.x * rnj : div_cr(z.x, njtot, rnj);
    const double vb = jp2 ? z.y * rnj : div_cr(z.y, njtot, rnj);
    const bool edge = j == 1 || j == g.jtot || ia == 1 || ia + 1 >= g.itot;
    if (AL && !edge) {
      // both columns in one 16-byte store (no halo or wall copies here)
      *reinterpret_cast<double2 *>(psi + c_idx(g, ia, j)) = make_double2(va, vb);
    } else {
      if (ia <= g.itot) put_psi(g, psi, ia, j, va);
      if (ia + 1 <= g.itot) put_psi(g, psi, ia + 1, j, vb);
    }
  }
}

// --------------------------------------------------- x solve per wavenumber

// Matrix of wavenumber J (src/fluid/compute_potential.c:117-127): l_i =
// 1/(DXC(i)DXF(i)), u_i = 1/(DXC(i+1)DXF(i)), c_i = -l_i - u_i + lambda_J with
// the Neumann walls folded into c_0 and c_{n-1}.  It does not change during
// a run, so its Thomas factorisation (src/linalg.c:32-50: u'_i and the
// reciprocal pivots, in the reference's sequential order) is computed once
// per instance by k_poisson_factor, one thread per wavenumber.  Only the
// reciprocal pivots r_i depend on J; the substitutions need al_i = -(r_i l_i)
// and be_i = -u'_i = -(r_i u_i) for 0 < i < n-1, so the factor table holds r
// (interleaved as in sns_lu.h, ntm entries per J) followed by u'_0 = u_0/c_0
// of every J, then per J the last row pfx <= n-2 whose pivot differs from
// the previous row's and that pivot (rows pfx..n-2 share it: on a grid with
// identical interior rows the recurrence reaches a fixed point, bit for bit,
// after a few rows for all but the lowest J); l, u are shared interleaved
// tables (LuTabLUr reads only rows up to pfx and the last row).
// The J = 0 system is singular; its last equation is replaced by
// psi_hat(itot) = 0 (r = 0 on that row).  Each solve then needs only the two
// substitutions, evaluated by lu_solve_shared for the real and imaginary
// parts together; with the factors exact, the substitutions keep the
// accuracy of the sequential sweep on the nearly singular low-J systems.
__device__ __forceinline__ int lu_slot(int i, int NT, int M) { return (i % M) * NT + i / M; }

__global__ void __launch_bounds__(64)
    k_poisson_factor(Geo g, Tables tb, const double *__restrict__ eigen, int NT, int M, int ntm,
                     double *__restrict__ fac) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  const int nJ = g.jtot / 2 + 1;
  if (J >= nJ) return;
  const int n = g.itot;
  const double lam = eigen[J];
  double *R = fac + (size_t)J * ntm;
  auto coef = [&](int i, double &l, double &c, double &u) {
    l = tb.pl()[i];
    u = tb.pu()[i];
    c = -l - u + lam;
    if (i == 0) c += l;
    if (i == n - 1) c += u;
  };
  double l, c, u;
  coef(0, l, c, u);
  double up = u / c;  // u'_0
  double r = 1. / c;
  R[lu_slot(0, NT, M)] = r;
  fac[(size_t)nJ * ntm + J] = up;
  int pfx = 0;
  double rpfx = r;
  for (int i = 1; i < n; i++) {
    coef(i, l, c, u);
    const double val = c - l * up;
    const double rp = r;
    if (fabs(val) < DBL_EPSILON || (J == 0 && i == n - 1)) {
      up = 0.;
      r = 0.;
    } else {
      r = 1. / val;
      up = r * u;
    }
    R[lu_slot(i, NT, M)] = r;
    if (i <= n - 2 && __double_as_longlong(r) != __double_as_longlong(rp)) {
      pfx = i;
      rpfx = r;
    }
  }
  fac[(size_t)nJ * ntm + nJ + J] = (double)pfx;
  fac[(size_t)nJ * ntm + 2 * nJ + J] = rpfx;
}

template <int NT, int M>
struct PxLds {
  double2 row[NT * (M + 1)];
  double scr[4 * 2 * (NT / 64)];
};

// One workgroup per wavenumber row; the row is staged through LDS (coalesced
// global access, chunk order for the substitutions).
template <int NT, int M>
__device__ __forceinline__ void poisson_x3(const Geo &g, const Tables &tb,
                                           const double *__restrict__ fac,
                                           const double *__restrict__ plu, int ntm,
                                           double2 *__restrict__ psihat) {
  extern __shared__ double lds_raw[];
  PxLds<NT, M> &L = *reinterpret_cast<PxLds<NT, M> *>(lds_raw);
  const int J = xcd_remap(blockIdx.x, gridDim.x);
  const int n = g.itot;
  double2 *row = psihat + (size_t)J * n;
  for (int r = threadIdx.x; r < n; r += NT) L.row[r + r / M] = row[r];
  __syncthreads();
  double d[2][M];
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    const double2 v = r < n ? L.row[r + r / M] : make_double2(0., 0.);
    d[0][k] = v.x;
    d[1][k] = v.y;
  }
  const int nJ = g.jtot / 2 + 1;
  const double *tail = fac + (size_t)nJ * ntm;
  const LuTabLUr<M> tab(fac + (size_t)J * ntm, plu, plu + ntm, n, tail[J], (int)tail[nJ + J],
                        tail[2 * nJ + J], NT, tb.plu_const != 0, tb.plc, tb.puc);
  lu_solve_shared<NT, M, 2>(d, tab, L.scr);
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int r = threadIdx.x * M + k;
    if (r < n) L.row[r + r / M] = make_double2(d[0][k], d[1][k]);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += NT) row[r] = L.row[r + r / M];
}

// ---------------------------------------------------------------- kernels
// Argument source S: ArgOne<PoisArgs> (one instance) or ArgBatch<PoisArgs>
// (instance blockIdx.z of a batch).

template <int L, class S>
__global__ void __launch_bounds__(NTF) k_pfwd_a(S src) {
  const PoisArgs a = src.get();
  pfwd_a<L, false, false>(a.g, a.tb, a.pl, a.gamma, a.dtp, a.ux, a.uy, a.tw, a.Y);
}

// the same on the padded state layout (pair loads), with the division
// sequence or, P2, with power-of-two divisors
template <int L, class S>
__global__ void __launch_bounds__(NTF) k_pfwd_a_al(S src) {
  const PoisArgs a = src.get();
  pfwd_a<L, false, true>(a.g, a.tb, a.pl, a.gamma, a.dtp, a.ux, a.uy, a.tw, a.Y);
}
template <int L, class S>
__global__ void __launch_bounds__(NTF) k_pfwd_a_p2(S src) {
  const PoisArgs a = src.get();
  pfwd_a<L, true, true>(a.g, a.tb, a.pl, a.gamma, a.dtp, a.ux, a.uy, a.tw, a.Y);
}

template <int L, class S>
__global__ void __launch_bounds__(NTF, 8) k_pfwd_b(S src) {
  const PoisArgs a = src.get();
  pfwd_b<L>(a.g, a.pl, a.tw, a.Y, a.psihat);
}

template <int NT, int M, class S>
__global__ void __launch_bounds__(NT) k_poisson_x3(S src) {
  const PoisArgs a = src.get();
  poisson_x3<NT, M>(a.g, a.tb, a.fac, a.plu, NT * M, a.psihat);
}

template <int L, class S>
__global__ void __launch_bounds__(NTF) k_pinv_b(S src) {
  const PoisArgs a = src.get();
  pinv_b<L>(a.g, a.pl, a.tw, a.psihat, a.Y);
}

// the same on the padded state layout (psi's column 1 on a 16-byte boundary)
template <int L, class S>
__global__ void __launch_bounds__(NTF, 8) k_pinv_a_al(S src) {
  const PoisArgs a = src.get();
  pinv_a<L, true>(a.g, a.pl, a.tw, a.Y, a.psi);
}

template <int L, class S>
__global__ void __launch_bounds__(NTF, 8) k_pinv_a(S src) {
  const PoisArgs a = src.get();
  pinv_a<L, false>(a.g, a.pl, a.tw, a.Y, a.psi);
}

// Threads NT and unknowns per thread M of the x solve for n = itot unknowns
// (NT*M >= n).  One workgroup solves one wavenumber row, so with only
// jtot/2+1 rows the narrower grids take 4 unknowns per thread (twice the
// waves per row) and the wider ones 8.
static void px_shape(int n, int &NT, int &M) {
  M = n <= 1024 ? 4 : 8;
  NT = 64;
  while (NT < 1024 && NT * M < n) NT *= 2;
}

int poisson_factor_entries(int itot) {
  int NT, M;
  px_shape(itot, NT, M);
  return NT * M;
}

size_t poisson_factor_doubles(int itot, int jtot) {
  return (size_t)(jtot / 2 + 1) * (poisson_factor_entries(itot) + 3);
}

void poisson_lu_coefs(int itot, const double *pl, const double *pu, double *plu) {
  int NT, M;
  px_shape(itot, NT, M);
  const int ntm = NT * M;
  for (int k = 0; k < 2 * ntm; k++) plu[k] = 0.;
  for (int i = 0; i < itot; i++) {
    const int slot = (i % M) * NT + i / M;
    plu[slot] = pl[i];
    plu[ntm + slot] = pu[i];
  }
}

void launch_poisson_factor(hipStream_t s, const Geo &g, const Tables &tb, const double *eigen,
                           double *fac) {
  int NT, M;
  px_shape(g.itot, NT, M);
  const int ntm = NT * M;
  const int nJ = g.jtot / 2 + 1;
  hipLaunchKernelGGL(k_poisson_factor, dim3((nJ + 63) / 64), dim3(64), 0, s, g, tb, eigen, NT,
                     M, ntm, fac);
}

template <int NT, int M, class S>
static void px_launch(hipStream_t s, const Geo &g, S src, int nz) {
  const size_t lds = sizeof(PxLds<NT, M>);
  static std::atomic<unsigned> attr{0};
  lds_limit_once((const void *)k_poisson_x3<NT, M, S>, lds, attr);
  hipLaunchKernelGGL((k_poisson_x3<NT, M, S>), dim3(g.jtot / 2 + 1, 1, nz), dim3(NT), lds, s, src);
}

// ------------------------------------------------------------- plan, launch

static bool radix_code(int m, int &np, unsigned long long &code) {
  int rest = m;
  np = 0;
  code = 0;
  auto push = [&](int r) {
    code |= (unsigned long long)r << (4 * np);
    np++;
    rest /= r;
  };
  while (rest % 8 == 0 && np < 16) push(8);
  while (rest % 4 == 0 && np < 16) push(4);
  while (rest % 2 == 0 && np < 16) push(2);
  while (rest % 3 == 0 && np < 16) push(3);
  while (rest % 5 == 0 && np < 16) push(5);
  return rest == 1;
}

bool fft2_plan(int n, int itot, Fft2Plan &pl, int n1) {
  if (n < 4) return false;
  // n1 = the given factor, else the largest divisor of n not above sqrt(n)
  // that the radices cover
  int best = -1;
  if (n1 > 0) {
    int np1, np2;
    unsigned long long c1, c2;
    if (n % n1 || n1 < 2 || n / n1 < 2 || !radix_code(n1, np1, c1) ||
        !radix_code(n / n1, np2, c2))
      return false;
    best = n1;
  } else {
    for (int d = 1; d * d <= n; d++) {
      if (n % d) continue;
      int np1, np2;
      unsigned long long c1, c2;
      if (radix_code(d, np1, c1) && radix_code(n / d, np2, c2)) best = d;
    }
  }
  if (best < 1) return false;
  pl.n = n;
  pl.n1 = best;
  pl.n2 = n / best;
  if (pl.n1 == 1) {  // keep both passes non-trivial only when possible
    return false;
  }
  radix_code(pl.n1, pl.np1, pl.code1);
  radix_code(pl.n2, pl.np2, pl.code2);
  if (2 * TILE * pl.n2 > EMAXF * NTF) return false;  // pass B must fit its LDS tile
  int G = 1;
  for (int c = 1; c <= pl.n2; c++)
    if (pl.n2 % c == 0 && c * TILE * pl.n1 <= EMAXF * NTF) G = c;
  pl.G = G;
  pl.npair = (itot + 1) / 2;
  return TILE * pl.n1 <= EMAXF * NTF;
}

static dim3 grid_a(const Fft2Plan &pl) { return dim3((pl.npair + TILE - 1) / TILE, pl.n2 / pl.G); }
static dim3 grid_b(const Fft2Plan &pl) { return dim3((pl.npair + TILE - 1) / TILE, pl.n1 / 2 + 1); }

static dim3 with_z(dim3 d, int nz) { return dim3(d.x, d.y, nz); }

// Compile-time length of a pass: the factor itself when it is a power of
// two in 8..128 and the pass's sequence count is a power of two, else 0 (the
// runtime radix schedule).
static int p2_len(int m, int S) {
  const bool pow2 = (m & (m - 1)) == 0 && (S & (S - 1)) == 0;
  return pow2 && m >= 8 && m <= 128 ? m : 0;
}

// Launch kernel template K<L, S> with L the compile-time length of factor m
#define SNS_P2_LAUNCH(K, L_, grid)                                                        \
  switch (L_) {                                                                           \
    case 8: hipLaunchKernelGGL((K<8, S>), grid, dim3(NTF), 0, s, src); break;              \
    case 16: hipLaunchKernelGGL((K<16, S>), grid, dim3(NTF), 0, s, src); break;            \
    case 32: hipLaunchKernelGGL((K<32, S>), grid, dim3(NTF), 0, s, src); break;            \
    case 64: hipLaunchKernelGGL((K<64, S>), grid, dim3(NTF), 0, s, src); break;            \
    case 128: hipLaunchKernelGGL((K<128, S>), grid, dim3(NTF), 0, s, src); break;          \
    default: hipLaunchKernelGGL((K<0, S>), grid, dim3(NTF), 0, s, src); break;             \
  }

template <class S>
static void poisson_stage(hipStream_t s, const Geo &g, const Fft2Plan &pl, bool pow2,
                          bool aligned, S src, int nz, int stage) {
  // pass A runs G*TILE sequences of length n1, pass B up to 2*TILE of length n2
  const int La = p2_len(pl.n1, pl.G * TILE);
  const int Lb = p2_len(pl.n2, TILE);
  if (stage == 0 || stage < 0) {
    if (aligned && pow2) SNS_P2_LAUNCH(k_pfwd_a_p2, La, with_z(grid_a(pl), nz))
    else if (aligned) SNS_P2_LAUNCH(k_pfwd_a_al, La, with_z(grid_a(pl), nz))
    else SNS_P2_LAUNCH(k_pfwd_a, La, with_z(grid_a(pl), nz));
    SNS_P2_LAUNCH(k_pfwd_b, Lb, with_z(grid_b(pl), nz));
  }
  if (stage == 1 || stage < 0) {
    int NT, M;
    px_shape(g.itot, NT, M);
    if (M == 4) {
      switch (NT) {
        case 64: px_launch<64, 4>(s, g, src, nz); break;
        case 128: px_launch<128, 4>(s, g, src, nz); break;
        default: px_launch<256, 4>(s, g, src, nz); break;
      }
    } else {
      switch (NT) {
        case 256: px_launch<256, 8>(s, g, src, nz); break;
        case 512: px_launch<512, 8>(s, g, src, nz); break;
        default: px_launch<1024, 8>(s, g, src, nz); break;
      }
    }
  }
  if (stage == 2 || stage < 0) {
    SNS_P2_LAUNCH(k_pinv_b, Lb, with_z(grid_b(pl), nz));
    if (aligned) SNS_P2_LAUNCH(k_pinv_a_al, La, with_z(grid_a(pl), nz))
    else SNS_P2_LAUNCH(k_pinv_a, La, with_z(grid_a(pl), nz));
  }
}

// The padded state layout: even row pitches, ux rows and the column 1 of uy
// and psi on 16-byte boundaries (every instance of a batch is allocated
// alike).
static bool pair_layout(const PoisArgs &a) {
  auto al16 = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
  return a.g.pu % 2 == 0 && a.g.pc % 2 == 0 && al16(a.ux) && al16(a.uy + 1) && al16(a.psi + 1);
}

void launch_poisson(hipStream_t s, const PoisArgs &a, int stage) {
  poisson_stage(s, a.g, a.pl, pow2_division(a.tb), pair_layout(a), ArgOne<PoisArgs>{a}, 1,
                stage);
}

void launch_poisson_batch(hipStream_t s, const PoisArgs &shape, const PoisArgs *dev, int nz,
                          int stage) {
  poisson_stage(s, shape.g, shape.pl, pow2_division(shape.tb), pair_layout(shape),
                arg_batch<PoisArgs>(dev), nz, stage);
}

}  // namespace sns
