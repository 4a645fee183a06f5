// sns_fft.h -- batched complex FFT staged in LDS (gfx950).
//
// Stockham auto-sort passes (no bit reversal): the pass with radix R and
// running sub-transform length Ns reads x_r = buf[j + r*n/R], applies the
// twiddles w^(r*k), k = j mod Ns, w = exp(sign*2*pi*i/(Ns*R)), performs an
// R-point DFT and writes buf[(j-k)*R + k + r*Ns].  Every thread loads all
// its butterfly inputs before a barrier and stores after it, so one LDS
// buffer is reused in place.  Radices 2, 3, 4, 5 and 8 are supported; the
// pass schedule is packed four bits per pass into FftPlan::code.
//
// Twiddles come from a table tw[m] = exp(-2*pi*i*m/n) evaluated on the host;
// the inverse transform uses their conjugates.  Forward: sign = -1,
// unnormalised, i.e. the FFTW r2c/c2r convention of the reference
// (src/fluid/compute_potential.c:53-67).
#pragma once
#include <hip/hip_runtime.h>

namespace sns {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// multiply by sign*i
template <int SIGN>
__device__ __forceinline__ double2 cmul_si(double2 a) {
  return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// Quotients of small non-negative integers by a divisor fixed for a loop:
// for 0 <= x < 2^16 and 1 <= d < 2^16, floor(x * ceil(2^32/d) / 2^32) =
// floor(x/d) (the error term x*(ceil(2^32/d) - 2^32/d)/2^32 stays below
// 2^-16, less than the distance 1/d of x/d to the next integer).
struct FastDiv {
  unsigned m;
  int d;
  __host__ __device__ explicit FastDiv(int dv)
      : m(dv == 1 ? 0u : (unsigned)(0xFFFFFFFFull / (unsigned)dv + 1)), d(dv) {}
  __device__ __forceinline__ int div(int x) const {
    return d == 1 ? x : (int)__umulhi((unsigned)x, m);
  }
};

template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<2, SIGN> {
  __device__ __forceinline__ static void run(double2 *x) {
    const double2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <int SIGN>
struct Dft<4, SIGN> {
  __device__ __forceinline__ static void run(double2 *x) {
    const double2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
    const double2 t2 = cadd(x[1], x[3]), t3 = cmul_si<SIGN>(csub(x[1], x[3]));
    x[0] = cadd(t0, t2);
    x[2] = csub(t0, t2);
    x[1] = cadd(t1, t3);
    x[3] = csub(t1, t3);
  }
};

template <int SIGN>
struct Dft<8, SIGN> {
  __device__ __forceinline__ static void run(double2 *x) {
    constexpr double c = 0.70710678118654752440;  // sqrt(2)/2
    double2 e[4] = {x[0], x[2], x[4], x[6]};
    double2 o[4] = {x[1], x[3], x[5], x[7]};
    Dft<4, SIGN>::run(e);
    Dft<4, SIGN>::run(o);
    // o_k *= exp(sign*2*pi*i*k/8)
    const double2 o1 = make_double2(c * (o[1].x - SIGN * o[1].y), c * (o[1].y + SIGN * o[1].x));
    const double2 o2 = cmul_si<SIGN>(o[2]);
    const double2 o3 = make_double2(-c * (o[3].x + SIGN * o[3].y), c * (SIGN * o[3].x - o[3].y));
    x[0] = cadd(e[0], o[0]);
    x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], o1);
    x[5] = csub(e[1], o1);
    x[2] = cadd(e[2], o2);
    x[6] = csub(e[2], o2);
    x[3] = cadd(e[3], o3);
    x[7] = csub(e[3], o3);
  }
};

template <int SIGN>
struct Dft<3, SIGN> {
  __device__ __forceinline__ static void run(double2 *x) {
    constexpr double h = 0.86602540378443864676;  // sqrt(3)/2
    const double2 t1 = cadd(x[1], x[2]);
    const double2 t2 = csub(x[0], cscale(t1, 0.5));
    const double2 t3 = cscale(cmul_si<SIGN>(csub(x[1], x[2])), h);
    x[0] = cadd(x[0], t1);
    x[1] = cadd(t2, t3);
    x[2] = csub(t2, t3);
  }
};

template <int SIGN>
struct Dft<5, SIGN> {
  __device__ __forceinline__ static void run(double2 *x) {
    constexpr double c1 = 0.30901699437494742410;   // cos(2 pi/5)
    constexpr double c2 = -0.80901699437494742410;  // cos(4 pi/5)
    constexpr double s1 = 0.95105651629515357212;   // sin(2 pi/5)
    constexpr double s2 = 0.58778525229247312917;   // sin(4 pi/5)
    const double2 t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
    const double2 t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
    const double2 a1 = cadd(x[0], cadd(cscale(t1, c1), cscale(t2, c2)));
    const double2 a2 = cadd(x[0], cadd(cscale(t1, c2), cscale(t2, c1)));
    const double2 b1 = cmul_si<SIGN>(cadd(cscale(t3, s1), cscale(t4, s2)));
    const double2 b2 = cmul_si<SIGN>(csub(cscale(t3, s2), cscale(t4, s1)));
    x[0] = cadd(x[0], cadd(t1, t2));
    x[1] = cadd(a1, b1);
    x[4] = csub(a1, b1);
    x[2] = cadd(a2, b2);
    x[3] = csub(a2, b2);
  }
};

// Interleaved layout: element idx of sequence s lives at buf[idx*S + s], so
// consecutive threads touch consecutive LDS words in every pass.  Twiddles of
// the length-n transform come from a table of length tw_n (a multiple of n):
// w_n^m = tw[m * (tw_n / n)].
// Not inlined: with every radix inlined into the pass loop the register
// allocation of the kernels grows to the union of the paths (about 200
// VGPRs); as calls each pass keeps its own (under 80).
template <int NT, int EMAX, int R, int SIGN>
__device__ __noinline__ void fft_pass_il(double2 *buf, int S, int n, int Ns,
                                            const double2 *__restrict__ tw, int twmul) {
  constexpr int BMAX = (EMAX + R - 1) / R;
  const int nb = n / R;
  const int total = S * nb;
  const int tstep = (n / (Ns * R)) * twmul;
  const FastDiv fs(S), fns(Ns);
  double2 x[BMAX][R];
#pragma unroll
  for (int k = 0; k < BMAX; k++) {
    const int bidx = threadIdx.x + k * NT;
    if (bidx < total) {
      const int j = fs.div(bidx), s = bidx - j * S;
#pragma unroll
      for (int r = 0; r < R; r++) x[k][r] = buf[(j + r * nb) * S + s];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < BMAX; k++) {
    const int bidx = threadIdx.x + k * NT;
    if (bidx < total) {
      const int j = fs.div(bidx), s = bidx - j * S;
      const int kk = j - fns.div(j) * Ns;
      if (kk != 0) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          double2 w = tw[r * kk * tstep];
          if (SIGN > 0) w.y = -w.y;
          x[k][r] = cmul(x[k][r], w);
        }
      }
      Dft<R, SIGN>::run(x[k]);
      const int base = (j - kk) * R + kk;
#pragma unroll
      for (int r = 0; r < R; r++) buf[(base + r * Ns) * S + s] = x[k][r];
    }
  }
  __syncthreads();
}

// Power-of-two lengths L = 8..128 with the pass schedule fixed at compile
// time (radix 8 while at least 8 remains, then the remainder) and S = 2^lgS
// sequences: the same Stockham passes as fft_pass_il, inlined, with every
// index formed by shifts and masks.
constexpr int p2_radix(int L, int p) {
  int rem = L;
  for (int q = 0; q < p; q++) rem /= rem >= 8 ? 8 : rem;
  return rem >= 8 ? 8 : rem;
}
constexpr int p2_npass(int L) {
  int np = 0;
  for (int rem = L; rem > 1; rem /= rem >= 8 ? 8 : rem) np++;
  return np;
}

template <int NT, int EMAX, int L, int R, int NS, int SIGN>
__device__ __forceinline__ void fft_pass_p2(double2 *buf, int lgS,
                                            const double2 *__restrict__ tw, int twmul) {
  constexpr int BMAX = (EMAX + R - 1) / R;
  constexpr int nb = L / R;
  const int S = 1 << lgS;
  const int total = S * nb;
  const int tstep = (L / (NS * R)) * twmul;
  double2 x[BMAX][R];
#pragma unroll
  for (int k = 0; k < BMAX; k++) {
    const int bidx = threadIdx.x + k * NT;
    if (bidx < total) {
      const int j = bidx >> lgS, s = bidx & (S - 1);
#pragma unroll
      for (int r = 0; r < R; r++) x[k][r] = buf[((j + r * nb) << lgS) + s];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < BMAX; k++) {
    const int bidx = threadIdx.x + k * NT;
    if (bidx < total) {
      const int j = bidx >> lgS, s = bidx & (S - 1);
      const int kk = j & (NS - 1);
      if (NS > 1 && kk != 0) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          double2 w = tw[r * kk * tstep];
          if (SIGN > 0) w.y = -w.y;
          x[k][r] = cmul(x[k][r], w);
        }
      }
      Dft<R, SIGN>::run(x[k]);
      const int base = (j - kk) * R + kk;
#pragma unroll
      for (int r = 0; r < R; r++) buf[((base + r * NS) << lgS) + s] = x[k][r];
    }
  }
  __syncthreads();
}

template <int NT, int EMAX, int L, int SIGN, int P = 0, int NS = 1>
__device__ __forceinline__ void fft_p2(double2 *buf, int lgS, const double2 *__restrict__ tw,
                                       int twmul) {
  if constexpr (P < p2_npass(L)) {
    constexpr int R = p2_radix(L, P);
    fft_pass_p2<NT, EMAX, L, R, NS, SIGN>(buf, lgS, tw, twmul);
    fft_p2<NT, EMAX, L, SIGN, P + 1, NS * R>(buf, lgS, tw, twmul);
  }
}

template <int NT, int EMAX, int SIGN>
__device__ void fft_il(double2 *buf, int S, int n, int npass, unsigned long long code,
                       const double2 *__restrict__ tw, int twmul) {
  int Ns = 1;
  for (int p = 0; p < npass; p++) {
    const int R = (int)((code >> (4 * p)) & 15ull);
    switch (R) {
      case 2: fft_pass_il<NT, EMAX, 2, SIGN>(buf, S, n, Ns, tw, twmul); break;
      case 3: fft_pass_il<NT, EMAX, 3, SIGN>(buf, S, n, Ns, tw, twmul); break;
      case 4: fft_pass_il<NT, EMAX, 4, SIGN>(buf, S, n, Ns, tw, twmul); break;
      case 5: fft_pass_il<NT, EMAX, 5, SIGN>(buf, S, n, Ns, tw, twmul); break;
      default: fft_pass_il<NT, EMAX, 8, SIGN>(buf, S, n, Ns, tw, twmul); break;
    }
    Ns *= R;
  }
}

}  // namespace sns
