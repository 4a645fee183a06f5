// sns_diag.hip -- monitoring diagnostics on the device (gfx950).
//
// One pass over ux, uy and T evaluates the per-cell terms of the reference's
// monitors with its operations and operand order:
//   divergence (max |div|, sum div)           src/logging.c:36-80
//   momentum (sum ux dV, sum uy dV)           src/logging.c:82-122
//   quadratic energies of ux, uy, T           src/logging.c:124-173
//   Nusselt numbers: wall flux, energy injection, kinetic and thermal
//   dissipation                               src/logging.c:175-272,
//                                             src/analyses.c:8-121
// Each workgroup reduces its rows in a fixed order into one slot of a partial
// buffer and a single workgroup reduces the slots, again in a fixed order, so
// results are reproducible run to run; the sums differ from the sequential
// loops only by summation order, the maximum is exact.
#include "sns_launch.h"
#include "sns_fast.h"
#include "sns_diag.h"

namespace sns {

constexpr int kDiagThreads = 256;

// dissipation terms (src/analyses.c:8-121); k_u = sqrt(Pr)/sqrt(Ra),
// k_t = 1./sqrt(Pr)/sqrt(Ra)
__device__ __forceinline__ double dissip_x(const Geo &g, const double *um, const double *uc,
                                           const double *up, int i, double dy, double k_u) {
  // ux at face i of row j; um, uc, up: rows j-1, j, j+1 (element I-1)
  double s = 0.;
  if (i != 1) {
    const double d = ldr(uc, i - 1) - ldr(uc, i - 2);
    const double dd = d / DXF(g, i - 1);
    s += 0.5 / DXC(g, i) * d * dd;
  }
  if (i != g.itot + 1) {
    const double d = ldr(uc, i) - ldr(uc, i - 1);
    const double dd = d / DXF(g, i);
    s += 0.5 / DXC(g, i) * d * dd;
  }
  {
    const double d = ldr(uc, i - 1) - ldr(um, i - 1);
    const double dd = d / dy;
    s += 0.5 / dy * d * dd;
  }
  {
    const double d = ldr(up, i - 1) - ldr(uc, i - 1);
    const double dd = d / dy;
    s += 0.5 / dy * d * dd;
  }
  return k_u * s;
}

// cell-centred field (uy or T) at cell i of row j: rows j-1, j, j+1
__device__ __forceinline__ double dissip_c(const Geo &g, const double *fm, const double *fc,
                                           const double *fp, int i, double dy, double k) {
  double s = 0.;
  {
    const double c = i == 1 ? 1. : 0.5;
    const double d = ldr(fc, i) - ldr(fc, i - 1);
    const double dd = d / DXC(g, i);
    s += c / DXF(g, i) * d * dd;
  }
  {
    const double c = i == g.itot ? 1. : 0.5;
    const double d = ldr(fc, i + 1) - ldr(fc, i);
    const double dd = d / DXC(g, i + 1);
    s += c / DXF(g, i) * d * dd;
  }
  {
    const double d = ldr(fc, i) - ldr(fm, i);
    const double dd = d / dy;
    s += 0.5 / dy * d * dd;
  }
  {
    const double d = ldr(fp, i) - ldr(fc, i);
    const double dd = d / dy;
    s += 0.5 / dy * d * dd;
  }
  return k * s;
}

// Slots: 0 max|div| | 1 sum div | 2, 3 momentum | 4, 5, 6 energies of ux,
// uy, T | 7 wall flux | 8 injection | 9 kinetic dissipation | 10 thermal
// dissipation (before the normalisations of the final kernel)
__global__ void __launch_bounds__(kDiagThreads)
    k_diag(Geo g, DiagArgs a, double *__restrict__ part) {
  const double dy = g.dy;
  const double k_u = g.sqrtPr / g.sqrtRa, k_t = 1. / g.sqrtPr / g.sqrtRa;
  const double prra = g.sqrtPr * g.sqrtRa;
  const size_t wu = (size_t)g.pu, wc = (size_t)g.pc;
  double acc[kDiagSlots];
#pragma unroll
  for (int q = 0; q < kDiagSlots; q++) acc[q] = 0.;
  for (int j = blockIdx.x + 1; j <= g.jtot; j += gridDim.x) {
    const double *um = a.ux + (j - 1) * wu, *uc = a.ux + j * wu, *up = a.ux + (j + 1) * wu;
    const double *vm = a.uy + (j - 1) * wc, *vc = a.uy + j * wc, *vp = a.uy + (j + 1) * wc;
    const double *tm = a.temp ? a.temp + (j - 1) * wc : nullptr;
    const double *tc = a.temp ? a.temp + j * wc : nullptr;
    const double *tp = a.temp ? a.temp + (j + 1) * wc : nullptr;
    if (threadIdx.x == 0 && tc) {
      const double dtdx0 = (ldr(tc, 1) - ldr(tc, 0)) / DXC(g, 1);
      const double dtdx1 = (ldr(tc, g.itot + 1) - ldr(tc, g.itot)) / DXC(g, g.itot + 1);
      acc[7] -= 0.5 * (dtdx0 + dtdx1) * dy;
    }
    for (int i = threadIdx.x + 1; i <= g.itot + 1; i += kDiagThreads) {
      // faces i = 1..itot+1 (ux)
      {
        const double u = ldr(uc, i - 1);
        const double cell = DXC(g, i) * dy;
        acc[2] += u * (DXC(g, i) * dy);
        acc[4] += 0.5 * (u * u) * (DXC(g, i) * dy);
        if (tc) {
          const double integrand = u * 0.5 * (ldr(tc, i - 1) + ldr(tc, i));
          acc[8] += prra * integrand * cell;
        }
        acc[9] += prra * dissip_x(g, um, uc, up, i, dy, k_u) * cell;
      }
      if (i > g.itot) continue;
      // cells i = 1..itot
      const double v = ldr(vc, i);
      const double div = (ldr(uc, i) - ldr(uc, i - 1)) / DXF(g, i) + (ldr(vp, i) - v) / dy;
      acc[0] = fmax(acc[0], fabs(div));
      acc[1] += div;
      const double cell = DXF(g, i) * dy;
      acc[3] += v * (DXF(g, i) * dy);
      acc[5] += 0.5 * (v * v) * (DXF(g, i) * dy);
      acc[9] += prra * dissip_c(g, vm, vc, vp, i, dy, k_u) * cell;
      if (tc) {
        const double t = ldr(tc, i);
        acc[6] += 0.5 * (t * t) * (DXF(g, i) * dy);
        acc[10] += prra * dissip_c(g, tm, tc, tp, i, dy, k_t) * cell;
      }
    }
  }
  // workgroup reduction in a fixed order: within the wave, then over waves
  __shared__ double red[kDiagThreads / 64][kDiagSlots];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kDiagSlots; q++) {
    double v = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(v, off);
      v = q == 0 ? fmax(v, o) : v + o;
    }
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kDiagSlots) {
    const int q = threadIdx.x;
    double v = red[0][q];
    for (int w = 1; w < kDiagThreads / 64; w++) v = q == 0 ? fmax(v, red[w][q]) : v + red[w][q];
    part[(size_t)blockIdx.x * kDiagSlots + q] = v;
  }
}

// Reduces nb partial slots and applies the normalisations of
// src/logging.c:201,219-220,247-248,266.
__global__ void __launch_bounds__(64)
    k_diag_final(const double *__restrict__ part, int nb, double lx, double ly,
                 double *__restrict__ out) {
  const int q = threadIdx.x;
  if (q >= kDiagSlots) return;
  double v = part[q];
  for (int b = 1; b < nb; b++) v = q == 0 ? fmax(v, part[(size_t)b * kDiagSlots + q]) : v + part[(size_t)b * kDiagSlots + q];
  if (q == 7) v /= ly;
  if (q == 8 || q == 9) {
    v /= lx * ly;
    v += 1.;
  }
  if (q == 10) v /= lx * ly;
  out[q] = v;
}

int diag_partials(const Geo &g) { return g.jtot < 2048 ? g.jtot : 2048; }

void launch_diag(hipStream_t s, const Geo &g, const DiagArgs &a, double *part, double lx,
                 double ly, double *out) {
  const int nb = diag_partials(g);
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(kDiagThreads), 0, s, g, a, part);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, part, nb, lx, ly, out);
}

}  // namespace sns

namespace sns {

// ---------------------------------------------------------- statistics

// src/statistics/collect.c:10-75: ux1 += ux, ux2 += ux^2 over faces
// I = 1..itot+1, and the same for uy and T over cells I = 0..itot+1, rows
// J = 1..jtot.  gf: layout of the fields; the accumulators use the reference
// layouts (row pitches itot+1 and itot+2).  Element-wise, so the sums equal
// the sequential ones bit for bit.
__global__ void __launch_bounds__(256)
    k_stat_collect(Geo gf, StatArgs a) {
  const int j = blockIdx.y + 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..itot+1
  const int itot = gf.itot;
  if (i > itot + 1) return;
  const size_t ru = (size_t)j * (itot + 1), rc = (size_t)j * (itot + 2);
  if (i >= 1) {
    const double v = a.ux[(size_t)j * gf.pu + (i - 1)];
    a.ux1[ru + (i - 1)] += v;
    a.ux2[ru + (i - 1)] += v * v;
  }
  {
    const double v = a.uy[(size_t)j * gf.pc + i];
    a.uy1[rc + i] += v;
    a.uy2[rc + i] += v * v;
  }
  if (a.temp) {
    const double v = a.temp[(size_t)j * gf.pc + i];
    a.temp1[rc + i] += v;
    a.temp2[rc + i] += v * v;
  }
}

void launch_stat_collect(hipStream_t s, const Geo &gf, const StatArgs &a) {
  hipLaunchKernelGGL(k_stat_collect, dim3((gf.itot + 2 + 255) / 256, gf.jtot), dim3(256), 0, s,
                     gf, a);
}

}  // namespace sns
