// sns_launch.h -- host-side launchers of the gfx950 kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <atomic>
#include "sns_device.h"

namespace sns {

// Weights of the explicit / implicit split of the diffusive terms, exactly
// (1.-1.*implicit) and (+1.*implicit) as in the reference.
struct Treat {
  double w_ex_x, w_ex_y, w_im_x, w_im_y;
  int implicitx, implicity;
};

inline Treat make_treat(int implicitx, int implicity) {
  Treat t;
  t.implicitx = implicitx;
  t.implicity = implicity;
  t.w_ex_x = 1. - 1. * implicitx;
  t.w_ex_y = 1. - 1. * implicity;
  t.w_im_x = +1. * implicitx;
  t.w_im_y = +1. * implicity;
  return t;
}

struct RK {
  double alpha, beta, gamma;
};

// Where a kernel finds its argument block.  ArgOne: passed by value, one
// instance.  ArgBatch: a device array of blocks, one per z slice of the grid,
// for a batch of instances of the same shape (itot, jtot, treatment) that one
// launch advances together.  The array is read through the constant address
// space, as kernel arguments are: the loads are scalar and invariant, and the
// pointers in the blocks are known to address global memory, so the batched
// kernels compile to the same code and register use as the single ones.
template <class A>
struct ArgOne {
  A a;
  __device__ __forceinline__ A get() const { return a; }
};
template <class A>
struct ArgBatch {
  const __attribute__((address_space(4))) A *p;
  __device__ __forceinline__ A get() const {
    A r;
    __builtin_memcpy(&r, (const A *)(p + blockIdx.z), sizeof(A));
    return r;
  }
};
// Raise kernel f's dynamic LDS limit to `bytes` on the current device, once
// per device (the attribute belongs to the device's copy of the kernel);
// `done` holds one bit per device index.
inline void lds_limit_once(const void *f, size_t bytes, std::atomic<unsigned> &done) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;  // the launch reports the error
  const unsigned bit = 1u << (dev & 31);
  if (done.load(std::memory_order_acquire) & bit) return;
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
      hipSuccess)
    done.fetch_or(bit, std::memory_order_release);
}

template <class A>
inline ArgBatch<A> arg_batch(const A *dev) {
  ArgBatch<A> b;
  b.p = (const __attribute__((address_space(4))) A *)(uintptr_t)dev;
  return b;
}

// Crank-Nicolson solve in y of one field of one instance (after the
// predictor): the x-aligned increments qf (jtot rows of n) are transposed
// into qy (n rows of jtot), solved per row, transposed back, and
// out = in + increment with boundary values.  The per-function path uses
// the solve on qy alone (in, qf, out unused).
struct YArgs {
  Geo g;
  RK rk;
  const double *dtp;
  int kind;          // field kind (KIND_TEMP, KIND_UX, KIND_UY)
  int n;             // unknowns per x row: itot-1 (ux) or itot
  const double *in;  // substage-start field
  double *qf, *qy;
  double *out;
  // partitioned solve on the row-major increments (sns_ysolve.hip) when L > 0:
  // S chunks of L rows, the factor table of this substage and field, and
  // 3S rows of n: the interface values, then the first and the last interior
  // value of each chunk
  int L, S;
  const double *yfac;
  double *qz;
  // launch shapes of the partitioned solve (host side; 0: the rule):
  // columns per workgroup of the interface solve, 1 recompute / 2 store
  int cols, recompute;
};

// ---- boundary values (src/fluid/boundary_conditions.c, src/temperature/boundary_conditions.c)
void launch_bc_ux(hipStream_t s, const Geo &g, double *ux);
void launch_bc_uy(hipStream_t s, const Geo &g, double *uy);
void launch_bc_p(hipStream_t s, const Geo &g, double *p);
void launch_bc_temp(hipStream_t s, const Geo &g, double *temp);

// ---- temperature
void launch_force(hipStream_t s, const Geo &g, int with_thermal_forcing, const double *temp,
                  double *tempforcex);
void launch_src_temp(hipStream_t s, const Geo &g, const Treat &tr, const double *ux,
                     const double *uy, const double *temp, double *srca, double *srcg);
// ---- fluid
void launch_src_vel(hipStream_t s, const Geo &g, const Treat &tr, const double *ux,
                    const double *uy, const double *p, const double *tempforcex, double *srcuxa,
                    double *srcuxg, double *srcuya, double *srcuyg);

// ---- RK increment of one field (kind): either added to the field with its
// boundary values refreshed (q == nullptr) or stored x-aligned in q.
void launch_incr(hipStream_t s, const Geo &g, int kind, RK rk, const double *dtp,
                 const double *srca, const double *srcb, const double *srcg, double *field,
                 double *q);
// field += q (x-aligned), then boundary values
void launch_addq(hipStream_t s, const Geo &g, int kind, const double *q, double *field);

// ---- Crank-Nicolson in y on y-aligned rows (after a transpose): periodic
// solve of each of nrows rows of length jtot, in place.
void launch_cny_rows(hipStream_t s, const Geo &g, int kind, RK rk, const double *dtp, int nrows,
                     double *qy);

// ---- transposes: recv[i*g_jsize + j] = send[j*g_isize + i]
void launch_transpose(hipStream_t s, int g_isize, int g_jsize, size_t elemsize, const void *send,
                      void *recv);

// ---- correction and pressure (src/fluid/correct_velocity.c, update_pressure.c)
void launch_correct(hipStream_t s, const Geo &g, double gamma, const double *dtp,
                    const double *psi, double *ux, double *uy);
void launch_pressure(hipStream_t s, const Geo &g, const Treat &tr, double gamma,
                     const double *dtp, const double *psi, double *p);

// ---- advective time-step limit (src/param/decide_dt.c:25-45): writes the
// minimum of DXC(i)/max(|UX|,1e-8), dy/max(|UY|,1e-8) and 5e-2 (before the
// safety factor) as its bit pattern to *out (which must hold 0x7ff0... on entry)
void launch_dt_reduce(hipStream_t s, const Geo &g, const double *ux, const double *uy,
                      unsigned long long *out);
void launch_fill_u64(hipStream_t s, unsigned long long *p, unsigned long long v, int n);
// seeded initial temperature noise in the interior, then boundary values
void launch_init_temp(hipStream_t s, const Geo &g, double *temp, uint64_t seed, double amplitude);
void launch_copy(hipStream_t s, const void *src, void *dst, size_t n16);
// tiled copy, loads in {1, 2, 4, 8} 16-byte words per thread; false: bad shape
bool launch_copy_shaped(hipStream_t s, const void *src, void *dst, size_t n16, int loads,
                        bool nontemporal);

}  // namespace sns
