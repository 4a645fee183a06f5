// sns_capi.hip -- instance context and the extern "C" entry points of
// libsns_hip.so (declared in include/sns_hip.h).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

#include <map>
#include <mutex>

#include "sns_ctx.h"
#include "sns_diag.h"

#ifndef M_PI
#define M_PI 3.1415926535897932
#endif


static thread_local std::string g_last_error;

int set_err(int code, const char *what) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s (code %d%s%s)", what, code, code > 0 ? ": " : "",
           code > 0 ? hipGetErrorString((hipError_t)code) : "");
  g_last_error = buf;
  return code;
}


int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err((int)e, what);
  return 0;
}

// Runs `fn` (which enqueues kernels on ctx->stream) and, when timing is on
// and the name matches, brackets it with two events.
template <typename F>
static int run(snsg_ctx *ctx, const char *name, F fn) {
  const bool t = ctx->timing && (ctx->filter.empty() || ctx->filter == name);
  if (t) {
    while (ctx->ev_pool.size() < ctx->ev_used + 2) {
      hipEvent_t e;
      HIPCHECK(hipEventCreate(&e));
      ctx->ev_pool.push_back(e);
    }
    HIPCHECK(hipEventRecord(ctx->ev_pool[ctx->ev_used], ctx->stream));
  }
  fn();
  int rc = check_launch(name);
  if (rc) return rc;
  if (t) {
    HIPCHECK(hipEventRecord(ctx->ev_pool[ctx->ev_used + 1], ctx->stream));
    ctx->ev_used += 2;
  }
  return 0;
}

// ------------------------------------------------------------- creation

// Fused-state rows are padded to a multiple of 16 elements (128 bytes); cell
// arrays start kStateCOff elements into their allocation so that cell I = 1
// of every row is 128-byte aligned.
static int state_pitch(int n) { return (n + 15) / 16 * 16; }
static constexpr size_t kStateCOff = 15;

// A non-blocking stream per device for the library's own fills.  The
// library never uses the null stream: its implicit synchronisation is not
// ordered with non-blocking streams, and it is refused while another
// thread records a graph.
static hipError_t util_stream(hipStream_t *out) {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu);
  auto it = streams.find(dev);
  if (it == streams.end()) {
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    it = streams.emplace(dev, s).first;
  }
  *out = it->second;
  return hipSuccess;
}

// Zero-filled device allocation: the fill is complete on return, so the
// buffer is ready for any stream.
int ctx_alloc(void **p, size_t bytes, size_t *count) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess) return set_err((int)e, "hipMalloc");
  hipStream_t u = nullptr;
  e = util_stream(&u);
  if (e == hipSuccess) e = hipMemsetAsync(*p, 0, bytes ? bytes : 8, u);
  if (e == hipSuccess) e = hipStreamSynchronize(u);
  if (e != hipSuccess) return set_err((int)e, "hipMemset");
  if (count) *count += bytes ? bytes : 8;
  return 0;
}

extern "C" int snsg_device_count(int *count) {
  HIPCHECK(hipGetDeviceCount(count));
  return 0;
}

extern "C" const char *snsg_last_error(void) { return g_last_error.c_str(); }

__global__ void k_set_f64(double *p, double v) { *p = v; }

static int check_sizes(const snsg_config *cfg, Fft2Plan &pl, int n1) {
  const int itot = cfg->itot, jtot = cfg->jtot;
  if (itot < 4 || jtot < 4) return set_err(SNSG_E_SIZE, "snsg_create: itot and jtot must be >= 4");
  if (itot > 8192) return set_err(SNSG_E_SIZE, "snsg_create: itot > 8192 is not supported");
  if (jtot > 16384 || !fft2_plan(jtot, itot, pl, n1))
    return set_err(SNSG_E_SIZE,
                   "snsg_create: jtot must be <= 16384, not prime, with prime factors 2, 3 "
                   "and 5 only");
  return 0;
}

// Everything of an instance that depends on its parameters and grid (the
// buffers exist): the grid and the column tables, the constants of Geo, the
// Poisson eigenvalues, twiddles and Thomas factors, the diffusive time-step
// bounds; time and dt restart from 0.  Used by snsg_create and to reuse the
// buffers of a finished instance of the same size for a new one.
int ctx_configure(snsg_ctx *c, const snsg_config *cfg, const double *dxf, const double *dxc) {
  const int itot = cfg->itot, jtot = cfg->jtot;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->cfg = *cfg;
  HIPCHECK(hipMemcpyAsync(c->d_dxf, dxf, sizeof(double) * itot, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->d_dxc, dxc, sizeof(double) * (itot + 1), hipMemcpyHostToDevice, c->stream));
  // host-evaluated constants, with the reference's operations
  const double sqrtPr = sqrt(cfg->Pr), sqrtRa = sqrt(cfg->Ra);
  const double dy = cfg->ly / jtot;  // src/param/init.c:187
  Geo &g = c->geo;
  g.itot = itot;
  g.jtot = jtot;
  g.pu = itot + 1;
  g.pc = itot + 2;
  g.dy = dy;
  g.dxf = c->d_dxf;
  g.dxc = c->d_dxc;
  g.sqrtPr = sqrtPr;
  g.sqrtRa = sqrtRa;
  g.kappa_u = sqrtPr / sqrtRa;
  g.kappa_T = 1. / sqrtPr / sqrtRa;
  g.two_sqrtRa = 2. * sqrtRa;
  g.two_sqrtPr_sqrtRa = 2. * sqrtPr * sqrtRa;
  g.two_Re = 2. * (sqrtRa / sqrtPr);
  c->geo_st = g;
  c->geo_st.pu = state_pitch(itot + 1);
  c->geo_st.pc = state_pitch(itot + 2);
  c->treat = make_treat(cfg->implicitx, cfg->implicity);
  // eigenvalues of the periodic second difference (src/fluid/compute_potential.c:112-115)
  std::vector<double> eig(jtot / 2 + 1);
  for (int J = 0; J <= jtot / 2; J++)
    eig[J] = -4. / pow(dy, 2.) * pow(sin(M_PI * (J) / (1. * jtot)), 2.);
  std::vector<double2> tw(jtot);
  for (int m = 0; m < jtot; m++) {
    const double a = 2. * M_PI * (double)m / (double)jtot;
    tw[m] = make_double2(cos(a), -sin(a));
  }
  HIPCHECK(hipMemcpyAsync(c->d_eigen, eig.data(), sizeof(double) * eig.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->d_tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice, c->stream));
  // per-column tables (see sns_fast.h), each evaluated as in the reference
  const int nt = kTables;
  const size_t stride = (size_t)itot + 2;
  std::vector<double> h(nt * stride, 0.);
  auto DXF = [&](int i) { return dxf[i - 1]; };
  auto DXC = [&](int i) { return dxc[i - 1]; };
  double *t[nt];
  for (int k = 0; k < nt; k++) t[k] = h.data() + k * stride;
  for (int i = 1; i <= itot; i++) t[0][i - 1] = DXF(i);
  for (int i = 1; i <= itot + 1; i++) t[1][i - 1] = DXC(i);
  for (int i = 1; i <= itot; i++) t[2][i - 1] = 1. / DXF(i);
  for (int i = 1; i <= itot + 1; i++) t[3][i - 1] = 1. / DXC(i);
  for (int i = 1; i <= itot; i++) {
    t[4][i - 1] = DXC(i) / (2. * DXF(i));
    t[5][i - 1] = DXC(i + 1) / (2. * DXF(i));
    t[8][i - 1] = 1. / DXF(i) / DXC(i);
    t[9][i - 1] = 1. / DXF(i) / DXC(i + 1);
    t[12][i - 1] = 1. / DXC(i) / DXF(i);
    t[13][i - 1] = 1. / DXC(i + 1) / DXF(i);
  }
  for (int i = 2; i <= itot; i++) {
    t[6][i - 1] = DXF(i - 1) / (2. * DXC(i));
    t[7][i - 1] = DXF(i) / (2. * DXC(i));
    t[10][i - 1] = 1. / DXC(i) / DXF(i - 1);
    t[11][i - 1] = 1. / DXC(i) / DXF(i);
  }
  HIPCHECK(hipMemcpyAsync(c->d_tab, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  // paired copies: (DXC, 1/DXC), (DXF, 1/DXF), (cT_xm, cT_xp), (cU_xm, cU_xp)
  std::vector<double2> h2(kTables2 * stride);
  for (size_t k = 0; k < stride; k++) {
    h2[0 * stride + k] = make_double2(t[1][k], t[3][k]);
    h2[1 * stride + k] = make_double2(t[0][k], t[2][k]);
    h2[2 * stride + k] = make_double2(t[4][k], t[5][k]);
    h2[3 * stride + k] = make_double2(t[6][k], t[7][k]);
  }
  HIPCHECK(hipMemcpyAsync(c->d_tab2, h2.data(), sizeof(double2) * h2.size(), hipMemcpyHostToDevice, c->stream));
  Tables &tb = c->tab;
  tb.t = c->d_tab;
  tb.t2 = c->d_tab2;
  tb.stride = (long)stride;
  tb.dy = dy;
  tb.rdy = 1. / dy;
  // uniform interior: columns I = 2..itot-1 (entry I-1) share every value
  // the paired tables hold for them; cT_xm(I) for I >= 2 and cT_xp(I) for
  // I <= itot-1 (the same weight), cU_xm and cU_xp likewise
  {
    auto same = [&](const double *v, int from, int to, double &val) {
      val = v[from];
      for (int k = from; k <= to; k++)
        if (memcmp(&v[k], &val, sizeof val) != 0) return false;
      return true;
    };
    double cxp = 0., uxp = 0.;
    tb.uniform = itot >= 4 && same(t[1], 1, itot - 1, tb.u_dxc) &&
                 same(t[3], 1, itot - 1, tb.u_rdxc) && same(t[0], 0, itot - 1, tb.u_dxf) &&
                 same(t[2], 0, itot - 1, tb.u_rdxf) && same(t[4], 1, itot - 1, tb.u_cT) &&
                 same(t[5], 0, itot - 2, cxp) && memcmp(&cxp, &tb.u_cT, sizeof cxp) == 0 &&
                 same(t[6], 1, itot - 1, tb.u_cU) && same(t[7], 1, itot - 1, uxp) &&
                 memcmp(&uxp, &tb.u_cU, sizeof uxp) == 0;
    // the Poisson substitutions read l for rows 1..itot-1 and u for rows
    // 0..itot-2 (al of row 0 and be of the last row are 0)
    tb.plu_const = itot >= 3 && same(t[12], 1, itot - 1, tb.plc) && same(t[13], 0, itot - 2, tb.puc);
  }
  // every divisor (DXF(1..itot), DXC(1..itot+1), dy) a power of two with an
  // exact reciprocal: divisions become multiplications, bit for bit
  {
    auto p2 = [](double v) {
      int e;
      return v > 0. && frexp(v, &e) == 0.5 && (1. / v) * v == 1.;
    };
    bool ok = p2(dy);
    for (int i = 1; i <= itot && ok; i++) ok = p2(DXF(i));
    for (int i = 1; i <= itot + 1 && ok; i++) ok = p2(DXC(i));
    tb.pow2 = ok;
  }
  // fixed diffusive time-step bounds (src/param/decide_dt.c:48-66)
  const double diffusivity = fmax(sqrtPr / sqrtRa, 1. / sqrtPr / sqrtRa);
  double dx = 1.;
  for (int i = 2; i <= itot; i++) dx = fmin(dx, DXC(i));
  c->dt_dif_x = 0.25 / diffusivity * pow(dx, 2.) * 0.9;
  c->dt_dif_y = 0.25 / diffusivity * pow(dy, 2.) * 0.9;
  {
    std::vector<double> plu(2 * (size_t)poisson_factor_entries(itot));
    poisson_lu_coefs(itot, t[12], t[13], plu.data());
    HIPCHECK(hipMemcpyAsync(c->d_plu, plu.data(), sizeof(double) * plu.size(),
                            hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  launch_poisson_factor(c->stream, c->geo, c->tab, c->d_eigen, c->d_pfac);
  hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, c->stream, c->d_time, 0.);
  hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, c->stream, c->d_dt, 0.);
  int rc = check_launch("instance set-up");
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------------------------------- scratch of the no-allocation calls
//
// snsg_diagnostics_noalloc keeps the monitors' per-workgroup partial sums in
// this file-scope device array instead of instance memory.  Each GPU holds
// its own copy for the life of the process, at its own address: the host
// takes that address per device with hipGetSymbolAddress, once, while the
// device is current (at the first snsg_create on it), and uses only that
// address.  The array's name in host code is a host-side stand-in, never a
// device address.
constexpr int kScratchPartials = 2048;  // partial slots (diag_partials <= 2048)
constexpr size_t kScratchDoubles = (size_t)kScratchPartials * kDiagSlots;
static __device__ __align__(256) double scratch[kScratchDoubles];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_addr;  // device -> address of scratch

// With `device` current: look up (once per device) the device address of
// scratch.
static int scratch_setup(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_addr.count(device)) return 0;
  double *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    return set_err(e != hipSuccess ? (int)e : SNSG_E_ARG,
                   "hipGetSymbolAddress(scratch) during instance set-up");
  g_scratch_addr[device] = ptr;
  return 0;
}

static double *scratch_addr(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  return it == g_scratch_addr.end() ? nullptr : it->second;
}

extern "C" int snsg_create(int device, const snsg_config *cfg, const double *dxf,
                           const double *dxc, snsg_ctx **out) {
  if (!cfg || !dxf || !dxc || !out) return set_err(SNSG_E_ARG, "snsg_create: null argument");
  const int itot = cfg->itot, jtot = cfg->jtot;
  Fft2Plan pl;
  const snsg_launch lt = tuned_launch(itot, jtot);
  int rc = check_sizes(cfg, pl, lt.fft_n1);
  if (rc) return rc;
  HIPCHECK(hipSetDevice(device));
  if ((rc = scratch_setup(device))) return rc;
  snsg_ctx *c = new snsg_ctx();
  c->device = device;
  c->cfg = *cfg;
  c->plan = pl;
  c->launch = lt;
  auto fail = [&](int code) {
    snsg_destroy(c);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_err(SNSG_E_ARG, "hipStreamCreate"));
  const size_t stride = (size_t)itot + 2;
  const size_t nfac = poisson_factor_doubles(itot, jtot);
  struct {
    void **p;
    size_t bytes;
  } bufs[] = {
      {(void **)&c->d_dxf, sizeof(double) * itot},
      {(void **)&c->d_dxc, sizeof(double) * (itot + 1)},
      {(void **)&c->d_dt, sizeof(double)},
      {(void **)&c->d_time, sizeof(double)},
      {(void **)&c->d_eigen, sizeof(double) * (jtot / 2 + 1)},
      {(void **)&c->d_tw, sizeof(double2) * jtot},
      {(void **)&c->d_psihat, sizeof(double2) * (size_t)itot * (jtot / 2 + 1)},
      {(void **)&c->d_Y, sizeof(double2) * (size_t)pl.npair * jtot},
      {(void **)&c->d_q, sizeof(double) * (size_t)(itot + 1) * jtot},
      {(void **)&c->d_qy, sizeof(double) * (size_t)(itot + 1) * jtot},
      {(void **)&c->d_red, sizeof(unsigned long long)},
      {(void **)&c->d_tab, sizeof(double) * kTables * stride},
      {(void **)&c->d_tab2, sizeof(double2) * kTables2 * stride},
      {(void **)&c->d_pfac, sizeof(double) * nfac},
      {(void **)&c->d_plu, sizeof(double) * 2 * poisson_factor_entries(itot)},
  };
  for (auto &b : bufs)
    if ((rc = ctx_alloc(b.p, b.bytes, &c->mem))) return fail(rc);
  if (hipHostMalloc((void **)&c->h_red, sizeof(unsigned long long), hipHostMallocDefault) !=
      hipSuccess)
    return fail(set_err(SNSG_E_NOMEM, "hipHostMalloc"));
  if ((rc = ctx_configure(c, cfg, dxf, dxc))) return fail(rc);
  *out = c;
  return 0;
}

extern "C" int snsg_destroy(snsg_ctx *c) {
  if (!c) return 0;
  if (c->in_batch)
    return set_err(SNSG_E_ARG, "snsg_destroy: the instance is in a batch (destroy the batch "
                               "first: its graphs hold the instance's buffers)");
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *bufs[] = {c->d_dxf, c->d_dxc, c->d_dt, c->d_eigen, c->d_tw, c->d_psihat, c->d_Y, c->d_q,
                  c->d_qy, c->d_red, c->d_tab, c->d_time, c->d_qf[0], c->d_qf[1],
                  c->d_qf[2], c->d_lu, c->d_part, c->d_pfac, c->d_tab2, c->d_diag, c->d_plu,
                  c->d_yfac, c->d_qz};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (c->h_red) hipHostFree(c->h_red);
  if (c->h_diag) hipHostFree(c->h_diag);
  for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" int snsg_set_treatment(snsg_ctx *c, int implicitx, int implicity) {
  if (!c) return set_err(SNSG_E_ARG, "null ctx");
  if (c->in_batch)
    return set_err(SNSG_E_ARG, "snsg_set_treatment: the instance is in a batch (the treatment "
                               "is part of its shape class and recorded graphs)");
  c->cfg.implicitx = implicitx ? 1 : 0;
  c->cfg.implicity = implicity ? 1 : 0;
  c->treat = make_treat(c->cfg.implicitx, c->cfg.implicity);
  return 0;
}


// ------------------------------------------------------- launch shapes

static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, snsg_launch> g_tune;

snsg_launch tuned_launch(int itot, int jtot) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find({itot, jtot});
  return it == g_tune.end() ? snsg_launch{} : it->second;
}

LaunchKey launch_key(int itot, int jtot, const snsg_launch &req) {
  LaunchKey k{};
  Fft2Plan pl;
  k.n1 = fft2_plan(jtot, itot, pl, req.fft_n1) ? pl.n1 : 0;
  // chunk 1: the column solves on the transposed increments (also where
  // jtot has no chunk length the partitioned solve takes)
  k.chunk = req.ysolve_chunk ? req.ysolve_chunk : ysolve_chunk(jtot);
  if (k.chunk == 0) k.chunk = 1;
  const bool part = k.chunk > 1;
  // the width applies to interface systems of up to 256 rows
  const int SR = part ? jtot / k.chunk - 1 : 0;
  k.cols = part && SR <= 256 ? (req.ysolve_cols ? req.ysolve_cols : ysolve_cols_rule(SR)) : 0;
  k.recompute = part ? req.ysolve_recompute : 0;
  k.split = req.split_corr;
  return k;
}

LaunchKey launch_key(const snsg_ctx *c) {
  LaunchKey k = launch_key(c->geo.itot, c->geo.jtot, c->launch);
  k.n1 = c->plan.n1;
  return k;
}

extern "C" int snsg_launch_check(int itot, int jtot, const snsg_launch *l) {
  if (!l) return set_err(SNSG_E_ARG, "snsg_launch_check: null argument");
  if (itot < 4 || jtot < 4) return set_err(SNSG_E_SIZE, "snsg_launch_check: itot, jtot < 4");
  const int w = l->ysolve_cols;
  if (l->fft_n1 < 0 || l->ysolve_chunk < 0 || (w && w != 16 && w != 32 && w != 64) ||
      l->ysolve_recompute < 0 || l->ysolve_recompute > 2 || l->split_corr < 0 ||
      l->split_corr > 1)
    return set_err(SNSG_E_ARG, "snsg_launch_check: value out of range");
  Fft2Plan pl;
  if (l->fft_n1 && !fft2_plan(jtot, itot, pl, l->fft_n1))
    return set_err(SNSG_E_SIZE, "snsg_launch_check: fft_n1 does not give a four-step plan of jtot");
  if (l->ysolve_chunk > 1 && !ysolve_chunk_ok(jtot, l->ysolve_chunk))
    return set_err(SNSG_E_SIZE, "snsg_launch_check: ysolve_chunk must divide jtot into 4..513 "
                                "chunks of 4..32 rows");
  return 0;
}

extern "C" int snsg_set_launch(snsg_ctx *c, const snsg_launch *l) {
  if (!c) return set_err(SNSG_E_ARG, "null ctx");
  const int itot = c->geo.itot, jtot = c->geo.jtot;
  int rc = snsg_launch_check(itot, jtot, l);
  if (rc) return rc;
  if (c->in_batch)
    return set_err(SNSG_E_ARG, "snsg_set_launch: the instance is in a batch (its shape class "
                               "and recorded graphs depend on the shapes)");
  Fft2Plan pl;
  if (!fft2_plan(jtot, itot, pl, l->fft_n1))
    return set_err(SNSG_E_SIZE, "snsg_set_launch: no four-step plan");
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->plan = pl;
  c->launch = *l;  // the chunk length takes effect at the next fused step
  return 0;
}

extern "C" int snsg_get_launch(const snsg_ctx *c, snsg_launch *l) {
  if (!c || !l) return set_err(SNSG_E_ARG, "snsg_get_launch: null argument");
  const LaunchKey k = launch_key(c);
  l->fft_n1 = k.n1;
  l->ysolve_chunk = k.chunk;
  l->ysolve_cols = k.cols;
  l->ysolve_recompute = k.recompute;
  l->split_corr = k.split;
  return 0;
}

extern "C" int snsg_tuning_add(int itot, int jtot, const snsg_launch *l) {
  int rc = snsg_launch_check(itot, jtot, l);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[{itot, jtot}] = *l;
  return 0;
}

extern "C" int snsg_tuning_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  return 0;
}

extern "C" int snsg_set_dt(snsg_ctx *c, double dt) {
  if (!c) return set_err(SNSG_E_ARG, "null ctx");
  HIPCHECK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, c->stream, c->d_dt, dt);
  return check_launch("snsg_set_dt");
}

extern "C" int snsg_synchronize(snsg_ctx *c) {
  if (!c) return set_err(SNSG_E_ARG, "null ctx");
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// --------------------------------------------------------------- memory

extern "C" int snsg_malloc(snsg_ctx *c, size_t bytes, void **dptr) {
  if (!c || !dptr) return set_err(SNSG_E_ARG, "snsg_malloc: null argument");
  HIPCHECK(hipSetDevice(c->device));
  return ctx_alloc(dptr, bytes);
}

extern "C" int snsg_free(snsg_ctx *c, void *dptr) {
  if (!c) return set_err(SNSG_E_ARG, "null ctx");
  if (!dptr) return 0;
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipFree(dptr));
  return 0;
}

extern "C" int snsg_upload(snsg_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || !dst || !src) return set_err(SNSG_E_ARG, "snsg_upload: null argument");
  HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int snsg_download(snsg_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || !dst || !src) return set_err(SNSG_E_ARG, "snsg_download: null argument");
  HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int snsg_copy(snsg_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || !dst || !src) return set_err(SNSG_E_ARG, "snsg_copy: null argument");
  if (bytes % 16 || (uintptr_t)dst % 16 || (uintptr_t)src % 16)
    return set_err(SNSG_E_ARG, "snsg_copy: size and addresses must be multiples of 16 bytes");
  return run(c, "copy", [&] { launch_copy(c->stream, src, dst, bytes / 16); });
}

extern "C" int snsg_copy_shaped(snsg_ctx *c, void *dst, const void *src, size_t bytes,
                                int loads, int nontemporal) {
  if (!c || !dst || !src) return set_err(SNSG_E_ARG, "snsg_copy_shaped: null argument");
  if (bytes % 16 || (uintptr_t)dst % 16 || (uintptr_t)src % 16)
    return set_err(SNSG_E_ARG, "snsg_copy_shaped: size and addresses must be multiples of 16 "
                               "bytes");
  if (loads != 1 && loads != 2 && loads != 4 && loads != 8)
    return set_err(SNSG_E_ARG, "snsg_copy_shaped: loads must be 1, 2, 4 or 8");
  bool ok = true;
  int rc = run(c, "copy", [&] {
    ok = launch_copy_shaped(c->stream, src, dst, bytes / 16, loads, nontemporal != 0);
  });
  if (!ok) return set_err(SNSG_E_SIZE, "snsg_copy_shaped: too many blocks");
  return rc;
}

// ------------------------------------------------------ boundary values

#define CHECK_CTX(c, p) \
  if (!(c) || !(p)) return set_err(SNSG_E_ARG, __func__)

extern "C" int snsg_fluid_update_boundaries_ux(snsg_ctx *c, double *ux) {
  CHECK_CTX(c, ux);
  return run(c, "bc_ux", [&] { launch_bc_ux(c->stream, c->geo, ux); });
}
extern "C" int snsg_fluid_update_boundaries_uy(snsg_ctx *c, double *uy) {
  CHECK_CTX(c, uy);
  return run(c, "bc_uy", [&] { launch_bc_uy(c->stream, c->geo, uy); });
}
extern "C" int snsg_fluid_update_boundaries_p(snsg_ctx *c, double *p) {
  CHECK_CTX(c, p);
  return run(c, "bc_p", [&] { launch_bc_p(c->stream, c->geo, p); });
}
extern "C" int snsg_temperature_update_boundaries_temp(snsg_ctx *c, double *temp) {
  CHECK_CTX(c, temp);
  return run(c, "bc_temp", [&] { launch_bc_temp(c->stream, c->geo, temp); });
}

// ------------------------------------------------- predictor internals

RK rk_of(const snsg_ctx *c, int rkstep) {
  RK r;
  r.alpha = c->cfg.alpha[rkstep];
  r.beta = c->cfg.beta[rkstep];
  r.gamma = c->cfg.gamma[rkstep];
  return r;
}

// Apply the RK increment of one field with the configured diffusive
// treatment: src/temperature/update.c:89-231, src/fluid/update_velocity.c:194-488
// src/fluid/compute_potential.c:15-168 (three kernel groups: forward
// transform, x solves, inverse transform)
// g: the layout of ux, uy and psi (c->geo or c->geo_st)
PoisArgs pois_args(const snsg_ctx *c, const Geo &g, double gamma, const double *ux,
                          const double *uy, double *psi) {
  PoisArgs a;
  a.g = g;
  a.tb = c->tab;
  a.pl = c->plan;
  a.gamma = gamma;
  a.dtp = c->d_dt;
  a.ux = ux;
  a.uy = uy;
  a.tw = c->d_tw;
  a.fac = c->d_pfac;
  a.plu = c->d_plu;
  a.Y = c->d_Y;
  a.psihat = c->d_psihat;
  a.psi = psi;
  return a;
}

static int potential(snsg_ctx *c, const Geo &g, double gamma, const double *ux, const double *uy,
                     double *psi) {
  const PoisArgs a = pois_args(c, g, gamma, ux, uy, psi);
  int rc = run(c, "poisson_fwd", [&] { launch_poisson(c->stream, a, 0); });
  if (rc) return rc;
  if ((rc = run(c, "poisson_x", [&] { launch_poisson(c->stream, a, 1); }))) return rc;
  return run(c, "poisson_inv", [&] { launch_poisson(c->stream, a, 2); });
}

// Crank-Nicolson in x of one field on the reference layouts: the LU tables
// of all substages for the current dt (k_lu_factor), then the row solves of
// substage rkstep, into the field (qout == nullptr) or into qout.
static int cnx_solve(snsg_ctx *c, int kind, int rkstep, const double *srca, const double *srcb,
                     const double *srcg, double *field, double *qout) {
  const Geo &g = c->geo;
  int rc;
  if (!lu_row_threads(g.itot))
    return set_err(SNSG_E_SIZE, "implicit x needs itot <= 8192");
  if (!c->d_lu) {
    c->lu_ntm = lu_table_entries(g.itot);
    if ((rc = ctx_alloc((void **)&c->d_lu, sizeof(double) * lu_table_doubles(c->lu_ntm),
                        &c->mem)))
      return rc;
  }
  LuFactorArgs lf{};
  lf.g = g;
  lf.tb = c->tab;
  lf.dtp = c->d_dt;
  for (int k = 0; k < 3; k++) lf.gamma[k] = c->cfg.gamma[k];
  lf.out = c->d_lu;
  lf.ntm = c->lu_ntm;
  if ((rc = run(c, "lu_factor", [&] { launch_lu_factor(c->stream, lf); }))) return rc;
  const int f = kind == KIND_TEMP ? 0 : (kind == KIND_UX ? 1 : 2);
  CnxLuArgs a{};
  a.g = g;
  a.rk = rk_of(c, rkstep);
  a.dtp = c->d_dt;
  a.srca = srca;
  a.srcb = srcb;
  a.srcg = srcg;
  a.field = field;
  a.qout = qout;
  a.ntm = c->lu_ntm;
  a.lu = c->d_lu + (size_t)(rkstep * 3 * 3 + f * 3) * c->lu_ntm;
  a.lu_hdr = c->d_lu + 27 * (size_t)c->lu_ntm + (rkstep * 3 + f) * kLuHdr;
  return run(c, "cnx", [&] { launch_cnx_lu(c->stream, kind, a); });
}

static int ysolve_prepare(snsg_ctx *c);
static YFactorArgs yfactor_args(const snsg_ctx *c, const Geo &g);
static YArgs y_args(const snsg_ctx *c, const Geo &g, int rk, int f, const double *in, double *qf,
                    double *out);

static int update_field(snsg_ctx *c, int kind, int rkstep, const double *srca, const double *srcb,
                        const double *srcg, double *field) {
  const Geo &g = c->geo;
  const RK rk = rk_of(c, rkstep);
  const int n = kind == KIND_UX ? g.itot - 1 : g.itot;
  int rc;
  if (!c->treat.implicity) {
    if (!c->treat.implicitx)
      return run(c, "incr", [&] {
        launch_incr(c->stream, g, kind, rk, c->d_dt, srca, srcb, srcg, field, nullptr);
      });
    return cnx_solve(c, kind, rkstep, srca, srcb, srcg, field, nullptr);
  }
  if (c->treat.implicitx)
    rc = cnx_solve(c, kind, rkstep, srca, srcb, srcg, nullptr, c->d_q);
  else
    rc = run(c, "incr", [&] {
      launch_incr(c->stream, g, kind, rk, c->d_dt, srca, srcb, srcg, nullptr, c->d_q);
    });
  if (rc) return rc;
  if ((rc = ysolve_prepare(c))) return rc;
  if (c->ychunk) {
    // the partitioned solve of the fused step on the x-aligned increments;
    // its apply adds the solution into the field in place (each thread reads
    // the interior value it then overwrites)
    const int f = kind == KIND_TEMP ? 0 : (kind == KIND_UX ? 1 : 2);
    const YFactorArgs yf = yfactor_args(c, g);
    if ((rc = run(c, "yfactor", [&] { launch_yfactor(c->stream, yf); }))) return rc;
    const YArgs ya = y_args(c, g, rkstep, f, field, c->d_q, field);
    return run(c, "ysolve", [&] { launch_ypart_fields(c->stream, &ya, 1); });
  }
  if ((rc = run(c, "transpose", [&] { launch_transpose(c->stream, n, g.jtot, 8, c->d_q, c->d_qy); })))
    return rc;
  if ((rc = run(c, "cny", [&] { launch_cny_rows(c->stream, g, kind, rk, c->d_dt, n, c->d_qy); })))
    return rc;
  if ((rc = run(c, "transpose", [&] { launch_transpose(c->stream, g.jtot, n, 8, c->d_qy, c->d_q); })))
    return rc;
  return run(c, "addq", [&] { launch_addq(c->stream, g, kind, c->d_q, field); });
}

extern "C" int snsg_temperature_compute_force(snsg_ctx *c, const double *temp, double *tf) {
  CHECK_CTX(c, temp);
  if (!tf) return set_err(SNSG_E_ARG, __func__);
  return run(c, "force", [&] {
    launch_force(c->stream, c->geo, c->cfg.with_thermal_forcing, temp, tf);
  });
}

extern "C" int snsg_temperature_update_temp(snsg_ctx *c, int rkstep, const double *ux,
                                            const double *uy, double *temp, double *srca,
                                            const double *srcb, double *srcg) {
  CHECK_CTX(c, temp);
  if (!ux || !uy || !srca || !srcb || !srcg || rkstep < 0 || rkstep > 2)
    return set_err(SNSG_E_ARG, __func__);
  int rc = run(c, "src_temp", [&] {
    launch_src_temp(c->stream, c->geo, c->treat, ux, uy, temp, srca, srcg);
  });
  if (rc) return rc;
  return update_field(c, KIND_TEMP, rkstep, srca, srcb, srcg, temp);
}

extern "C" int snsg_fluid_update_velocity(snsg_ctx *c, int rkstep, double *ux, double *uy,
                                          const double *p, const double *tf, double *srcuxa,
                                          const double *srcuxb, double *srcuxg, double *srcuya,
                                          const double *srcuyb, double *srcuyg) {
  CHECK_CTX(c, ux);
  if (!uy || !p || !tf || !srcuxa || !srcuxb || !srcuxg || !srcuya || !srcuyb || !srcuyg ||
      rkstep < 0 || rkstep > 2)
    return set_err(SNSG_E_ARG, __func__);
  int rc = run(c, "src_vel", [&] {
    launch_src_vel(c->stream, c->geo, c->treat, ux, uy, p, tf, srcuxa, srcuxg, srcuya, srcuyg);
  });
  if (rc) return rc;
  if ((rc = update_field(c, KIND_UX, rkstep, srcuxa, srcuxb, srcuxg, ux))) return rc;
  return update_field(c, KIND_UY, rkstep, srcuya, srcuyb, srcuyg, uy);
}

extern "C" int snsg_fluid_compute_potential(snsg_ctx *c, int rkstep, const double *ux,
                                            const double *uy, double *psi) {
  CHECK_CTX(c, psi);
  if (!ux || !uy || rkstep < 0 || rkstep > 2) return set_err(SNSG_E_ARG, __func__);
  return potential(c, c->geo, c->cfg.gamma[rkstep], ux, uy, psi);
}

extern "C" int snsg_fluid_correct_velocity(snsg_ctx *c, int rkstep, const double *psi, double *ux,
                                           double *uy) {
  CHECK_CTX(c, psi);
  if (!ux || !uy || rkstep < 0 || rkstep > 2) return set_err(SNSG_E_ARG, __func__);
  return run(c, "correct", [&] {
    launch_correct(c->stream, c->geo, c->cfg.gamma[rkstep], c->d_dt, psi, ux, uy);
  });
}

extern "C" int snsg_fluid_update_pressure(snsg_ctx *c, int rkstep, const double *psi, double *p) {
  CHECK_CTX(c, psi);
  if (!p || rkstep < 0 || rkstep > 2) return set_err(SNSG_E_ARG, __func__);
  return run(c, "pressure", [&] {
    launch_pressure(c->stream, c->geo, c->treat, c->cfg.gamma[rkstep], c->d_dt, psi, p);
  });
}

extern "C" int snsg_param_decide_dt(snsg_ctx *c, const double *ux, const double *uy,
                                    double *dt_adv) {
  CHECK_CTX(c, ux);
  if (!uy || !dt_adv) return set_err(SNSG_E_ARG, __func__);
  const unsigned long long inf = 0x7ff0000000000000ull;
  int rc = run(c, "dt_reduce", [&] {
    launch_fill_u64(c->stream, c->d_red, inf, 1);
    launch_dt_reduce(c->stream, c->geo, ux, uy, c->d_red);
  });
  if (rc) return rc;
  HIPCHECK(hipMemcpyAsync(c->h_red, c->d_red, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  double v;
  memcpy(&v, c->h_red, sizeof v);
  *dt_adv = v;
  return 0;
}

extern "C" int snsg_transpose(snsg_ctx *c, int g_isize, int g_jsize, size_t elemsize,
                              const void *send, void *recv) {
  if (!c || !send || !recv || g_isize < 0 || g_jsize < 0 ||
      (elemsize != 4 && elemsize != 8 && elemsize != 16))
    return set_err(SNSG_E_ARG, __func__);
  if (g_isize == 0 || g_jsize == 0) return 0;
  return run(c, "transpose", [&] {
    launch_transpose(c->stream, g_isize, g_jsize, elemsize, send, recv);
  });
}

// ---------------------------------------------------------- fused step

static size_t bytes_ux(const snsg_ctx *c) {
  return sizeof(double) * (size_t)c->geo_st.pu * (c->geo_st.jtot + 2);
}
static size_t bytes_c(const snsg_ctx *c) {
  return sizeof(double) * ((size_t)c->geo_st.pc * (c->geo_st.jtot + 2) + 16);
}
static int alloc_c(snsg_ctx *c, double **p) {
  int rc = ctx_alloc((void **)p, bytes_c(c), &c->mem);
  if (!rc) *p += kStateCOff;
  return rc;
}

extern "C" int snsg_state_alloc(snsg_ctx *c, snsg_state *st) {
  if (!c || !st) return set_err(SNSG_E_ARG, __func__);
  HIPCHECK(hipSetDevice(c->device));
  memset(st, 0, sizeof *st);
  const int itot = c->geo.itot, jtot = c->geo.jtot;
  st->ld_ux = c->geo_st.pu;
  st->ld_c = c->geo_st.pc;
  int rc;
  if ((rc = ctx_alloc((void **)&st->ux, bytes_ux(c), &c->mem))) return rc;
  if ((rc = alloc_c(c, &st->uy))) return rc;
  if ((rc = alloc_c(c, &st->p))) return rc;
  if ((rc = alloc_c(c, &st->psi))) return rc;
  if ((rc = alloc_c(c, &st->temp))) return rc;
  if ((rc = alloc_c(c, &st->temp_alt))) return rc;
  if ((rc = ctx_alloc((void **)&st->ux_star, bytes_ux(c), &c->mem))) return rc;
  if ((rc = alloc_c(c, &st->uy_star))) return rc;
  if ((rc = alloc_c(c, &st->p_alt))) return rc;
  if ((rc = ctx_alloc((void **)&st->srcuxa, sizeof(double) * (size_t)(itot - 1) * jtot, &c->mem))) return rc;
  if ((rc = ctx_alloc((void **)&st->srcuya, sizeof(double) * (size_t)itot * jtot, &c->mem))) return rc;
  if ((rc = ctx_alloc((void **)&st->srctempa, sizeof(double) * (size_t)itot * jtot, &c->mem))) return rc;
  return 0;
}

extern "C" int snsg_state_free(snsg_ctx *c, snsg_state *st) {
  if (!c || !st) return set_err(SNSG_E_ARG, __func__);
  if (c->in_batch)
    return set_err(SNSG_E_ARG, "snsg_state_free: the instance is in a batch (destroy the batch "
                               "first: its graphs hold the state's fields)");
  HIPCHECK(hipStreamSynchronize(c->stream));
  double *plain[] = {st->ux, st->ux_star, st->srcuxa, st->srcuya, st->srctempa};
  double *cell[] = {st->uy, st->p, st->psi, st->temp, st->temp_alt, st->uy_star, st->p_alt};
  for (double *p : plain)
    if (p) hipFree(p);
  for (double *p : cell)
    if (p) hipFree(p - kStateCOff);
  if (st->ux) {
    const size_t sx = sizeof(double) * (size_t)(c->geo.itot - 1) * c->geo.jtot;
    const size_t sy = sizeof(double) * (size_t)c->geo.itot * c->geo.jtot;
    c->mem -= 2 * bytes_ux(c) + 7 * bytes_c(c) + sx + 2 * sy;
  }
  memset(st, 0, sizeof *st);
  return 0;
}

// Zeroes every array of a state (on the instance's stream), as a new one is.
int state_reset(snsg_ctx *c, snsg_state *st) {
  const size_t bu = bytes_ux(c), bc = bytes_c(c);
  const size_t sx = sizeof(double) * (size_t)(c->geo.itot - 1) * c->geo.jtot;
  const size_t sy = sizeof(double) * (size_t)c->geo.itot * c->geo.jtot;
  struct {
    void *p;
    size_t bytes;
  } a[] = {{st->ux, bu},           {st->ux_star, bu},       {st->srcuxa, sx},
           {st->srcuya, sy},       {st->srctempa, sy},      {st->uy - kStateCOff, bc},
           {st->p - kStateCOff, bc}, {st->psi - kStateCOff, bc}, {st->temp - kStateCOff, bc},
           {st->temp_alt - kStateCOff, bc}, {st->uy_star - kStateCOff, bc},
           {st->p_alt - kStateCOff, bc}};
  for (auto &x : a) HIPCHECK(hipMemsetAsync(x.p, 0, x.bytes, c->stream));
  return 0;
}

extern "C" int snsg_uniform_grid(const snsg_ctx *c, int *uniform) {
  if (!c || !uniform) return set_err(SNSG_E_ARG, __func__);
  *uniform = c->tab.uniform;
  return 0;
}

extern "C" int snsg_memory(const snsg_ctx *c, size_t *bytes) {
  if (!c || !bytes) return set_err(SNSG_E_ARG, __func__);
  *bytes = c->mem;
  return 0;
}

extern "C" int snsg_state_init_temperature(snsg_ctx *c, snsg_state *st, uint64_t seed,
                                           double amplitude) {
  if (!c || !st || !st->temp) return set_err(SNSG_E_ARG, __func__);
  HIPCHECK(hipSetDevice(c->device));
  launch_init_temp(c->stream, c->geo_st, st->temp, seed, amplitude);
  return check_launch(__func__);
}

// Device bytes of an instance of cfg's size and treatment with its state and
// the scratch of the fused step: the allocations of snsg_create,
// snsg_state_alloc and the fused step's first use, each as ctx_alloc counts it.
extern "C" int snsg_instance_bytes(const snsg_config *cfg, size_t *bytes) {
  if (!cfg || !bytes) return set_err(SNSG_E_ARG, __func__);
  Fft2Plan pl;
  int rc = check_sizes(cfg, pl, tuned_launch(cfg->itot, cfg->jtot).fft_n1);
  if (rc) return rc;
  const size_t itot = cfg->itot, jtot = cfg->jtot, d = sizeof(double), z = sizeof(double2);
  const size_t stride = itot + 2;
  const size_t pu = (size_t)state_pitch((int)itot + 1), pc = (size_t)state_pitch((int)itot + 2);
  size_t n = 0;
  // snsg_create
  n += d * itot + d * (itot + 1) + d + d + d * (jtot / 2 + 1) + z * jtot;
  n += z * itot * (jtot / 2 + 1) + z * (size_t)pl.npair * jtot;
  n += 2 * d * (itot + 1) * jtot + sizeof(unsigned long long);
  n += d * kTables * stride + z * kTables2 * stride;
  n += d * poisson_factor_doubles((int)itot, (int)jtot) + d * 2 * (size_t)poisson_factor_entries((int)itot);
  // snsg_state_alloc
  n += 2 * d * pu * (jtot + 2) + 7 * d * (pc * (jtot + 2) + 16);
  n += d * (itot - 1) * jtot + 2 * d * itot * jtot;
  // first fused step
  if (cfg->implicity) {
    n += d * (2 * itot + itot - 1) * jtot;
    const int L = launch_key((int)itot, (int)jtot, tuned_launch((int)itot, (int)jtot)).chunk;
    if (L > 1) n += d * 9 * ysolve_table_doubles(L) + 9 * d * (jtot / L) * itot;
  }
  if (cfg->implicitx) n += d * lu_table_doubles(lu_table_entries((int)itot));
  Geo g{};
  g.itot = (int)itot;
  g.jtot = (int)jtot;
  n += d * (size_t)corr_partials(g);
  *bytes = n;
  return 0;
}

// device pointer, row pitch, row length and row count of a state field
static int state_field(const snsg_ctx *c, const snsg_state *st, int field, double **dev,
                       size_t *pitch, size_t *width, size_t *rows) {
  const int itot = c->geo.itot, jtot = c->geo.jtot;
  *rows = (size_t)jtot + 2;
  switch (field) {
    case SNSG_UX: *dev = st->ux; break;
    case SNSG_UX_STAR: *dev = st->ux_star; break;
    case SNSG_UY: *dev = st->uy; break;
    case SNSG_UY_STAR: *dev = st->uy_star; break;
    case SNSG_P: *dev = st->p; break;
    case SNSG_PSI: *dev = st->psi; break;
    case SNSG_TEMP: *dev = st->temp; break;
    case SNSG_SRCUXA: *dev = st->srcuxa; *pitch = *width = itot - 1; *rows = jtot; return 0;
    case SNSG_SRCUYA: *dev = st->srcuya; *pitch = *width = itot; *rows = jtot; return 0;
    case SNSG_SRCTEMPA: *dev = st->srctempa; *pitch = *width = itot; *rows = jtot; return 0;
    default: return set_err(SNSG_E_ARG, "state field id");
  }
  const bool ux = field == SNSG_UX || field == SNSG_UX_STAR;
  *pitch = ux ? c->geo_st.pu : c->geo_st.pc;
  *width = ux ? itot + 1 : itot + 2;
  return 0;
}

extern "C" int snsg_state_upload(snsg_ctx *c, snsg_state *st, int field, const double *host) {
  if (!c || !st || !host) return set_err(SNSG_E_ARG, __func__);
  double *dev;
  size_t pitch, width, rows;
  int rc = state_field(c, st, field, &dev, &pitch, &width, &rows);
  if (rc) return rc;
  if (!dev) return set_err(SNSG_E_ARG, "snsg_state_upload: state not allocated");
  HIPCHECK(hipMemcpy2DAsync(dev, pitch * sizeof(double), host, width * sizeof(double),
                            width * sizeof(double), rows, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int snsg_state_download(snsg_ctx *c, const snsg_state *st, int field, double *host) {
  if (!c || !st || !host) return set_err(SNSG_E_ARG, __func__);
  double *dev;
  size_t pitch, width, rows;
  int rc = state_field(c, st, field, &dev, &pitch, &width, &rows);
  if (rc) return rc;
  if (!dev) return set_err(SNSG_E_ARG, "snsg_state_download: state not allocated");
  HIPCHECK(hipMemcpy2DAsync(host, width * sizeof(double), dev, pitch * sizeof(double),
                            width * sizeof(double), rows, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// monitors of src/logging.c on the device (see sns_diag.hip); blocking
static int diagnostics(snsg_ctx *c, const Geo &g, const double *ux, const double *uy,
                       const double *temp, double *out) {
  int rc;
  const size_t np = (size_t)diag_partials(g) * kDiagSlots;
  if (!c->d_diag) {
    if ((rc = ctx_alloc((void **)&c->d_diag, sizeof(double) * (np + kDiagSlots), &c->mem))) return rc;
    HIPCHECK(hipHostMalloc((void **)&c->h_diag, sizeof(double) * kDiagSlots, hipHostMallocDefault));
  }
  DiagArgs a{ux, uy, c->cfg.with_temperature ? temp : nullptr};
  rc = run(c, "diag",
           [&] { launch_diag(c->stream, g, a, c->d_diag, 1., c->cfg.ly, c->d_diag + np); });
  if (rc) return rc;
  HIPCHECK(hipMemcpyAsync(c->h_diag, c->d_diag + np, sizeof(double) * kDiagSlots,
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  memcpy(out, c->h_diag, sizeof(double) * kDiagSlots);
  return 0;
}

extern "C" int snsg_diagnostics(snsg_ctx *c, const double *ux, const double *uy,
                                const double *temp, double *out) {
  if (!c || !ux || !uy || !out || (c->cfg.with_temperature && !temp))
    return set_err(SNSG_E_ARG, __func__);
  return diagnostics(c, c->geo, ux, uy, temp, out);
}

extern "C" int snsg_state_diagnostics(snsg_ctx *c, const snsg_state *st, double *out) {
  if (!c || !st || !st->ux || !out) return set_err(SNSG_E_ARG, __func__);
  return diagnostics(c, c->geo_st, st->ux, st->uy, st->temp, out);
}

// snsg_diagnostics with no allocation and no host synchronisation: the
// same kernels and partial count (so the same bits), the partials in the
// device's scratch, the 11 results written to the caller's device array.
// Nothing in scratch needs to start at zero: every partial slot the final
// reduction reads is written by the first kernel of the same call.
extern "C" int snsg_diagnostics_noalloc(snsg_ctx *c, const double *ux, const double *uy,
                                        const double *temp, double *out_dev) {
  if (!c || !ux || !uy || !out_dev || (c->cfg.with_temperature && !temp))
    return set_err(SNSG_E_ARG, __func__);
  const int nb = diag_partials(c->geo);
  if (nb > kScratchPartials) {
    char msg[160];
    snprintf(msg, sizeof msg,
             "snsg_diagnostics_noalloc: %d partial slots exceed the scratch's %d", nb,
             kScratchPartials);
    return set_err(SNSG_E_SIZE, msg);
  }
  double *part = scratch_addr(c->device);
  if (!part)
    return set_err(SNSG_E_ARG, "snsg_diagnostics_noalloc: no scratch address for the "
                               "instance's device");
  const DiagArgs a{ux, uy, c->cfg.with_temperature ? temp : nullptr};
  launch_diag(c->stream, c->geo, a, part, 1., c->cfg.ly, out_dev);
  return check_launch(__func__);
}

extern "C" int snsg_statistics_collect(snsg_ctx *c, const double *ux, const double *uy,
                                       const double *temp, double *ux1, double *ux2, double *uy1,
                                       double *uy2, double *temp1, double *temp2) {
  if (!c || !ux || !uy || !ux1 || !ux2 || !uy1 || !uy2) return set_err(SNSG_E_ARG, __func__);
  const bool t = c->cfg.with_temperature != 0;
  if (t && (!temp || !temp1 || !temp2)) return set_err(SNSG_E_ARG, __func__);
  StatArgs a{ux, uy, t ? temp : nullptr, ux1, ux2, uy1, uy2, t ? temp1 : nullptr,
             t ? temp2 : nullptr};
  return run(c, "stat", [&] { launch_stat_collect(c->stream, c->geo, a); });
}

extern "C" int snsg_state_statistics_collect(snsg_ctx *c, const snsg_state *st, double *ux1,
                                             double *ux2, double *uy1, double *uy2,
                                             double *temp1, double *temp2) {
  if (!c || !st || !st->ux || !ux1 || !ux2 || !uy1 || !uy2) return set_err(SNSG_E_ARG, __func__);
  const bool t = c->cfg.with_temperature != 0;
  if (t && (!temp1 || !temp2)) return set_err(SNSG_E_ARG, __func__);
  StatArgs a{st->ux, st->uy, t ? st->temp : nullptr, ux1, ux2, uy1, uy2, t ? temp1 : nullptr,
             t ? temp2 : nullptr};
  return run(c, "stat", [&] { launch_stat_collect(c->stream, c->geo_st, a); });
}

extern "C" int snsg_state_decide_dt(snsg_ctx *c, const snsg_state *st, double *dt_adv) {
  if (!c || !st || !st->ux || !dt_adv) return set_err(SNSG_E_ARG, __func__);
  int rc = run(c, "dt_reduce", [&] {
    launch_fill_u64(c->stream, c->d_red, 0x7ff0000000000000ull, 1);
    launch_dt_reduce(c->stream, c->geo_st, st->ux, st->uy, c->d_red);
  });
  if (rc) return rc;
  HIPCHECK(hipMemcpyAsync(c->h_red, c->d_red, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  memcpy(dt_adv, c->h_red, sizeof *dt_adv);
  return 0;
}

const int kYKinds[3] = {KIND_TEMP, KIND_UX, KIND_UY};

// Scratch of the fused step, allocated on first use
// Factor tables and interface rows of the partitioned y solve, sized by the
// chunk length (the launch shapes may have changed it); c->ychunk 0 means
// the transposed column solves.
static int ysolve_prepare(snsg_ctx *c) {
  const int itot = c->geo.itot, jtot = c->geo.jtot;
  int rc;
  auto ybytes = [&](int L, size_t &nt, size_t &nz) {
    nt = sizeof(double) * 9 * ysolve_table_doubles(L);
    // interface values and the first and last interior values of each
    // chunk, for each of the three fields (their solves run concurrently)
    nz = sizeof(double) * 9 * (size_t)(jtot / L) * itot;
  };
  int L = c->treat.implicity ? launch_key(c).chunk : c->ychunk;
  if (L == 1) L = 0;  // transposed solves
  if (L != c->ychunk || (L && !c->d_qz)) {
    if (c->d_qz) {
      HIPCHECK(hipStreamSynchronize(c->stream));
      size_t nt, nz;
      ybytes(c->ychunk, nt, nz);
      (void)hipFree(c->d_yfac);
      (void)hipFree(c->d_qz);
      c->d_yfac = c->d_qz = nullptr;
      c->mem -= nt + nz;
    }
    c->ychunk = L;
    if (L) {
      size_t nt, nz;
      ybytes(L, nt, nz);
      if ((rc = ctx_alloc((void **)&c->d_yfac, nt, &c->mem))) return rc;
      if ((rc = ctx_alloc((void **)&c->d_qz, nz, &c->mem))) return rc;
    }
  }
  return 0;
}

// Factors of the partitioned y solves of all substages and fields (g: the
// layout the solves run on)
static YFactorArgs yfactor_args(const snsg_ctx *c, const Geo &g) {
  const int yL = c->treat.implicity ? c->ychunk : 0;
  YFactorArgs yf{};
  yf.g = g;
  yf.dtp = c->d_dt;
  for (int k = 0; k < 3; k++) yf.gamma[k] = c->cfg.gamma[k];
  yf.out = c->d_yfac;
  yf.L = yL;
  yf.S = yL ? g.jtot / yL : 0;
  yf.stride = yL ? (long)ysolve_table_doubles(yL) : 0;
  return yf;
}

// The y solve of field f (order T, ux, uy) in substage rk: increments qf
// (x-aligned), out = in + solution with boundary values
static YArgs y_args(const snsg_ctx *c, const Geo &g, int rk, int f, const double *in, double *qf,
                    double *out) {
  const int yL = c->treat.implicity ? c->ychunk : 0;
  YArgs y{};
  y.g = g;
  y.rk = rk_of(c, rk);
  y.dtp = c->d_dt;
  y.kind = kYKinds[f];
  y.n = kYKinds[f] == KIND_UX ? g.itot - 1 : g.itot;
  y.in = in;
  y.qf = qf;
  y.qy = c->d_qy;
  y.out = out;
  y.L = yL;
  y.S = yL ? g.jtot / yL : 0;
  y.yfac = yL ? c->d_yfac + (size_t)(rk * 3 + f) * ysolve_table_doubles(yL) : nullptr;
  y.qz = yL ? c->d_qz + (size_t)f * 3 * (g.jtot / yL) * g.itot : nullptr;
  y.cols = c->launch.ysolve_cols;
  y.recompute = c->launch.ysolve_recompute;
  return y;
}

int step_prepare(snsg_ctx *c) {
  const Geo &g = c->geo_st;
  const int itot = g.itot, jtot = g.jtot;
  int rc;
  if (c->treat.implicitx && !pred_impx_supported(itot))
    return set_err(SNSG_E_SIZE, "fused step: implicit x needs itot <= 8192");
  if (c->treat.implicity && !c->d_qf[0]) {
    const size_t nb[3] = {(size_t)itot * jtot, (size_t)(itot - 1) * jtot, (size_t)itot * jtot};
    for (int k = 0; k < 3; k++)
      if ((rc = ctx_alloc((void **)&c->d_qf[k], sizeof(double) * nb[k], &c->mem))) return rc;
  }
  if ((rc = ysolve_prepare(c))) return rc;
  if (c->treat.implicitx && !c->d_lu) {
    c->lu_ntm = lu_table_entries(itot);
    if ((rc = ctx_alloc((void **)&c->d_lu, sizeof(double) * lu_table_doubles(c->lu_ntm), &c->mem))) return rc;
  }
  if (!c->d_part) {
    c->nparts = corr_partials(g);
    if ((rc = ctx_alloc((void **)&c->d_part, sizeof(double) * (size_t)c->nparts, &c->mem))) return rc;
  }
  return 0;
}

void step_args(const snsg_ctx *c, const snsg_state *st, bool reduce, StepArgs &sa) {
  const Geo &g = c->geo_st;
  const bool impx = c->treat.implicitx, impy = c->treat.implicity;
  const bool with_temp = c->cfg.with_temperature != 0;
  sa = StepArgs{};
  // the CN-x operators of all three substages depend on dt only
  sa.lu.g = g;
  sa.lu.tb = c->tab;
  sa.lu.dtp = c->d_dt;
  for (int k = 0; k < 3; k++) sa.lu.gamma[k] = c->cfg.gamma[k];
  sa.lu.out = c->d_lu;
  sa.lu.ntm = c->lu_ntm;
  sa.dt.part = c->d_part;
  sa.dt.n = c->nparts;
  sa.dt.dt = c->d_dt;
  sa.dt.time = c->d_time;
  sa.dt.dt_dif_x = c->dt_dif_x;
  sa.dt.dt_dif_y = c->dt_dif_y;
  sa.dt.implicitx = impx;
  sa.dt.implicity = impy;
  sa.bound.g = g;
  sa.bound.tb = c->tab;
  sa.bound.tr = c->treat;
  sa.bound.ux = st->ux;
  sa.bound.uy = st->uy;
  sa.bound.red = c->d_part;
  sa.yf = yfactor_args(c, g);
  double *temp = st->temp, *temp_alt = st->temp_alt;
  // Folded corrections (implicit x, explicit y): substage 1 predicts from
  // (ux, uy, p) into (ux*, uy*) and stops at psi; substage 2's predictor
  // corrects ux*, uy* and p with that psi as it loads them, predicts into
  // (ux, uy) and stores the corrected p in p_alt; substage 3 does the same
  // from (ux, uy, p_alt) into (ux*, uy*) and p, and its correction kernel
  // writes ux, uy, p as in the split sequence.
  const bool fold = impx && !impy && !c->launch.split_corr;
  double *in_u = st->ux, *in_v = st->uy, *out_u = st->ux_star, *out_v = st->uy_star;
  double *in_p = st->p;
  for (int rk = 0; rk < 3; rk++) {
    const bool folded = fold && rk > 0;  // applies the correction of substage rk-1
    PredArgs &pa = sa.pred[rk];
    pa.g = g;
    pa.tb = c->tab;
    pa.tr = c->treat;
    pa.ux = in_u; pa.uy = in_v; pa.p = in_p; pa.temp = temp;
    pa.hist_ux = st->srcuxa; pa.hist_uy = st->srcuya; pa.hist_t = st->srctempa;
    pa.out_u = out_u; pa.out_v = out_v; pa.out_t = temp_alt;
    pa.psi = folded ? st->psi : nullptr;
    pa.p_out = folded ? (rk == 1 ? st->p_alt : st->p) : nullptr;
    pa.corr_gamma = folded ? c->cfg.gamma[rk - 1] : 0.;
    pa.q_u = impy ? c->d_qf[1] : nullptr;
    pa.q_v = impy ? c->d_qf[2] : nullptr;
    pa.q_t = impy ? c->d_qf[0] : nullptr;
    pa.dtp = c->d_dt;
    pa.alpha = c->cfg.alpha[rk]; pa.beta = c->cfg.beta[rk]; pa.gamma = c->cfg.gamma[rk];
    pa.read_hist = c->cfg.beta[rk] != 0.;
    pa.write_hist = rk < 2;
    pa.with_temp = with_temp;
    pa.forcing = with_temp && c->cfg.with_thermal_forcing;
    pa.lu = c->d_lu ? c->d_lu + (size_t)(rk * 3 * 3) * c->lu_ntm : nullptr;
    pa.lu_hdr = c->d_lu ? c->d_lu + 27 * (size_t)c->lu_ntm + rk * 3 * kLuHdr : nullptr;
    pa.ntm = c->lu_ntm;
    const double *ins[3] = {temp, in_u, in_v};
    double *outs[3] = {temp_alt, out_u, out_v};
    for (int f = 0; f < 3; f++) sa.y[rk][f] = y_args(c, g, rk, f, ins[f], c->d_qf[f], outs[f]);
    sa.pois[rk] = pois_args(c, g, c->cfg.gamma[rk], out_u, out_v, st->psi);
    sa.corr_on[rk] = !fold || rk == 2;
    CorrArgs &ca = sa.corr[rk];
    ca.g = g;
    ca.tb = c->tab;
    ca.tr = c->treat;
    ca.ux_star = out_u; ca.uy_star = out_v; ca.psi = st->psi;
    ca.ux = st->ux; ca.uy = st->uy; ca.p = st->p;
    ca.dtp = c->d_dt;
    ca.gamma = c->cfg.gamma[rk];
    ca.red = (reduce && rk == 2) ? c->d_part : nullptr;
    if (with_temp) std::swap(temp, temp_alt);
    if (fold) {
      std::swap(in_u, out_u);
      std::swap(in_v, out_v);
      if (pa.p_out) in_p = pa.p_out;
    }
  }
}

// One RK3 step with the fused kernels; dt is read from the device slot
// (snsg_set_dt or the device time-step kernel).
static int integrate_fused(snsg_ctx *c, snsg_state *st, bool reduce) {
  int rc;
  if (c->geo_st.jtot < 2) return set_err(SNSG_E_SIZE, "fused step: jtot must be at least 2");
  if ((rc = step_prepare(c))) return rc;
  StepArgs sa;
  step_args(c, st, reduce, sa);
  const bool impx = c->treat.implicitx, impy = c->treat.implicity;
  const bool with_temp = c->cfg.with_temperature != 0;
  if (impx && (rc = run(c, "lu_factor", [&] { launch_lu_factor(c->stream, sa.lu); }))) return rc;
  if (impy && sa.yf.L && (rc = run(c, "ysolve", [&] { launch_yfactor(c->stream, sa.yf); })))
    return rc;
  for (int rk = 0; rk < 3; rk++) {
    if ((rc = run(c, "pred", [&] { launch_pred(c->stream, sa.pred[rk], impx); }))) return rc;
    if (impy) {
      // the partitioned solves of all fields in one launch per kernel
      const int f0 = with_temp ? 0 : 1;
      if (sa.y[rk][f0].L) {
        rc = run(c, "ysolve", [&] { launch_ypart_fields(c->stream, &sa.y[rk][f0], 3 - f0); });
        if (rc) return rc;
      } else {
        for (int f = f0; f < 3; f++) {
          rc = run(c, "ysolve", [&] { launch_ysolve(c->stream, kYKinds[f], sa.y[rk][f]); });
          if (rc) return rc;
        }
      }
    }
    const PoisArgs &pa = sa.pois[rk];
    if ((rc = run(c, "poisson_fwd", [&] { launch_poisson(c->stream, pa, 0); }))) return rc;
    if ((rc = run(c, "poisson_x", [&] { launch_poisson(c->stream, pa, 1); }))) return rc;
    if ((rc = run(c, "poisson_inv", [&] { launch_poisson(c->stream, pa, 2); }))) return rc;
    if (sa.corr_on[rk] && (rc = run(c, "corr", [&] { launch_corr(c->stream, sa.corr[rk]); })))
      return rc;
    if (with_temp) std::swap(st->temp, st->temp_alt);
  }
  return 0;
}

extern "C" int snsg_integrate(snsg_ctx *c, snsg_state *st) {
  if (!c || !st || !st->ux) return set_err(SNSG_E_ARG, __func__);
  return integrate_fused(c, st, false);
}

// Each step decides dt from the partial minima of the advective bound: the
// first from those of the state as passed in (k_adv_bound), later ones from
// those the previous step's last correction stored.
extern "C" int snsg_advance(snsg_ctx *c, snsg_state *st, int nsteps) {
  if (!c || !st || !st->ux || nsteps < 0) return set_err(SNSG_E_ARG, __func__);
  int rc;
  if ((rc = step_prepare(c))) return rc;
  StepArgs sa;
  step_args(c, st, true, sa);
  if ((rc = run(c, "dt_reduce", [&] { launch_adv_bound(c->stream, sa.bound); }))) return rc;
  for (int n = 0; n < nsteps; n++) {
    if ((rc = run(c, "dt", [&] { launch_dt(c->stream, sa.dt); }))) return rc;
    if ((rc = integrate_fused(c, st, true))) return rc;
  }
  return 0;
}

extern "C" int snsg_get_time(snsg_ctx *c, double *time, double *dt) {
  if (!c) return set_err(SNSG_E_ARG, __func__);
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (time) HIPCHECK(hipMemcpyAsync(time, c->d_time, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (dt) HIPCHECK(hipMemcpyAsync(dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int snsg_check_division(snsg_ctx *c, int n, const double *a, const double *b,
                                   const double *r, double *q_cr, double *q_ieee) {
  if (!c || n < 0 || !a || !b || !r || !q_cr || !q_ieee) return set_err(SNSG_E_ARG, __func__);
  if (n == 0) return 0;
  return run(c, "check_div", [&] { launch_check_div(c->stream, n, a, b, r, q_cr, q_ieee); });
}

extern "C" int snsg_set_time(snsg_ctx *c, double time) {
  if (!c) return set_err(SNSG_E_ARG, __func__);
  hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, c->stream, c->d_time, time);
  return check_launch("snsg_set_time");
}

// --------------------------------------------------------------- timing

extern "C" int snsg_timing_enable(snsg_ctx *c, const char *filter) {
  if (!c) return set_err(SNSG_E_ARG, __func__);
  c->timing = true;
  c->filter = filter ? filter : "";
  c->ev_used = 0;
  return 0;
}

extern "C" int snsg_timing_disable(snsg_ctx *c) {
  if (!c) return set_err(SNSG_E_ARG, __func__);
  c->timing = false;
  return 0;
}

extern "C" int snsg_timing_read(snsg_ctx *c, int *launches, double *total_ms) {
  if (!c || !launches || !total_ms) return set_err(SNSG_E_ARG, __func__);
  HIPCHECK(hipStreamSynchronize(c->stream));
  double tot = 0.;
  for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev_pool[k], c->ev_pool[k + 1]));
    tot += ms;
  }
  *launches = (int)(c->ev_used / 2);
  *total_ms = tot;
  c->ev_used = 0;
  return 0;
}
