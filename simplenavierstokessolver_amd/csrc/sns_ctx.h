// sns_ctx.h -- the instance context behind the C-ABI and the helpers the
// C-ABI translation units share (internal).
#pragma once
#include <stddef.h>
#include <string>
#include <vector>

#include "sns_hip.h"
#include "sns_launch.h"
#include "sns_fused.h"
#include "sns_poisson.h"

using namespace sns;

struct snsg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  snsg_config cfg{};
  Geo geo{};     // reference layouts (per-function entry points)
  Geo geo_st{};  // padded row pitches of the fused state (snsg_state)
  Treat treat{};
  double *d_dxf = nullptr, *d_dxc = nullptr;
  double *d_dt = nullptr;
  double *d_eigen = nullptr;
  double *d_pfac = nullptr;             // Thomas factors of the Poisson x systems
  double *d_plu = nullptr;              // their shared row coefficients l, u (interleaved)
  double2 *d_tw = nullptr;
  double2 *d_psihat = nullptr;
  double *d_q = nullptr, *d_qy = nullptr;
  unsigned long long *d_red = nullptr;
  unsigned long long *h_red = nullptr;  // pinned
  double *d_part = nullptr;             // per-workgroup minima of the fused correction
  double *d_diag = nullptr;             // diagnostics: partial slots, then the results
  double *h_diag = nullptr;             // pinned copy of the results
  int nparts = 0;
  Fft2Plan plan{};
  snsg_launch launch{};  // requested launch shapes (0: the rule)
  int in_batch = 0;      // batches that hold this instance
  double2 *d_Y = nullptr;  // four-step FFT intermediate, jtot rows of (itot+1)/2 pairs
  // per-column tables of the fused kernels (one allocation)
  double *d_tab = nullptr;
  double2 *d_tab2 = nullptr;           // paired column tables
  Tables tab{};
  // fused-step scratch: x-aligned increments of T, ux, uy when a y solve follows
  double *d_qf[3] = {nullptr, nullptr, nullptr};
  // partitioned y solves: chunk length (0: transposed solves), factor
  // tables of the 3 substages x 3 fields, interface values
  int ychunk = 0;
  double *d_yfac = nullptr, *d_qz = nullptr;
  // LU tables of the CN-x operators (3 substages x 3 fields x 5 tables)
  double *d_lu = nullptr;
  int lu_ntm = 0;
  double *d_time = nullptr;
  size_t mem = 0;  // device bytes held by the context and its state
  double dt_dif_x = 0., dt_dif_y = 0.;
  // timing
  bool timing = false;
  std::string filter;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
};

// Records `what` as the thread's last error and returns code.
int set_err(int code, const char *what);
int check_launch(const char *what);

#define HIPCHECK(expr)                                          \
  do {                                                          \
    hipError_t e_ = (expr);                                     \
    if (e_ != hipSuccess) return set_err((int)e_, #expr);       \
  } while (0)

// Zero-filled device allocation, ready for any stream on return; adds the
// bytes to *count when count is given.
int ctx_alloc(void **p, size_t bytes, size_t *count = nullptr);

// (Re)initialises an instance whose buffers exist for cfg's itot and jtot.
int ctx_configure(snsg_ctx *c, const snsg_config *cfg, const double *dxf, const double *dxc);

RK rk_of(const snsg_ctx *c, int rkstep);
PoisArgs pois_args(const snsg_ctx *c, const Geo &g, double gamma, const double *ux,
                   const double *uy, double *psi);

// Argument blocks of one fused RK3 step (src/main.c:79-101) of instance c on
// state st, with st->temp and st->temp_alt as they are at the start of the
// step (T and T' trade places after every substage).  `reduce`: the last
// correction stores the advective-bound minima for the next step in d_part.
struct StepArgs {
  LuFactorArgs lu;
  DtArgs dt;
  CorrArgs bound;  // advective bound of the state as it is (d_part)
  YFactorArgs yf;  // factors of the partitioned y solves (ychunk > 0)
  PredArgs pred[3];
  YArgs y[3][3];   // [substage][T, ux, uy]
  PoisArgs pois[3];
  CorrArgs corr[3];
  // substage ends with the correction kernel (otherwise the next substage's
  // predictor applies it as it loads the fields)
  bool corr_on[3];
};

// field kinds of the y solves, in the order T, ux, uy
extern const int kYKinds[3];

// Zeroes every array of a state on the instance's stream.
int state_reset(snsg_ctx *c, snsg_state *st);

// Scratch of the fused step, allocated on first use
int step_prepare(snsg_ctx *c);

// Launch shapes: the table entry of (itot, jtot) (zeros when there is none),
// and the shapes that decide a batch's shape classes, rules resolved where
// they are static: FFT factor n1, chunk length, interface-solve width,
// recompute choice (0: decided at launch).
snsg_launch tuned_launch(int itot, int jtot);
struct LaunchKey {
  int n1, chunk, cols, recompute, split;
};
LaunchKey launch_key(int itot, int jtot, const snsg_launch &req);
LaunchKey launch_key(const snsg_ctx *c);
void step_args(const snsg_ctx *c, const snsg_state *st, bool reduce, StepArgs &sa);
