// sns_fused.h -- arguments and launchers of the fused substage kernels.
#pragma once
#include "sns_launch.h"
#include "sns_fast.h"

namespace sns {

struct PredArgs {
  Geo g;
  Tables tb;
  Treat tr;
  const double *ux, *uy, *p, *temp;       // substage-start fields
  double *hist_ux, *hist_uy, *hist_t;     // RK history of the explicit terms (in place)
  double *out_u, *out_v, *out_t;          // ux*, uy*, T' with boundary values
  double *q_u, *q_v, *q_t;                // x-aligned increments (set when a y solve follows)
  const double *dtp;
  double alpha, beta, gamma;
  int read_hist, write_hist;
  int with_temp;                          // temperature equation integrated
  int forcing;                            // buoyancy term from T
};

struct CorrArgs {
  Geo g;
  Tables tb;
  Treat tr;
  const double *ux_star, *uy_star, *psi;
  double *ux, *uy, *p;
  const double *dtp;
  double gamma;
  double *red;  // per-workgroup advective-bound minima (corr_partials entries), or nullptr
};

// Thomas factorisation tables of the three CN-x operators for the three
// substages: out holds 9 groups (substage s, field f at (s*3+f)) of 3 tables
// of ntm entries each (see sns_lu.h)
struct LuFactorArgs {
  Geo g;
  Tables tb;
  const double *dtp;
  double gamma[3];
  double *out;
  int ntm;
};

bool pred_impx_supported(int itot);
int lu_table_entries(int itot);
void launch_lu_factor(hipStream_t s, const LuFactorArgs &a);
void launch_pred(hipStream_t s, const PredArgs &a, int implicitx, const double *lu, int ntm);
// implicit-x predictor (sns_pred.hip); lu: the 3 field groups of one substage
void launch_pred_rows(hipStream_t s, const PredArgs &a, const double *lu, int ntm);
void launch_add_out(hipStream_t s, const Geo &g, int kind, const double *in, const double *q,
                    double *out);
int corr_partials(const Geo &g);
void launch_corr(hipStream_t s, const CorrArgs &a);
void launch_dt(hipStream_t s, const double *part, int n, double *dt, double *time,
               double dt_dif_x, double dt_dif_y, int implicitx, int implicity);
void launch_check_div(hipStream_t s, int n, const double *a, const double *b, const double *r,
                      double *q_cr, double *q_ieee);

}  // namespace sns
