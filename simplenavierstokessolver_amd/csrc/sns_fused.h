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
  const double *lu;                       // implicit x: LU tables of the substage (3 fields)
  const double *lu_hdr;                   // their headers (kLuHdr doubles per field)
  int ntm;                                // entries per LU table
  // Folded correction (implicit x, explicit y): when psi is set, ux, uy and
  // p are the previous substage's ux*, uy* and pressure, corrected with psi
  // and corr_gamma*dt as they are loaded (k_corr's operations); the
  // uy workgroups store the corrected pressure of their rows in p_out.
  const double *psi;
  double *p_out;
  double corr_gamma;
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

// Time-step decision of one instance from the partial minima of the
// advective bound (k_corr / k_adv_bound)
struct DtArgs {
  const double *part;
  int n;
  double *dt, *time;
  double dt_dif_x, dt_dif_y;
  int implicitx, implicity;
};

// Thomas factorisation tables of the three CN-x operators for the three
// substages: out holds 9 groups (substage s, field f at (s*3+f)) of 3 tables
// of ntm entries each (see sns_lu.h), then 9 headers of kLuHdr doubles:
// the row P from which the rows up to n-2 all repeat one (r, al, be) triple
// bit for bit, and that triple (lu_table_doubles in all)
constexpr int kLuHdr = 4;
inline size_t lu_table_doubles(int ntm) { return 27 * (size_t)ntm + 9 * kLuHdr; }
// threads of the row kernels for itot (the tables interleave lu_table_entries
// = threads x 8 unknowns per thread); 0 where itot is not supported
int lu_row_threads(int itot);

// Crank-Nicolson in x on the reference layouts (the per-function entry
// points): the increment of each row from the RK sources (or qin), solved
// with the LU tables of one substage and field (lu, lu_hdr: as PredArgs),
// then added to `field` with boundary values, or stored to qout.
struct CnxLuArgs {
  Geo g;
  RK rk;
  const double *dtp;
  const double *srca, *srcb, *srcg, *qin;
  double *field, *qout;
  const double *lu, *lu_hdr;
  int ntm;
};
void launch_cnx_lu(hipStream_t s, int kind, const CnxLuArgs &a);
struct LuFactorArgs {
  Geo g;
  Tables tb;
  const double *dtp;
  double gamma[3];
  double *out;
  int ntm;
};

// Factors of the partitioned Crank-Nicolson y solves of all three substages
// and fields (sns_ysolve.hip): table (s*3+f) at out + (s*3+f)*stride
struct YFactorArgs {
  Geo g;
  const double *dtp;
  double gamma[3];
  double *out;
  int L, S;
  long stride;
};
// chunk length of the partitioned y solve for jtot rows (0: not applicable)
int ysolve_chunk(int jtot);
bool ysolve_chunk_ok(int jtot, int L);
int ysolve_cols_rule(int SR);
size_t ysolve_table_doubles(int L);
void launch_yfactor(hipStream_t s, const YFactorArgs &a);
void launch_yfactor_batch(hipStream_t s, const YFactorArgs *dev, int nz);
// The partitioned CN-y solve (a.L > 0) on the row-major increments a.qf,
// then out = in + increment with boundary values, for nf consecutive fields
// of the order T, ux, uy (a[0..nf-1]), each kernel launched once for all
void launch_ypart_fields(hipStream_t s, const YArgs *a, int nf);
// and for nz instances: dev holds nf blocks of nz arguments, field-major
void launch_ypart_fields_batch(hipStream_t s, const YArgs &shape, const YArgs *dev, int nz,
                               int nf);

bool pred_impx_supported(int itot);
int lu_table_entries(int itot);
void launch_lu_factor(hipStream_t s, const LuFactorArgs &a);
// predictor of one substage; a.lu, a.ntm: the LU tables when implicitx
void launch_pred(hipStream_t s, const PredArgs &a, int implicitx);
// implicit-x predictor (sns_pred.hip)
void launch_pred_rows(hipStream_t s, const PredArgs &a);
void launch_pred_rows_batch(hipStream_t s, const PredArgs &shape, const PredArgs *dev, int nz);
// y solve of one field (kind) after the predictor: transpose the x-aligned
// increments, periodic Crank-Nicolson solve per column, transpose back,
// out = in + increment with boundary values
void launch_ysolve(hipStream_t s, int kind, const YArgs &a);
int corr_partials(const Geo &g);
void launch_corr(hipStream_t s, const CorrArgs &a);
// the advective bound of ux, uy over the workgroups of k_corr into a.red
// (the same partial minima the last correction of a step leaves)
void launch_adv_bound(hipStream_t s, const CorrArgs &a);
void launch_dt(hipStream_t s, const DtArgs &a);
void launch_check_div(hipStream_t s, int n, const double *a, const double *b, const double *r,
                      double *q_cr, double *q_ieee);

// The same kernels for nz instances of one shape (itot, jtot, treatment,
// temperature equation): `shape` gives the launch configuration, dev the
// per-instance arguments (a device array of nz blocks, instance = blockIdx.z).
void launch_lu_factor_batch(hipStream_t s, const LuFactorArgs &shape, const LuFactorArgs *dev,
                            int nz);
void launch_pred_batch(hipStream_t s, const PredArgs &shape, int implicitx, const PredArgs *dev,
                       int nz);
void launch_ysolve_batch(hipStream_t s, int kind, const YArgs &shape, const YArgs *dev, int nz);
void launch_corr_batch(hipStream_t s, const CorrArgs &shape, const CorrArgs *dev, int nz);
void launch_adv_bound_batch(hipStream_t s, const CorrArgs &shape, const CorrArgs *dev, int nz);
void launch_dt_batch(hipStream_t s, const DtArgs *dev, int nz);

}  // namespace sns
