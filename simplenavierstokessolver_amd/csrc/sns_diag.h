// sns_diag.h -- monitoring diagnostics on the device.
#pragma once
#include "sns_launch.h"

namespace sns {

// divmax, divsum, momentum x/y, energies ux/uy/T, Nusselt wall/injection/
// kinetic dissipation/thermal dissipation
constexpr int kDiagSlots = 11;

struct DiagArgs {
  const double *ux, *uy, *temp;  // temp may be null (no temperature equation)
};

// statistics accumulators (reference layouts); temp / temp1 / temp2 may be null
struct StatArgs {
  const double *ux, *uy, *temp;
  double *ux1, *ux2, *uy1, *uy2, *temp1, *temp2;
};
void launch_stat_collect(hipStream_t s, const Geo &gf, const StatArgs &a);

int diag_partials(const Geo &g);
// out (device, kDiagSlots doubles); part: diag_partials(g)*kDiagSlots doubles
void launch_diag(hipStream_t s, const Geo &g, const DiagArgs &a, double *part, double lx,
                 double ly, double *out);

}  // namespace sns
