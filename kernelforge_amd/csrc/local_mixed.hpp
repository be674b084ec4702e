// local_mixed.hpp -- entry points of the BF16x3 kernels (local_mixed.hip), called by the
// C ABI in local_scalar.hip for precision KF_PREC_BF16X3.
#pragma once

#include "kf_common.hpp"
#include "local_plan.hpp"

namespace kf {
int local_mixed_symm(int dense, const double *X, int nm, int ma, int rep, double sigma,
                     const LocalPlan &p, void *out, hipStream_t stream);
int local_mixed_rect(const double *X1, const double *X2, int nm1, int nm2, int ma1, int ma2,
                     int rep, double sigma, const LocalPlan &p1, const LocalPlan &p2, float *out,
                     hipStream_t stream);
}  // namespace kf
