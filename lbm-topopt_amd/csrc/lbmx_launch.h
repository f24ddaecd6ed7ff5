// Host entry points of the sweep kernels, one instantiation set per lattice.
#pragma once
#include <hip/hip_runtime.h>

#include "lbmx_common.h"

namespace lbmx {

// lattice-specific dispatchers (defined in kernels_<lattice>.hip); they return
// hipErrorInvalidValue for an operator/precision combination that is not instantiated.
#define LBMX_DECLARE_LAUNCHERS(SUFFIX)                                                          \
  hipError_t launch_fwd_##SUFFIX(int coll, int prec, const FwdArgs& a, hipStream_t s);          \
  hipError_t launch_init_##SUFFIX(int coll, int prec, const FwdArgs& a, const double* macros,   \
                                  const double* u, double rho0, double T0, hipStream_t s);      \
  hipError_t launch_adj_##SUFFIX(int coll, int prec, const AdjArgs& a, hipStream_t s);
LBMX_DECLARE_LAUNCHERS(d2q9)
LBMX_DECLARE_LAUNCHERS(d3q19)
LBMX_DECLARE_LAUNCHERS(thermal)
#undef LBMX_DECLARE_LAUNCHERS

}  // namespace lbmx
