// Host entry points of the sweep kernels, one instantiation set per (lattice, storage type).
#pragma once
#include <hip/hip_runtime.h>

#include "lbmx_common.h"

namespace lbmx {

// Each returns hipErrorInvalidValue for an operator/coupling combination that is not built.
#define LBMX_DECLARE_LAUNCHERS(SUFFIX)                                                                   \
  hipError_t launch_fwd_##SUFFIX(int coll, int des, const FwdArgs& a, hipStream_t s);                    \
  hipError_t launch_init_##SUFFIX(int coll, int des, const FwdArgs& a, const double* macros,             \
                                  const double* u, double rho0, double T0, hipStream_t s);               \
  hipError_t launch_adj_##SUFFIX(int coll, int des, const AdjArgs& a, hipStream_t s);
#define LBMX_DECLARE_LATTICE(LAT) \
  LBMX_DECLARE_LAUNCHERS(LAT##_f64) LBMX_DECLARE_LAUNCHERS(LAT##_f32) LBMX_DECLARE_LAUNCHERS(LAT##_f16) LBMX_DECLARE_LAUNCHERS(LAT##_bf16)
LBMX_DECLARE_LATTICE(d2q9)
LBMX_DECLARE_LATTICE(d3q19)
LBMX_DECLARE_LATTICE(thermal)
#undef LBMX_DECLARE_LATTICE
#undef LBMX_DECLARE_LAUNCHERS

// Zero fill of a device range (bytes a multiple of 4) by a kernel.
hipError_t fill_zero(void* p, size_t bytes, hipStream_t s);

// Batched sweeps (kind 0 forward, 1 adjoint) over every instance of a batch; ty = tile rows.
hipError_t launch_batch_d2q9_f32(int coll, int des, int ty, int kind, const BatchStep& st, hipStream_t s);
hipError_t launch_batch_d2q9_f64(int coll, int des, int ty, int kind, const BatchStep& st, hipStream_t s);

}  // namespace lbmx
