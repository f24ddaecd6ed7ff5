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
  hipError_t launch_adj_##SUFFIX(int coll, int des, const AdjArgs& a, hipStream_t s);                    \
  hipError_t launch_neumann_fix_##SUFFIX(const DevBC& e, void* lam, const Layout& ll, const double* nside,  \
                                         hipStream_t s);
#define LBMX_DECLARE_LATTICE(LAT) \
  LBMX_DECLARE_LAUNCHERS(LAT##_f64) LBMX_DECLARE_LAUNCHERS(LAT##_f32) LBMX_DECLARE_LAUNCHERS(LAT##_f16) LBMX_DECLARE_LAUNCHERS(LAT##_bf16)
LBMX_DECLARE_LATTICE(d2q9)
LBMX_DECLARE_LATTICE(d3q19)
LBMX_DECLARE_LATTICE(thermal)
#undef LBMX_DECLARE_LATTICE
#undef LBMX_DECLARE_LAUNCHERS

// Zero fill of a device range (bytes a multiple of 4) by a kernel.
hipError_t fill_zero(void* p, size_t bytes, hipStream_t s);

// FP64 -> FP32 conversion of n values (design fields of the reduced-storage forms).
hipError_t convert_f64_f32(float* out, const double* in, long long n, hipStream_t s);

// Design pipeline kernels (natural-order FP64 fields on the full grid).
struct ConeArgs {
  const double* in;    // field to filter (forward) or s / (W 1) (transposed)
  double* out;
  const double* den;   // forward: divide by den inside the box (nullptr: no division)
  const double* wt;    // weights [(dz+R)(2R+1) + dy+R][dx+R], (2R+1)^3 entries (2D: dz = 0 only)
  const int* rx;       // per (dz, dy): largest |dx| with dx^2+dy^2+dz^2 <= R^2, or -1
  int nx, ny, nz, R;
  int box[6];          // filter box, inclusive, clipped to the grid
  int transposed;      // 0: out = W in (rows in the box), 1: out = W^T in
};
hipError_t launch_cone(const ConeArgs& a, hipStream_t s);
// out[i] = inside design box ? design[box-local i] : background[i]
hipError_t launch_map_to_full(double* out, const double* design, const double* background, int nx, int ny, int nz,
                              const int box[6], hipStream_t s);
// design[box-local i] = full[i] over the design box
hipError_t launch_extract_box(double* design, const double* full, int nx, int ny, int nz, const int box[6],
                              hipStream_t s);
// out = inside box ? (tb + tanh(beta (x - eta))) / tsum : x, with tb = tanh(beta eta) and
// tsum = tb + tanh(beta (1 - eta)); with a per-node eta field (etaf != nullptr) eta, tb and
// tsum are those of etaf[i]
hipError_t launch_project(double* out, const double* x, int nx, int ny, int nz, const int box[6], double beta,
                          double eta, double tb, double tsum, const double* etaf, hipStream_t s);
// out = (inside box && project ? sens * (1 - tanh(beta (x - eta))^2) beta / tsum : sens) / den
hipError_t launch_sens_scale(double* out, const double* sens, const double* x, const double* den, int nx, int ny,
                             int nz, const int box[6], int project, double beta, double eta, double tsum,
                             const double* etaf, hipStream_t s);

// Method of moving asymptotes: per-variable arrays of one optimiser (n variables, m <= 8
// constraints; pij, qij, dgdx hold m rows of n) and the multipliers of a dual evaluation.
constexpr int kMmaMaxM = 8;
struct MmaArgs {
  long long n;
  int m, k, con_mod, asym_type;
  double movlim, asyminit, asymdec, asyminc;
  double* x;
  const double* dfdx;
  const double* dgdx;
  const double* xo1;
  const double* xo2;
  double *lower, *upper, *L, *U, *alpha, *beta, *p0, *q0, *pij, *qij;
};
struct MmaLambda {
  double lam[kMmaMaxM];
};
int mma_blocks(long long n);  // blocks of the reductions (partials per block)
// subproblem at a.x; out[0..m) = sum_i pij / (U - x) + qij / (x - L)
hipError_t launch_mma_gensub(const MmaArgs& a, double* part, double* out, hipStream_t s);
// a.x = x(lam); out[0..m) = the same sums at x(lam); with hess, out[m + j*m + l] = the
// Gauss-Newton part of the dual Hessian
hipError_t launch_mma_dual(const MmaArgs& a, const MmaLambda& lm, bool hess, double* part, double* out,
                           hipStream_t s);
// out[0] = max |xnew - xo1|
hipError_t launch_mma_change(const double* xnew, const double* xo1, long long n, double* part, double* out,
                             hipStream_t s);

// Batched sweeps (kind 0 forward, 1 adjoint) over every instance of a batch; ty = tile rows,
// pair = the forward of a 2- or 4-row tile takes two nodes per thread.
#define LBMX_DECLARE_BATCH(SUFFIX) \
  hipError_t launch_batch_##SUFFIX(int coll, int des, int ty, bool pair, int kind, const BatchStep& st, hipStream_t s);
LBMX_DECLARE_BATCH(d2q9_f64) LBMX_DECLARE_BATCH(d2q9_f32) LBMX_DECLARE_BATCH(d2q9_f16) LBMX_DECLARE_BATCH(d2q9_bf16)
LBMX_DECLARE_BATCH(d3q19_f64) LBMX_DECLARE_BATCH(d3q19_f32) LBMX_DECLARE_BATCH(d3q19_f16) LBMX_DECLARE_BATCH(d3q19_bf16)
LBMX_DECLARE_BATCH(thermal_f64) LBMX_DECLARE_BATCH(thermal_f32) LBMX_DECLARE_BATCH(thermal_f16) LBMX_DECLARE_BATCH(thermal_bf16)
#undef LBMX_DECLARE_BATCH

}  // namespace lbmx
