// Method of moving asymptotes on the device: the per-variable work of one update (move limits,
// asymptotes, the convex separable approximation, the primal point of the dual multipliers)
// and the sums over the design variables that the dual Newton iteration needs. Sums are
// reduced in a fixed order (per-block partials, then one block over the partials), so an
// update is deterministic. Arithmetic is separately rounded, as in the reference's loops.
#include <hip/hip_runtime.h>

#include "lbmx_launch.h"

namespace lbmx {

namespace {

constexpr int kMmaThreads = 256;

__device__ inline double dmax(double a, double b) { return a > b ? a : b; }
__device__ inline double dmin(double a, double b) { return a < b ? a : b; }

// Sums `val[0..K)` of every thread of the block into part[blockIdx.x * K + k] (thread 0 order:
// a fixed tree over the 256 threads).
template <int K>
__device__ inline void block_sums(double (&val)[K], double* part) {
  __shared__ double red[kMmaThreads];
  for (int k = 0; k < K; ++k) {
    red[threadIdx.x] = val[k];
    __syncthreads();
    for (int s = kMmaThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * K + k] = red[0];
    __syncthreads();
  }
}

// GenSub: move limits clipped to [0, 1], asymptotes, alpha/beta, p0/q0, pij/qij, and the
// partial sums of pij / (U - x) + qij / (x - L) for b.
template <int M>
__global__ __launch_bounds__(kMmaThreads) void k_mma_gensub(MmaArgs a, double* part) {
#pragma clang fp contract(off)
  double acc[M];
  for (int j = 0; j < M; ++j) acc[j] = 0.0;
  const long long n = a.n;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double x = a.x[i];
    const double up = dmin(1.0, x + a.movlim);
    const double lo = dmax(0.0, x - a.movlim);
    a.upper[i] = up;
    a.lower[i] = lo;
    double Li, Ui;
    if (a.k < 3) {
      Li = (x - a.asyminit * up) + a.asyminit * lo;
      Ui = (x + a.asyminit * up) - a.asyminit * lo;
    } else {
      const double x1 = a.xo1[i], x2 = a.xo2[i];
      const double trend = (x - x1) * (x1 - x2);
      const double g = trend < 0.0 ? a.asymdec : (trend > 0.0 ? a.asyminc : 1.0);
      Li = x - g * (x1 - a.L[i]);
      Ui = x + g * (a.U[i] - x1);
      const double span = dmax(1.0e-5, up - lo);
      if (a.asym_type == 0) {
        Li = dmin(dmax(Li, x - 10.0 * span), x - 0.01 * span);
        Ui = dmin(dmax(Ui, x + 0.01 * span), x + 10.0 * span);
      } else {
        Li = dmin(dmax(Li, x - 100.0 * span), x - 1.0e-4 * span);
        Ui = dmin(dmax(Ui, x + 1.0e-4 * span), x + 100.0 * span);
        const double l2 = lo - 1.0e-5, h2 = up + 1.0e-5;
        if (x < l2) {
          Li = x - (h2 - x) / 0.9;
          Ui = x + (h2 - x) / 0.9;
        }
        if (x > h2) {
          Li = x - (x - l2) / 0.9;
          Ui = x + (x - l2) / 0.9;
        }
      }
    }
    a.L[i] = Li;
    a.U[i] = Ui;
    a.alpha[i] = dmax(lo, 0.9 * Li + 0.1 * x);
    a.beta[i] = dmin(up, 0.9 * Ui + 0.1 * x);
    const double ux = Ui - x, xl = x - Li, ul = Ui - Li;
    const double feps = 1.0e-6;
    const double df = a.dfdx[i];
    a.p0[i] = (ux * ux) * (dmax(0.0, df) + 0.001 * fabs(df) + 0.5 * feps / ul);
    a.q0[i] = (xl * xl) * (dmax(0.0, -df) + 0.001 * fabs(df) + 0.5 * feps / ul);
    for (int j = 0; j < M; ++j) {
      const double g = a.dgdx[(size_t)j * n + i];
      double p, q;
      if (a.con_mod) {
        p = (ux * ux) * (dmax(0.0, g) + 0.001 * fabs(g) + 0.5 * feps / ul);
        q = (xl * xl) * (dmax(0.0, -g) + 0.001 * fabs(g) + 0.5 * feps / ul);
      } else {
        p = (ux * ux) * dmax(0.0, g);
        q = (xl * xl) * dmax(0.0, -g);
      }
      a.pij[(size_t)j * n + i] = p;
      a.qij[(size_t)j * n + i] = q;
      acc[j] = acc[j] + (p / ux + q / xl);
    }
  }
  block_sums<M>(acc, part);
}

// The primal point x(lam) (clamped to [alpha, beta]) and the sums of the dual Newton step:
// the constraint approximations sum_i pij / (U - x) + qij / (x - L) (K = M) and, with HESS,
// the Gauss-Newton part of the dual Hessian sum_i PQ_ij d2_i PQ_il (K = M + M*M).
template <int M, bool HESS>
__global__ __launch_bounds__(kMmaThreads) void k_mma_dual(MmaArgs a, MmaLambda lm, double* part) {
#pragma clang fp contract(off)
  constexpr int K = HESS ? M + M * M : M;
  double acc[K];
  for (int j = 0; j < K; ++j) acc[j] = 0.0;
  const long long n = a.n;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double pl = a.p0[i], ql = a.q0[i];
    double pj[M], qj[M];
    for (int j = 0; j < M; ++j) {
      pj[j] = a.pij[(size_t)j * n + i];
      qj[j] = a.qij[(size_t)j * n + i];
      pl = pl + pj[j] * lm.lam[j];
      ql = ql + qj[j] * lm.lam[j];
    }
    const double Li = a.L[i], Ui = a.U[i];
    const double sp = sqrt(pl), sq = sqrt(ql);
    const double xp = (sp * Li + sq * Ui) / (sp + sq);
    double x = xp;
    if (x < a.alpha[i]) x = a.alpha[i];
    if (x > a.beta[i]) x = a.beta[i];
    a.x[i] = x;
    const double ux = Ui - x, xl = x - Li;
    for (int j = 0; j < M; ++j) acc[j] = acc[j] + (pj[j] / ux + qj[j] / xl);
    if constexpr (HESS) {
      double d2 = -1.0 / (2.0 * pl / pow(ux, 3.0) + 2.0 * ql / pow(xl, 3.0));
      if (xp < a.alpha[i] || xp > a.beta[i]) d2 = 0.0;
      double pq[M];
      for (int j = 0; j < M; ++j) pq[j] = pj[j] / (ux * ux) - qj[j] / (xl * xl);
      for (int j = 0; j < M; ++j) {
        const double t = pq[j] * d2;
        for (int l = 0; l < M; ++l) acc[M + j * M + l] = acc[M + j * M + l] + t * pq[l];
      }
    }
  }
  block_sums<K>(acc, part);
}

// out[k] = sum over blocks b of part[b * K + k]: thread t adds the partials t, t + 256, ...
// in order, then a fixed tree over the 256 threads (the same order on every call)
__global__ __launch_bounds__(kMmaThreads) void k_mma_finish(const double* part, int nblocks, int K, double* out) {
#pragma clang fp contract(off)
  __shared__ double red[kMmaThreads];
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kMmaThreads) s = s + part[(size_t)b * K + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kMmaThreads / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
    __syncthreads();
  }
}

// per-block max |xnew - xo1|
__global__ __launch_bounds__(kMmaThreads) void k_mma_change(const double* xnew, const double* xo1, long long n,
                                                              double* part) {
  __shared__ double red[kMmaThreads];
  double v = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    v = dmax(v, fabs(xnew[i] - xo1[i]));
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kMmaThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = dmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kMmaThreads) void k_mma_max(const double* part, int nblocks, double* out) {
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int b = 0; b < nblocks; ++b) v = dmax(v, part[b]);
    out[0] = v;
  }
}

template <int M>
hipError_t gensub_m(const MmaArgs& a, double* part, double* out, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_gensub<M>, dim3(nblocks), dim3(kMmaThreads), 0, s, a, part);
  hipLaunchKernelGGL(k_mma_finish, dim3(1), dim3(kMmaThreads), 0, s, (const double*)part, nblocks, M, out);
  return hipGetLastError();
}

template <int M>
hipError_t dual_m(const MmaArgs& a, const MmaLambda& lm, bool hess, double* part, double* out, int nblocks,
                  hipStream_t s) {
  if (hess) {
    hipLaunchKernelGGL((k_mma_dual<M, true>), dim3(nblocks), dim3(kMmaThreads), 0, s, a, lm, part);
    hipLaunchKernelGGL(k_mma_finish, dim3(1), dim3(kMmaThreads), 0, s, (const double*)part, nblocks, M + M * M, out);
  } else {
    hipLaunchKernelGGL((k_mma_dual<M, false>), dim3(nblocks), dim3(kMmaThreads), 0, s, a, lm, part);
    hipLaunchKernelGGL(k_mma_finish, dim3(1), dim3(kMmaThreads), 0, s, (const double*)part, nblocks, M, out);
  }
  return hipGetLastError();
}

}  // namespace

int mma_blocks(long long n) {
  const long long b = (n + kMmaThreads - 1) / kMmaThreads;
  return (int)(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

hipError_t launch_mma_gensub(const MmaArgs& a, double* part, double* out, hipStream_t s) {
  const int nb = mma_blocks(a.n);
  switch (a.m) {
    case 1: return gensub_m<1>(a, part, out, nb, s);
    case 2: return gensub_m<2>(a, part, out, nb, s);
    case 3: return gensub_m<3>(a, part, out, nb, s);
    case 4: return gensub_m<4>(a, part, out, nb, s);
    case 5: return gensub_m<5>(a, part, out, nb, s);
    case 6: return gensub_m<6>(a, part, out, nb, s);
    case 7: return gensub_m<7>(a, part, out, nb, s);
    case 8: return gensub_m<8>(a, part, out, nb, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mma_dual(const MmaArgs& a, const MmaLambda& lm, bool hess, double* part, double* out,
                           hipStream_t s) {
  const int nb = mma_blocks(a.n);
  switch (a.m) {
    case 1: return dual_m<1>(a, lm, hess, part, out, nb, s);
    case 2: return dual_m<2>(a, lm, hess, part, out, nb, s);
    case 3: return dual_m<3>(a, lm, hess, part, out, nb, s);
    case 4: return dual_m<4>(a, lm, hess, part, out, nb, s);
    case 5: return dual_m<5>(a, lm, hess, part, out, nb, s);
    case 6: return dual_m<6>(a, lm, hess, part, out, nb, s);
    case 7: return dual_m<7>(a, lm, hess, part, out, nb, s);
    case 8: return dual_m<8>(a, lm, hess, part, out, nb, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mma_change(const double* xnew, const double* xo1, long long n, double* part, double* out,
                             hipStream_t s) {
  const int nb = mma_blocks(n);
  hipLaunchKernelGGL(k_mma_change, dim3(nb), dim3(kMmaThreads), 0, s, xnew, xo1, n, part);
  hipLaunchKernelGGL(k_mma_max, dim3(1), dim3(kMmaThreads), 0, s, (const double*)part, nb, out);
  return hipGetLastError();
}

}  // namespace lbmx
