// Method of moving asymptotes behind the lbmx_mma_* entry points (src/topOpt/MMA.cc, the
// reference's PETSc MMA with robust constants). Per-variable arrays live in HBM; the m
// multipliers and the m x m dual Newton system stay on the host, fed by the device sums of
// each dual evaluation.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "host_common.hpp"
#include "lbmx.h"
#include "lbmx_launch.h"

namespace lbmx {

struct Mma {
  long long n = 0;
  int m = 0, k = 0;
  double asyminit = 0.5, asymdec = 0.7, asyminc = 1.2;
  int con_mod = 0, asym_type = 0;
  std::vector<double> a, c, y, lam, mu, s, b, grad, hess;
  double z = 0.0;
  int newton_iters = 0;
  double last_ms = 0.0;
  hipStream_t st = nullptr;
  DeviceBuffer x, dfdx, dgdx, xo1, xo2, lower, upper, L, U, alpha, beta, p0, q0, pij, qij, part, out, tmp;

  Mma(long long nn, int mm) : n(nn), m(mm) {
    if (n < 1) throw Error{LBMX_ERR_ARG, "MMA needs at least one design variable"};
    if (m < 1 || m > LBMX_MMA_MAX_CONSTRAINTS) throw Error{LBMX_ERR_ARG, "MMA constraints out of range 1..8"};
    a.assign(m, 0.0);
    c.assign(m, 1000.0);
    y.assign(m, 0.0);
    lam.assign(m, 0.0);
    mu.assign(m, 0.0);
    s.assign(2 * m, 0.0);
    b.assign(m, 0.0);
    grad.assign(m, 0.0);
    hess.assign((size_t)m * m, 0.0);
    const size_t nb = (size_t)n * 8;
    for (DeviceBuffer* d : {&x, &dfdx, &xo1, &xo2, &lower, &upper, &L, &U, &alpha, &beta, &p0, &q0, &tmp}) d->ensure(nb);
    for (DeviceBuffer* d : {&dgdx, &pij, &qij}) d->ensure(nb * m);
    part.ensure((size_t)mma_blocks(n) * (m + m * m) * 8);
    out.ensure((size_t)(m + m * m) * 8);
    HIP_OK(hipStreamCreate(&st));
  }
  ~Mma() {
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }

  MmaArgs args(double movlim) {
    MmaArgs r{};
    r.n = n;
    r.m = m;
    r.k = k;
    r.con_mod = con_mod;
    r.asym_type = asym_type;
    r.movlim = movlim;
    r.asyminit = asyminit;
    r.asymdec = asymdec;
    r.asyminc = asyminc;
    r.x = x.as<double>();
    r.dfdx = dfdx.as<double>();
    r.dgdx = dgdx.as<double>();
    r.xo1 = xo1.as<double>();
    r.xo2 = xo2.as<double>();
    r.lower = lower.as<double>();
    r.upper = upper.as<double>();
    r.L = L.as<double>();
    r.U = U.as<double>();
    r.alpha = alpha.as<double>();
    r.beta = beta.as<double>();
    r.p0 = p0.as<double>();
    r.q0 = q0.as<double>();
    r.pij = pij.as<double>();
    r.qij = qij.as<double>();
    return r;
  }

  void fetch(double* host, int count) {
    HIP_OK(hipMemcpyAsync(host, out.p, (size_t)count * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
  }

  // y, z of the current multipliers (the scalar part of XYZofLAMBDA, MMA.cc:697-716)
  void primal_scalars() {
    double la = 0.0;
    for (int j = 0; j < m; ++j) {
      if (lam[j] < 0.0) lam[j] = 0.0;
      y[j] = std::max(0.0, lam[j] - c[j]);
      la += lam[j] * a[j];
    }
    z = std::max(0.0, 10.0 * (la - 1.0));
  }

  // x(lam) on the device and the sums of the dual gradient (and Hessian) at it
  void dual_eval(const MmaArgs& ar, bool with_hess, double* sums) {
    MmaLambda lm{};
    for (int j = 0; j < m; ++j) lm.lam[j] = lam[j];
    HIP_OK(launch_mma_dual(ar, lm, with_hess, part.as<double>(), out.as<double>(), st));
    fetch(sums, with_hess ? m + m * m : m);
  }

  // the host part of DualHess (MMA.cc:843-874) on top of the device sums
  void finish_hessian(const double* h) {
    for (int i = 0; i < m * m; ++i) hess[i] = h[i];
    double la = 0.0;
    for (int j = 0; j < m; ++j) {
      if (lam[j] < 0.0) lam[j] = 0.0;
      la += lam[j] * a[j];
      if (lam[j] > c[j]) hess[j * m + j] += -1.0;
      hess[j * m + j] += -mu[j] / lam[j];
    }
    if (la > 0.0)
      for (int j = 0; j < m; ++j)
        for (int l = 0; l < m; ++l) hess[j * m + l] += -10.0 * a[j] * a[l];
    double tr = 0.0;
    for (int j = 0; j < m; ++j) tr += hess[j * m + j];
    double corr = 1e-4 * tr / m;
    if (-corr < 1.0e-7) corr = -1.0e-7;
    for (int j = 0; j < m; ++j) hess[j * m + j] += corr;
  }

  // LU without pivoting of the m x m dual Hessian and the solve (MMA.cc:952-986)
  void lu_solve(std::vector<double>& K, std::vector<double>& r) {
    const int nn = m;
    for (int p = 0; p < nn - 1; ++p)
      for (int i = p + 1; i < nn; ++i) {
        K[i * nn + p] = K[i * nn + p] / K[p * nn + p];
        for (int j = p + 1; j < nn; ++j) K[i * nn + j] = K[i * nn + j] - K[i * nn + p] * K[p * nn + j];
      }
    for (int i = 1; i < nn; ++i) {
      double acc = 0.0;
      for (int j = 0; j < i; ++j) acc = acc - K[i * nn + j] * r[j];
      r[i] = r[i] + acc;
    }
    r[nn - 1] = r[nn - 1] / K[(nn - 1) * nn + (nn - 1)];
    for (int i = nn - 2; i >= 0; --i) {
      double acc = r[i];
      for (int j = i + 1; j < nn; ++j) acc = acc - K[i * nn + j] * r[j];
      r[i] = acc / K[i * nn + i];
    }
  }

  // Update (MMA.cc:485-524): GenSub, history shift, dual interior-point solve. The caller's
  // arrays are host arrays, or device arrays when dev is set.
  void update(double* xh, const double* dfdxh, const double* gx, const double* const* dgdxh, double movlim,
              bool dev) {
    auto t0 = std::chrono::steady_clock::now();
    const size_t nb = (size_t)n * 8;
    const hipMemcpyKind in = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIP_OK(hipMemcpyAsync(x.p, xh, nb, in, st));
    HIP_OK(hipMemcpyAsync(dfdx.p, dfdxh, nb, in, st));
    for (int j = 0; j < m; ++j)
      HIP_OK(hipMemcpyAsync(dgdx.as<double>() + (size_t)j * n, dgdxh[j], nb, in, st));
    k += 1;
    MmaArgs ar = args(movlim);
    HIP_OK(launch_mma_gensub(ar, part.as<double>(), out.as<double>(), st));
    std::vector<double> sums((size_t)m + (size_t)m * m);
    fetch(sums.data(), m);
    for (int j = 0; j < m; ++j) b[j] = sums[j] - gx[j];
    // xo2 <- xo1, xo1 <- x (the old xo2 buffer is reused)
    std::swap(xo1.p, xo2.p);
    HIP_OK(hipMemcpyAsync(xo1.p, x.p, nb, hipMemcpyDeviceToDevice, st));
    ar = args(movlim);
    solve_dual(ar, sums);
    HIP_OK(hipMemcpyAsync(xh, x.p, nb, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // SolveDIP with DualGrad, DualHess, DualLineSearch and DualResidual (MMA.cc:655-950)
  void solve_dual(const MmaArgs& ar, std::vector<double>& sums) {
    for (int j = 0; j < m; ++j) {
      lam[j] = c[j] / 2.0;
      mu[j] = 1.0;
    }
    const double tol = 1.0e-9 * std::sqrt((double)(m + n));
    double epsi = 1.0, err = 1.0;
    newton_iters = 0;
    while (epsi > tol) {
      int loop = 0;
      while (err > 0.9 * epsi && loop < 100) {
        ++loop;
        ++newton_iters;
        primal_scalars();
        dual_eval(ar, true, sums.data());
        for (int j = 0; j < m; ++j) grad[j] = sums[j];
        for (int j = 0; j < m; ++j) grad[j] += -b[j] - a[j] * z - y[j];
        for (int j = 0; j < m; ++j) grad[j] = -1.0 * grad[j] - epsi / lam[j];
        finish_hessian(sums.data() + m);
        lu_solve(hess, grad);
        for (int j = 0; j < m; ++j) s[j] = grad[j];
        for (int j = 0; j < m; ++j) s[m + j] = -mu[j] + epsi / lam[j] - s[j] * mu[j] / lam[j];
        double theta = 1.005;
        for (int j = 0; j < m; ++j) {
          if (theta < -1.01 * s[j] / lam[j]) theta = -1.01 * s[j] / lam[j];
          if (theta < -1.01 * s[j + m] / mu[j]) theta = -1.01 * s[j + m] / mu[j];
        }
        theta = 1.0 / theta;
        for (int j = 0; j < m; ++j) {
          lam[j] = lam[j] + theta * s[j];
          mu[j] = mu[j] + theta * s[j + m];
        }
        primal_scalars();
        dual_eval(ar, false, sums.data());
        double nr = 0.0;
        for (int j = 0; j < m; ++j) {
          const double r1 = sums[j] + (-b[j] - a[j] * z - y[j] + mu[j]);
          const double r2 = mu[j] * lam[j] - epsi;
          nr = std::max(nr, std::max(std::fabs(r1), std::fabs(r2)));
        }
        err = nr;
      }
      epsi = epsi * 0.1;
    }
  }

  double design_change(const double* xnew, bool dev) {
    const double* xn = x.as<double>();
    if (xnew && dev) {
      xn = xnew;
    } else if (xnew) {
      HIP_OK(hipMemcpyAsync(tmp.p, xnew, (size_t)n * 8, hipMemcpyHostToDevice, st));
      xn = tmp.as<double>();
    }
    HIP_OK(launch_mma_change(xn, xo1.as<double>(), n, part.as<double>(), out.as<double>(), st));
    double ch = 0.0;
    fetch(&ch, 1);
    return ch;
  }
};

}  // namespace lbmx

struct lbmx_mma {
  std::unique_ptr<lbmx::Mma> o;
};

// Scratch of lbmx_design_change_noalloc: the per-block partial maxima of the change reduction
// (mma_blocks() launches at most kNoAllocPartials blocks). Each device holds its own copy for
// the life of the process; its address on the current device comes from hipGetSymbolAddress
// (the array's name is a host-side stand-in, never a device address).
static constexpr int kNoAllocPartials = 2048;
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device, filled on the first call there

double* scratch_on_current_device() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw lbmx::Error{LBMX_ERR_DEVICE, std::string("hipGetSymbolAddress of the scratch array failed: ") +
                                            hipGetErrorString(e)};
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

using lbmx::Error;
using lbmx::guarded;

#define CHECK_MMA(h)                                    \
  if (!(h) || !(h)->o) {                                \
    lbmx::g_last_error = "null MMA handle";             \
    return LBMX_ERR_ARG;                                \
  }

extern "C" {

lbmx_status lbmx_mma_create(long long n, int m, lbmx_mma** out) {
  if (!out) {
    lbmx::g_last_error = "null argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    auto h = std::make_unique<lbmx_mma>();
    h->o = std::make_unique<lbmx::Mma>(n, m);
    *out = h.release();
  });
}

lbmx_status lbmx_mma_destroy(lbmx_mma* h) {
  delete h;
  return LBMX_OK;
}

lbmx_status lbmx_mma_set_robust_constants(lbmx_mma* h, const double* a) {
  CHECK_MMA(h);
  if (!a) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  for (int j = 0; j < h->o->m; ++j) h->o->a[j] = a[j];
  return LBMX_OK;
}

lbmx_status lbmx_mma_set_asymptotes(lbmx_mma* h, double init, double decrease, double increase) {
  CHECK_MMA(h);
  h->o->asyminit = init;
  h->o->asymdec = decrease;
  h->o->asyminc = increase;
  return LBMX_OK;
}

lbmx_status lbmx_mma_set_options(lbmx_mma* h, int constraint_modification, int robust_asymptotes_type) {
  CHECK_MMA(h);
  if (robust_asymptotes_type != 0 && robust_asymptotes_type != 1) {
    lbmx::g_last_error = "robust asymptotes type must be 0 or 1";
    return LBMX_ERR_ARG;
  }
  h->o->con_mod = constraint_modification != 0;
  h->o->asym_type = robust_asymptotes_type;
  return LBMX_OK;
}

lbmx_status lbmx_mma_update(lbmx_mma* h, double* x, const double* dfdx, const double* gx, const double* const* dgdx,
                            double movlim) {
  CHECK_MMA(h);
  if (!x || !dfdx || !gx || !dgdx) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  for (int j = 0; j < h->o->m; ++j)
    if (!dgdx[j]) { lbmx::g_last_error = "null constraint sensitivity"; return LBMX_ERR_ARG; }
  return guarded([&] { h->o->update(x, dfdx, gx, dgdx, movlim, false); });
}

lbmx_status lbmx_mma_update_device(lbmx_mma* h, double* x, const double* dfdx, const double* gx,
                                   const double* const* dgdx, double movlim) {
  CHECK_MMA(h);
  if (!x || !dfdx || !gx || !dgdx) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  for (int j = 0; j < h->o->m; ++j)
    if (!dgdx[j]) { lbmx::g_last_error = "null constraint sensitivity"; return LBMX_ERR_ARG; }
  return guarded([&] { h->o->update(x, dfdx, gx, dgdx, movlim, true); });
}

lbmx_status lbmx_mma_design_change(lbmx_mma* h, const double* xnew, double* ch) {
  CHECK_MMA(h);
  if (!ch) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { *ch = h->o->design_change(xnew, false); });
}

lbmx_status lbmx_mma_design_change_device(lbmx_mma* h, const double* xnew, double* ch) {
  CHECK_MMA(h);
  if (!ch) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  return guarded([&] { *ch = h->o->design_change(xnew, true); });
}

lbmx_status lbmx_design_change_noalloc(const double* xnew, const double* xold, long long n, double* ch,
                                       void* hip_stream) {
  if (!xnew || !xold || !ch) { lbmx::g_last_error = "null argument"; return LBMX_ERR_ARG; }
  if (n < 1 || n > LBMX_NOALLOC_MAX_N) {
    lbmx::g_last_error = "design length " + std::to_string(n) + " outside 1.." + std::to_string(LBMX_NOALLOC_MAX_N);
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    static_assert(kNoAllocPartials >= 2048, "the scratch must hold mma_blocks() partials");
    if (lbmx::mma_blocks(n) > kNoAllocPartials) throw Error{LBMX_ERR_ARG, "too many reduction blocks"};
    double* part = scratch_on_current_device();
    HIP_OK(lbmx::launch_mma_change(xnew, xold, n, part, ch, static_cast<hipStream_t>(hip_stream)));
  });
}

lbmx_status lbmx_mma_get_state(lbmx_mma* h, double* L, double* U, double* lam, int* newton_iters, double* ms) {
  CHECK_MMA(h);
  return guarded([&] {
    lbmx::Mma& o = *h->o;
    if (L) HIP_OK(hipMemcpy(L, o.L.p, (size_t)o.n * 8, hipMemcpyDeviceToHost));
    if (U) HIP_OK(hipMemcpy(U, o.U.p, (size_t)o.n * 8, hipMemcpyDeviceToHost));
    if (lam)
      for (int j = 0; j < o.m; ++j) lam[j] = o.lam[j];
    if (newton_iters) *newton_iters = o.newton_iters;
    if (ms) *ms = o.last_ms;
  });
}

}  // extern "C"
