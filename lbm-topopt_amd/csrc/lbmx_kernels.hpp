// Sweep kernels. Layout in HBM: structure-of-arrays, plane q at offset q*N, node
// n = (z*ny + y)*nx + x (x fastest), so every wave reads/writes 64 consecutive nodes of one
// population plane. Workgroups are 64 (x) x 4 (y) threads.
//
//  k_fwd : pull-stream from the post-collision state p_{t-1} (out-of-domain sources take the
//          node's own opposite population: mid-way bounce-back), boundary pieces in insertion
//          order, functional of the new state, collision, store p_t.
//  k_adj : one reverse step: recompute s_t = B_t S p_{t-1} (pull + boundaries), reverse
//          collision on nu_t, functional source, transposed boundaries in reverse order, then
//          the transposed stream as a push into nu_{t-1}; dJ/dgamma accumulated in FP64.
#pragma once
#include "lbmx_device.hpp"

namespace lbmx {

template <class L, class P>
__device__ __forceinline__ void pull_node(const typename P::S* __restrict__ src, long long N, int nx,
                                          int ny, int nz, int x, int y, int z, long long n,
                                          typename P::T (&f)[L::Q]) {
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = L::ex[q], cy = L::ey[q], cz = L::ez[q];
    if constexpr (cx == 0 && cy == 0 && cz == 0) {
      f[q] = P::template dec<L, q>(src[(long long)q * N + n]);
    } else {
      bool inside = true;
      if constexpr (cx == 1) inside = inside && (x > 0);
      if constexpr (cx == -1) inside = inside && (x < nx - 1);
      if constexpr (cy == 1) inside = inside && (y > 0);
      if constexpr (cy == -1) inside = inside && (y < ny - 1);
      if constexpr (cz == 1) inside = inside && (z > 0);
      if constexpr (cz == -1) inside = inside && (z < nz - 1);
      const long long shift = (long long)cx + (long long)cy * nx + (long long)cz * nx * ny;
      const long long off = inside ? (long long)q * N + n - shift : (long long)L::opp[q] * N + n;
      f[q] = P::template dec<L, q>(src[off]);
    }
  });
}

template <class L, class T>
__device__ __forceinline__ bool on_boundary(int x, int y, int z, int nx, int ny, int nz) {
  bool b = x == 0 || x == nx - 1 || y == 0 || y == ny - 1;
  if constexpr (L::D == 3) b = b || z == 0 || z == nz - 1;
  return b;
}

template <class L, class T>
__device__ __forceinline__ void apply_bcs(T (&f)[L::Q], const DevBC* __restrict__ bcs, int nbc, int t,
                                          int x, int y, int z) {
  for (int b = 0; b < nbc; ++b) {
    const DevBC& e = bcs[b];
    if (in_box(e.box, x, y, z)) bc_node<L, T>(f, e, t, x, y, z);
  }
}

template <class L, class T>
__device__ __forceinline__ void apply_bcs_adj(T (&l)[L::Q], const DevBC* __restrict__ bcs, int nbc,
                                              int x, int y, int z) {
  for (int b = nbc - 1; b >= 0; --b) {
    const DevBC& e = bcs[b];
    if (in_box(e.box, x, y, z)) bc_node_adj<L, T>(l, e);
  }
}

__device__ __forceinline__ bool block_meets(const int* bb) {
  int x0 = blockIdx.x * kTileX, x1 = x0 + kTileX - 1;
  int y0 = blockIdx.y * kTileY, y1 = y0 + kTileY - 1;
  int z = blockIdx.z;
  return !(x1 < bb[0] || x0 > bb[1] || y1 < bb[2] || y0 > bb[3] || z < bb[4] || z > bb[5]);
}

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double part[kTileY];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (threadIdx.x == 0) part[threadIdx.y] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0 && threadIdx.y == 0)
    for (int w = 0; w < kTileY; ++w) s += part[w];
  return s;
}

template <class L, int COLL, class T>
__device__ __forceinline__ T node_objective(const ObjArgs& o, const T (&f)[L::Q], int x, int y, int z) {
  T rho, j[3];
  flow_macros<L>(f, rho, j);
  T acc = 0;
  for (int m = 0; m < o.nterm; ++m)
    if (in_box(o.terms[m].box, x, y, z))
      acc += T(o.terms[m].weight) * term_value<L, COLL, T>(o.terms[m].kind, f, rho, j);
  return acc * T(o.time_const);
}

template <class L, int COLL, class P>
__global__ __launch_bounds__(kTileX* kTileY) void k_fwd(FwdArgs a) {
  using T = typename P::T;
  using S = typename P::S;
  const int x = blockIdx.x * kTileX + threadIdx.x;
  const int y = blockIdx.y * kTileY + threadIdx.y;
  const int z = blockIdx.z;
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  const long long N = a.g.N;
  const bool active = x < nx && y < ny;
  const bool objblk = a.obj.acc != nullptr && block_meets(a.obj.bbox);
  double objv = 0.0;
  if (active) {
    const long long n = ((long long)z * ny + y) * nx + x;
    T f[L::Q];
    if (a.mode == 2) {
      const double* s = static_cast<const double*>(a.src);
      static_for<0, L::Q>([&](auto qc) { f[decltype(qc)::value] = T(s[n * L::Q + decltype(qc)::value]); });
    } else {
      pull_node<L, P>(static_cast<const S*>(a.src), N, nx, ny, nz, x, y, z, n, f);
      if (on_boundary<L, T>(x, y, z, nx, ny, nz)) apply_bcs<L, T>(f, a.bcs, a.nbc, a.t, x, y, z);
    }
    if (a.mode == 1) {
      double* out = static_cast<double*>(a.aos_out);
      static_for<0, L::Q>([&](auto qc) { out[n * L::Q + decltype(qc)::value] = double(f[decltype(qc)::value]); });
    } else {
      if (objblk && in_box(a.obj.bbox, x, y, z)) objv = double(node_objective<L, COLL, T>(a.obj, f, x, y, z));
      const T gam = T(static_cast<const typename P::G*>(a.gamma)[n]);
      T rho, j[3];
      collide<L, COLL, T>(f, gam, a.m, rho, j);
      S* dst = static_cast<S*>(a.dst);
      static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        dst[(long long)q * N + n] = P::template enc<L, q>(f[q]);
      });
    }
  }
  if (objblk && a.mode == 0) {
    double s = block_sum(objv);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      long long bid = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
      a.obj.acc[bid] += s;
    }
  }
}

// Initial state: s_0 = f_eq(rho, u[, T]) from uniform or per-node macros; writes p_0 = C(s_0)
// (IsoEquilibriumInitializationLoop2d, src/dynamics/isothermalLBLoops2d.hh:264-285).
template <class L, int COLL, class P>
__global__ __launch_bounds__(kTileX* kTileY) void k_init(FwdArgs a, const double* macros, double rho0,
                                                         double ux0, double uy0, double uz0, double T0) {
  using T = typename P::T;
  using S = typename P::S;
  const int x = blockIdx.x * kTileX + threadIdx.x;
  const int y = blockIdx.y * kTileY + threadIdx.y;
  const int z = blockIdx.z;
  const int nx = a.g.nx, ny = a.g.ny;
  if (x >= nx || y >= ny) return;
  const long long N = a.g.N;
  const long long n = ((long long)z * ny + y) * nx + x;
  double r = rho0, u[3] = {ux0, uy0, uz0}, tmp = T0;
  if (macros) {
    r = macros[5 * n]; u[0] = macros[5 * n + 1]; u[1] = macros[5 * n + 2]; u[2] = macros[5 * n + 3];
    tmp = macros[5 * n + 4];
  }
  constexpr bool inc = CollTraits<COLL>::inc;
  double fd[L::Q];
  const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  static_for<0, L::QF>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    double cu = cdot<L, q>(u);
    double poly = 3.0 * cu + 4.5 * cu * cu - 1.5 * u2;
    fd[q] = inc ? L::w[q] * (r + poly) : L::w[q] * r * (1.0 + poly);
  });
  if constexpr (L::QT > 0) {
    // ThermalMRTFeq2d (src/dynamics/thermalEquilibriumDistributions2d.hh:28-47): f = M^-1 m_eq
    const double meq[5] = {tmp, u[0] * tmp, u[1] * tmp, -2.0 * tmp, 0.0};
    static_for<0, 5>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      double s = 0;
      static_for<0, 5>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if constexpr (kB5.M[k][q] != 0) s += kB5.M[k][q] * kB5.invn2[k] * meq[k];
      });
      fd[L::QF + q] = s;
    });
  }
  if (a.aos_out) {
    double* out = static_cast<double*>(a.aos_out);
    static_for<0, L::Q>([&](auto qc) { out[n * L::Q + decltype(qc)::value] = fd[decltype(qc)::value]; });
  }
  if (a.dst) {
    T f[L::Q];
    static_for<0, L::Q>([&](auto qc) { f[decltype(qc)::value] = T(fd[decltype(qc)::value]); });
    const T gam = T(static_cast<const typename P::G*>(a.gamma)[n]);
    T rho, j[3];
    collide<L, COLL, T>(f, gam, a.m, rho, j);
    S* dst = static_cast<S*>(a.dst);
    static_for<0, L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      dst[(long long)q * N + n] = P::template enc<L, q>(f[q]);
    });
  }
}

template <class L, int COLL, class P>
__global__ __launch_bounds__(kTileX* kTileY) void k_adj(AdjArgs a) {
  using T = typename P::T;
  using S = typename P::S;
  using LS = typename P::LS;
  const int x = blockIdx.x * kTileX + threadIdx.x;
  const int y = blockIdx.y * kTileY + threadIdx.y;
  const int z = blockIdx.z;
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  if (x >= nx || y >= ny) return;
  const long long N = a.g.N;
  const long long n = ((long long)z * ny + y) * nx + x;
  T f[L::Q];
  pull_node<L, P>(static_cast<const S*>(a.fwd_src), N, nx, ny, nz, x, y, z, n, f);
  const bool bnd = on_boundary<L, T>(x, y, z, nx, ny, nz);
  if (bnd) apply_bcs<L, T>(f, a.bcs, a.nbc, a.t, x, y, z);
  T lam[L::Q];
  const LS* lin = static_cast<const LS*>(a.lam_in);
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    lam[q] = T(lin[(long long)q * N + n]);
  });
  const T gam = T(static_cast<const typename P::G*>(a.gamma)[n]);
  T dg = 0;
  reverse_collide<L, COLL, T>(f, gam, a.m, lam, dg);
  a.sens[n] += double(dg);
  if (a.obj.nterm > 0 && in_box(a.obj.bbox, x, y, z)) {
    T rho, j[3];
    flow_macros<L>(f, rho, j);
    for (int m = 0; m < a.obj.nterm; ++m)
      if (in_box(a.obj.terms[m].box, x, y, z))
        term_grad<L, T>(a.obj.terms[m].kind, f, rho, j, T(a.obj.time_const * a.obj.terms[m].weight), lam);
  }
  if (bnd) apply_bcs_adj<L, T>(lam, a.bcs, a.nbc, x, y, z);
  LS* lout = static_cast<LS*>(a.lam_out);
  static_for<0, L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = L::ex[q], cy = L::ey[q], cz = L::ez[q];
    if constexpr (cx == 0 && cy == 0 && cz == 0) {
      lout[(long long)q * N + n] = LS(lam[q]);
    } else {
      bool inside = true;
      if constexpr (cx == 1) inside = inside && (x > 0);
      if constexpr (cx == -1) inside = inside && (x < nx - 1);
      if constexpr (cy == 1) inside = inside && (y > 0);
      if constexpr (cy == -1) inside = inside && (y < ny - 1);
      if constexpr (cz == 1) inside = inside && (z > 0);
      if constexpr (cz == -1) inside = inside && (z < nz - 1);
      const long long shift = (long long)cx + (long long)cy * nx + (long long)cz * nx * ny;
      const long long off = inside ? (long long)q * N + n - shift : (long long)L::opp[q] * N + n;
      lout[off] = LS(lam[q]);
    }
  });
}

// ---- host-side launchers (one instantiation set per lattice translation unit) ----
template <class L, int COLL, class P>
hipError_t launch_fwd_t(const FwdArgs& a, hipStream_t s) {
  dim3 blk(kTileX, kTileY, 1);
  dim3 grd((a.g.nx + kTileX - 1) / kTileX, (a.g.ny + kTileY - 1) / kTileY, a.g.nz);
  hipLaunchKernelGGL((k_fwd<L, COLL, P>), grd, blk, 0, s, a);
  return hipGetLastError();
}

template <class L, int COLL, class P>
hipError_t launch_init_t(const FwdArgs& a, const double* macros, const double* u, double rho0, double T0,
                         hipStream_t s) {
  dim3 blk(kTileX, kTileY, 1);
  dim3 grd((a.g.nx + kTileX - 1) / kTileX, (a.g.ny + kTileY - 1) / kTileY, a.g.nz);
  hipLaunchKernelGGL((k_init<L, COLL, P>), grd, blk, 0, s, a, macros, rho0, u[0], u[1], u[2], T0);
  return hipGetLastError();
}

template <class L, int COLL, class P>
hipError_t launch_adj_t(const AdjArgs& a, hipStream_t s) {
  dim3 blk(kTileX, kTileY, 1);
  dim3 grd((a.g.nx + kTileX - 1) / kTileX, (a.g.ny + kTileY - 1) / kTileY, a.g.nz);
  hipLaunchKernelGGL((k_adj<L, COLL, P>), grd, blk, 0, s, a);
  return hipGetLastError();
}

}  // namespace lbmx
