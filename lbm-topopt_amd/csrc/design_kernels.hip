// Design pipeline kernels: design-box mapping, cone density filter (forward and transposed),
// tanh projection and its sensitivity scaling. All fields are FP64 in natural order
// (z*ny + y)*nx + x on the full grid.
//
// The cone filter sums its taps in ascending (dz, dy, dx) order with separately rounded
// multiplies and adds. That is the order in which a compressed-row MatMult visits the columns
// of a row, and the order in which MatMultTranspose accumulates into an entry, so the results
// equal a plain row loop (and a scatter-form transpose) bit for bit.
#include <hip/hip_runtime.h>

#include "lbmx_launch.h"

namespace lbmx {

namespace {

constexpr int kConeTX = 64;   // outputs along x per workgroup (one wave wide)
constexpr int kConeWY = 4;    // waves per workgroup, stacked along y
constexpr int kConeKY = 4;    // outputs per thread along y (four named sums in k_cone)
constexpr int kConeTY = kConeWY * kConeKY;  // outputs along y per workgroup
constexpr int kMaxR = 16;

__device__ inline bool in_box(int x, int y, int z, const int* b) {
  return x >= b[0] && x <= b[1] && y >= b[2] && y <= b[3] && z >= b[4] && z <= b[5];
}

// LDS bytes of one workgroup: the input region of one plane and that plane's cone weights
inline size_t cone_lds_bytes(int R) {
  const size_t W = kConeTX + 2 * R, H = kConeTY + 2 * R, S = 2 * R + 1;
  return (W * H + S * S) * sizeof(double);
}

// One workgroup: a 64 x 16 tile of outputs in one z plane. For each in-grid plane z + dz the
// (64 + 2R) x (16 + 2R) input region and the plane's (2R + 1)^2 weights are staged in LDS;
// then for each tap row (dy) and tap (dx) one weight is read and applied to the thread's four
// outputs (four independent sums, each in ascending (dz, dy, dx) order).
__global__ __launch_bounds__(256) void k_cone(ConeArgs a) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  const int R = a.R;
  const int W = kConeTX + 2 * R;
  const int H = kConeTY + 2 * R;
  const int S = 2 * R + 1;
  double* tile = lds;
  double* wl = lds + W * H;
  const int tx = threadIdx.x % kConeTX;
  const int ty = threadIdx.x / kConeTX;
  const int x0 = blockIdx.x * kConeTX;
  const int y0 = blockIdx.y * kConeTY;
  const int z = blockIdx.z;
  const int x = x0 + tx;
  const size_t plane = (size_t)a.nx * a.ny;

  double acc[kConeKY];
  double own[kConeKY];  // transposed: this output's own input (identity rows outside the box)
  bool mine[kConeKY];   // output inside the filter box
#pragma unroll
  for (int k = 0; k < kConeKY; ++k) {
    const int y = y0 + ty + kConeWY * k;
    acc[k] = 0.0;
    const bool valid = x < a.nx && y < a.ny;
    mine[k] = valid && in_box(x, y, z, a.box);
    own[k] = valid ? a.in[(size_t)z * plane + (size_t)y * a.nx + x] : 0.0;
  }

  for (int dz = -R; dz <= R; ++dz) {
    const int zz = z + dz;
    if (zz < 0 || zz >= a.nz) continue;  // uniform over the workgroup
    __syncthreads();
    for (int idx = threadIdx.x; idx < W * H; idx += blockDim.x) {
      const int lx = idx % W;
      const int ly = idx / W;
      const int gx = x0 - R + lx;
      const int gy = y0 - R + ly;
      double v = 0.0;
      if (gx >= 0 && gx < a.nx && gy >= 0 && gy < a.ny) {
        // transposed: only rows inside the box carry cone weights
        if (!a.transposed || in_box(gx, gy, zz, a.box)) v = a.in[(size_t)zz * plane + (size_t)gy * a.nx + gx];
      }
      tile[ly * W + lx] = v;
    }
    const double* wz = a.wt + (size_t)(dz + R) * S * S;
    for (int idx = threadIdx.x; idx < S * S; idx += blockDim.x) wl[idx] = wz[idx];
    __syncthreads();
    const int* rxz = a.rx + (dz + R) * S;
    for (int dy = -R; dy <= R; ++dy) {
      const int r = rxz[dy + R];
      if (r < 0) continue;
      const double* w = wl + (dy + R) * S + R;  // w[dx], dx in [-R, R]
      const double* row = tile + (ty + dy + R) * W + tx + R;
      constexpr int kStep = kConeWY;           // rows between this thread's outputs
      const double* r1 = row + kStep * W;
      const double* r2 = row + 2 * kStep * W;
      const double* r3 = row + 3 * kStep * W;
      double s0 = acc[0], s1 = acc[1], s2 = acc[2], s3 = acc[3];
      // taps [lo, hi) of this row, two at a time with all loads issued first
      auto taps = [&](int lo, int hi) {
        int dx = lo;
        for (; dx + 1 < hi; dx += 2) {
          const double wa = w[dx], wb = w[dx + 1];
          const double a0 = row[dx], b0 = row[dx + 1], a1 = r1[dx], b1 = r1[dx + 1];
          const double a2 = r2[dx], b2 = r2[dx + 1], a3 = r3[dx], b3 = r3[dx + 1];
          s0 = s0 + wa * a0;
          s1 = s1 + wa * a1;
          s2 = s2 + wa * a2;
          s3 = s3 + wa * a3;
          s0 = s0 + wb * b0;
          s1 = s1 + wb * b1;
          s2 = s2 + wb * b2;
          s3 = s3 + wb * b3;
        }
        if (dx < hi) {
          const double wa = w[dx];
          const double a0 = row[dx], a1 = r1[dx], a2 = r2[dx], a3 = r3[dx];
          s0 = s0 + wa * a0;
          s1 = s1 + wa * a1;
          s2 = s2 + wa * a2;
          s3 = s3 + wa * a3;
        }
      };
      if (!(dz == 0 && dy == 0)) {
        taps(-r, r + 1);
      } else {
        taps(-r, 0);
        // the diagonal entry: cone weight for a row in the box, 1 for an identity row
        const double wv = w[0];
        if (a.transposed) {
          s0 = s0 + (mine[0] ? wv : 1.0) * own[0];
          s1 = s1 + (mine[1] ? wv : 1.0) * own[1];
          s2 = s2 + (mine[2] ? wv : 1.0) * own[2];
          s3 = s3 + (mine[3] ? wv : 1.0) * own[3];
        } else {
          s0 = s0 + wv * row[0];
          s1 = s1 + wv * r1[0];
          s2 = s2 + wv * r2[0];
          s3 = s3 + wv * r3[0];
        }
        taps(1, r + 1);
      }
      acc[0] = s0;
      acc[1] = s1;
      acc[2] = s2;
      acc[3] = s3;
    }
  }

#pragma unroll
  for (int k = 0; k < kConeKY; ++k) {
    const int y = y0 + ty + kConeWY * k;
    if (x >= a.nx || y >= a.ny) continue;
    const size_t i = (size_t)z * plane + (size_t)y * a.nx + x;
    double v;
    if (a.transposed) v = acc[k];
    else if (!mine[k]) v = own[k];
    else v = a.den ? acc[k] / a.den[i] : acc[k];
    a.out[i] = v;
  }
}

struct Grid3 {
  int nx, ny, nz;
  __device__ bool locate(long long i, int& x, int& y, int& z) const {
    if (i >= (long long)nx * ny * nz) return false;
    x = (int)(i % nx);
    const long long r = i / nx;
    y = (int)(r % ny);
    z = (int)(r / ny);
    return true;
  }
};

struct Box6 {
  int b[6];
};

__global__ __launch_bounds__(256) void k_map_to_full(double* out, const double* design, const double* background,
                                                      Grid3 g, Box6 bx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int x, y, z;
  if (!g.locate(i, x, y, z)) return;
  const int* b = bx.b;
  if (in_box(x, y, z, b)) {
    const int nxd = b[1] - b[0] + 1, nyd = b[3] - b[2] + 1;
    out[i] = design[((size_t)(z - b[4]) * nyd + (y - b[2])) * nxd + (x - b[0])];
  } else {
    out[i] = background[i];
  }
}

__global__ __launch_bounds__(256) void k_extract_box(double* design, const double* full, Grid3 g, Box6 bx) {
  const int* b = bx.b;
  const int nxd = b[1] - b[0] + 1, nyd = b[3] - b[2] + 1, nzd = b[5] - b[4] + 1;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)nxd * nyd * nzd) return;
  const int xd = (int)(k % nxd);
  const long long r = k / nxd;
  const int yd = (int)(r % nyd);
  const int zd = (int)(r / nyd);
  design[k] = full[((size_t)(zd + b[4]) * g.ny + (yd + b[2])) * g.nx + (xd + b[0])];
}

// threshold of node i: the stage constant, or a per-node value with its own tanh terms
__device__ inline void node_eta(const double* etaf, long long i, double beta, double& eta, double& tb,
                                double& tsum) {
#pragma clang fp contract(off)
  if (etaf) {
    eta = etaf[i];
    tb = tanh(beta * eta);
    tsum = tb + tanh(beta * (1.0 - eta));
  }
}

__global__ __launch_bounds__(256) void k_project(double* out, const double* xin, Grid3 g, Box6 bx, double beta,
                                                  double eta, double tb, double tsum, const double* etaf) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int x, y, z;
  if (!g.locate(i, x, y, z)) return;
  const double v = xin[i];
  if (in_box(x, y, z, bx.b)) {
    node_eta(etaf, i, beta, eta, tb, tsum);
    out[i] = (tb + tanh(beta * (v - eta))) / tsum;
  } else {
    out[i] = v;
  }
}

__global__ __launch_bounds__(256) void k_sens_scale(double* out, const double* sens, const double* xin,
                                                     const double* den, Grid3 g, Box6 bx, int project, double beta,
                                                     double eta, double tsum, const double* etaf) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int x, y, z;
  if (!g.locate(i, x, y, z)) return;
  double s = sens[i];
  if (project && in_box(x, y, z, bx.b)) {
    double tb = 0.0;
    node_eta(etaf, i, beta, eta, tb, tsum);
    const double t = tanh(beta * (xin[i] - eta));
    s = s * (((1.0 - t * t) * beta) / tsum);
  }
  out[i] = s / den[i];
}

__global__ __launch_bounds__(256) void k_convert_f64_f32(float* out, const double* in, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

inline Box6 box6(const int b[6]) {
  Box6 r;
  for (int k = 0; k < 6; ++k) r.b[k] = b[k];
  return r;
}

}  // namespace

hipError_t launch_cone(const ConeArgs& a, hipStream_t s) {
  if (a.R < 1 || a.R > kMaxR || a.nx < 1 || a.ny < 1 || a.nz < 1) return hipErrorInvalidValue;
  dim3 grid((a.nx + kConeTX - 1) / kConeTX, (a.ny + kConeTY - 1) / kConeTY, a.nz);
  hipLaunchKernelGGL(k_cone, grid, dim3(kConeTX * kConeWY), cone_lds_bytes(a.R), s, a);
  return hipGetLastError();
}

hipError_t launch_map_to_full(double* out, const double* design, const double* background, int nx, int ny, int nz,
                              const int box[6], hipStream_t s) {
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_map_to_full, dim3(blocks_for(n)), dim3(256), 0, s, out, design, background,
                     Grid3{nx, ny, nz}, box6(box));
  return hipGetLastError();
}

hipError_t launch_extract_box(double* design, const double* full, int nx, int ny, int nz, const int box[6],
                              hipStream_t s) {
  const long long n = (long long)(box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1);
  hipLaunchKernelGGL(k_extract_box, dim3(blocks_for(n)), dim3(256), 0, s, design, full, Grid3{nx, ny, nz},
                     box6(box));
  return hipGetLastError();
}

hipError_t launch_project(double* out, const double* x, int nx, int ny, int nz, const int box[6], double beta,
                          double eta, double tb, double tsum, const double* etaf, hipStream_t s) {
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_project, dim3(blocks_for(n)), dim3(256), 0, s, out, x, Grid3{nx, ny, nz}, box6(box), beta,
                     eta, tb, tsum, etaf);
  return hipGetLastError();
}

hipError_t launch_sens_scale(double* out, const double* sens, const double* x, const double* den, int nx, int ny,
                             int nz, const int box[6], int project, double beta, double eta, double tsum,
                             const double* etaf, hipStream_t s) {
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(k_sens_scale, dim3(blocks_for(n)), dim3(256), 0, s, out, sens, x, den, Grid3{nx, ny, nz},
                     box6(box), project, beta, eta, tsum, etaf);
  return hipGetLastError();
}

hipError_t convert_f64_f32(float* out, const double* in, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_convert_f64_f32, dim3(blocks_for(n)), dim3(256), 0, s, out, in, n);
  return hipGetLastError();
}

}  // namespace lbmx
