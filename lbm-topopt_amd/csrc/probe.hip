// Streaming copy used as the measured HBM ceiling and as the calibration kernel for the
// FETCH_SIZE / WRITE_SIZE counters (known bytes read and written per launch).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "host_common.hpp"
#include "lbmx.h"

namespace lbmx {

// one 32-byte element per thread over an exact grid: on MI355X this reaches a higher copy
// rate than grid-stride loops, non-temporal accesses or several loads per thread
__global__ __launch_bounds__(256) void k_stream_copy(const double4* __restrict__ in, double4* __restrict__ out,
                                                      long long n4) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) out[i] = in[i];
}

// The access pattern of a Q-direction sweep without its arithmetic: Q planes read at the
// same node offset, Q planes written with non-temporal stores as the sweeps write their
// aligned rows (FP64, one node per thread).
template <int Q>
__global__ __launch_bounds__(256) void k_plane_copy(const double* __restrict__ in, double* __restrict__ out,
                                                    long long plane, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = in[q * plane + i];
#pragma unroll
  for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(v[q], out + q * plane + i);
}

// The same for 16-bit planes as the 16-bit forms' two-node forward sweep moves them: two nodes
// per thread, each plane's pair read and written as one 32-bit word.
template <int Q>
__global__ __launch_bounds__(256) void k_plane_copy16(const unsigned int* __restrict__ in, unsigned int* __restrict__ out,
                                                      long long plane_words, long long nwords) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwords) return;
  unsigned int v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = __builtin_nontemporal_load(in + q * plane_words + i);
#pragma unroll
  for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(v[q], out + q * plane_words + i);
}

}  // namespace lbmx

extern "C" lbmx_status lbmx_stream_copy_planes16(int planes, size_t nodes, int iters, double* gbs) {
  if (!gbs || iters < 1 || nodes < 512 || (planes != 9 && planes != 14 && planes != 19)) {
    lbmx::g_last_error = "stream copy planes16: needs gbs, iters >= 1, nodes >= 512 and planes 9, 14 or 19";
    return LBMX_ERR_ARG;
  }
  const long long nw = (long long)nodes / 2;           // 32-bit words (two nodes each)
  const long long plane = (nw + 63) / 64 * 64;         // 256-byte aligned planes
  const size_t bytes = (size_t)planes * plane * 4;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) {
    lbmx::g_last_error = "device allocation failed";
    return LBMX_ERR_MEM;
  }
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    lbmx::g_last_error = "device allocation failed";
    return LBMX_ERR_MEM;
  }
  (void)hipMemset(a, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const unsigned blocks = (unsigned)((nw + 255) / 256);
  auto launch = [&]() {
    const unsigned int* in = static_cast<const unsigned int*>(a);
    unsigned int* out = static_cast<unsigned int*>(b);
    if (planes == 19)
      hipLaunchKernelGGL(lbmx::k_plane_copy16<19>, dim3(blocks), dim3(256), 0, nullptr, in, out, plane, nw);
    else if (planes == 14)
      hipLaunchKernelGGL(lbmx::k_plane_copy16<14>, dim3(blocks), dim3(256), 0, nullptr, in, out, plane, nw);
    else
      hipLaunchKernelGGL(lbmx::k_plane_copy16<9>, dim3(blocks), dim3(256), 0, nullptr, in, out, plane, nw);
  };
  launch();
  (void)hipEventRecord(e0, nullptr);
  for (int k = 0; k < iters; ++k) launch();
  (void)hipEventRecord(e1, nullptr);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) {
    lbmx::g_last_error = hipGetErrorString(err);
    return LBMX_ERR_DEVICE;
  }
  *gbs = 2.0 * planes * nw * 4.0 * iters / (ms * 1e-3) / 1e9;
  return LBMX_OK;
}

extern "C" lbmx_status lbmx_stream_copy_planes(int planes, size_t nodes, int iters, double* gbs) {
  if (!gbs || iters < 1 || nodes < 256 || (planes != 9 && planes != 14 && planes != 19)) {
    lbmx::g_last_error = "stream copy planes: needs gbs, iters >= 1, nodes >= 256 and planes 9, 14 or 19";
    return LBMX_ERR_ARG;
  }
  const long long n = (long long)nodes;
  const long long plane = (n + 31) / 32 * 32;  // 256-byte aligned planes
  const size_t bytes = (size_t)planes * plane * 8;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) {
    lbmx::g_last_error = "device allocation failed";
    return LBMX_ERR_MEM;
  }
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    lbmx::g_last_error = "device allocation failed";
    return LBMX_ERR_MEM;
  }
  (void)hipMemset(a, 0, bytes);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  auto launch = [&]() {
    if (planes == 19)
      hipLaunchKernelGGL(lbmx::k_plane_copy<19>, dim3(blocks), dim3(256), 0, nullptr, (const double*)a, (double*)b,
                         plane, n);
    else if (planes == 14)
      hipLaunchKernelGGL(lbmx::k_plane_copy<14>, dim3(blocks), dim3(256), 0, nullptr, (const double*)a, (double*)b,
                         plane, n);
    else
      hipLaunchKernelGGL(lbmx::k_plane_copy<9>, dim3(blocks), dim3(256), 0, nullptr, (const double*)a, (double*)b,
                         plane, n);
  };
  launch();
  (void)hipEventRecord(e0, nullptr);
  for (int k = 0; k < iters; ++k) launch();
  (void)hipEventRecord(e1, nullptr);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) {
    lbmx::g_last_error = hipGetErrorString(err);
    return LBMX_ERR_DEVICE;
  }
  *gbs = 2.0 * planes * n * 8.0 * iters / (ms * 1e-3) / 1e9;
  return LBMX_OK;
}

extern "C" lbmx_status lbmx_stream_copy(size_t bytes, int iters, double* gbs) {
  if (!gbs || iters < 1 || bytes < 1024) {
    lbmx::g_last_error = "stream copy: needs gbs, iters >= 1 and bytes >= 1024";
    return LBMX_ERR_ARG;
  }
  const long long n4 = (long long)(bytes / 32);
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n4 * 32) != hipSuccess) {
    lbmx::g_last_error = "device allocation failed";
    return LBMX_ERR_MEM;
  }
  if (hipMalloc(&b, n4 * 32) != hipSuccess) {
    (void)hipFree(a);
    lbmx::g_last_error = "device allocation failed";
    return LBMX_ERR_MEM;
  }
  (void)hipMemset(a, 0, n4 * 32);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const unsigned blocks = (unsigned)((n4 + 255) / 256);
  hipLaunchKernelGGL(lbmx::k_stream_copy, dim3(blocks), dim3(256), 0, nullptr, (const double4*)a, (double4*)b, n4);
  (void)hipEventRecord(e0, nullptr);
  for (int k = 0; k < iters; ++k)
    hipLaunchKernelGGL(lbmx::k_stream_copy, dim3(blocks), dim3(256), 0, nullptr, (const double4*)a, (double4*)b, n4);
  (void)hipEventRecord(e1, nullptr);
  hipError_t err = hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (err != hipSuccess) {
    lbmx::g_last_error = hipGetErrorString(err);
    return LBMX_ERR_DEVICE;
  }
  *gbs = 2.0 * n4 * 32 * iters / (ms * 1e-3) / 1e9;
  return LBMX_OK;
}

extern "C" lbmx_status lbmx_device_count(int* count) {
  if (!count) return LBMX_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return LBMX_OK;
}

namespace lbmx {

// Zero fill by a kernel (usable inside captured graphs, where memset nodes are avoided).
__global__ __launch_bounds__(256) void k_fill_zero(unsigned int* __restrict__ p, long long nwords) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) p[i] = 0u;
}

hipError_t fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if (bytes % 4 != 0) return hipErrorInvalidValue;
  const long long nwords = (long long)(bytes / 4);
  const long long want = (nwords + 255) / 256;
  const unsigned blocks = (unsigned)std::min<long long>(want, 65536);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, s, static_cast<unsigned int*>(p), nwords);
  return hipGetLastError();
}

}  // namespace lbmx
