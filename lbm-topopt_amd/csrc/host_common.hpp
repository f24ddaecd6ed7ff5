// Host-side plumbing shared by the translation units behind the C ABI: the error type every
// internal function throws, the per-thread last-error message, device buffers and the wrapper
// that turns exceptions into lbmx_status codes.
#pragma once
#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

#include "lbmx.h"

namespace lbmx {

inline thread_local std::string g_last_error;

struct Error {
  int code;
  std::string msg;
};

#define HIP_OK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw ::lbmx::Error{LBMX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

struct DeviceBuffer {
  void* p = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t b) {
    if (bytes >= b && p) return;
    reset();
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error{LBMX_ERR_MEM, "hipMalloc of " + std::to_string(b) + " bytes failed: " + hipGetErrorString(e)};
    }
    bytes = b;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// A pair of timing events, destroyed on every exit path.
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EventPair() {
    HIP_OK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      throw Error{LBMX_ERR_DEVICE, "hipEventCreate failed"};
    }
  }
  EventPair(const EventPair&) = delete;
  EventPair& operator=(const EventPair&) = delete;
  ~EventPair() {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
};

// Runs f, mapping the exceptions of the engine to status codes and the last-error message.
template <class F>
inline lbmx_status guarded(F&& f) {
  try {
    f();
    return LBMX_OK;
  } catch (const Error& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return LBMX_ERR_MEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return LBMX_ERR_STATE;
  }
}

// Access to a solver handle's device fields for the design pipeline (implemented with the
// engine). All of them throw Error.
hipStream_t solver_stream(lbmx_solver* s);
int solver_lattice(lbmx_solver* s);
long long solver_nodes(lbmx_solver* s, int dims[3]);
// Sets the solver's design from a device FP64 array in natural order, on the solver's stream.
void solver_set_design_device(lbmx_solver* s, const double* dev);
// The solver's dJ/dgamma of its last gradient (device, FP64, natural order).
const double* solver_sensitivity_device(lbmx_solver* s);

}  // namespace lbmx
