// common.hpp -- error handling, logging and the per-device HIP context of
// libspined.  Everything here is host code shared by the C-ABI objects and
// the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/lattice_symmetries.h"

namespace spined {

// An error that maps onto an ls_error_code at the C boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char *expr, const char *file, int line);

#define SPINED_HIP(expr)                                                       \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) ::spined::throw_hip(_e, #expr, __FILE__, __LINE__);   \
  } while (0)

#define SPINED_CHECK_LAUNCH() SPINED_HIP(hipGetLastError())

bool logging_enabled();
void log_line(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Library-owned stream of the current device (created on first use, default
// priority, non-blocking w.r.t. the null stream).
hipStream_t stream();
int current_device();

// Device buffer with RAII release (any thread).
template <class T> struct DeviceBuffer {
  T *ptr = nullptr;
  size_t n = 0;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t count) { allocate(count); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), n(o.n) {
    o.ptr = nullptr;
    o.n = 0;
  }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      n = o.n;
      o.ptr = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }
  void allocate(size_t count) {
    release();
    if (count) {
      hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T));
      if (e != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        throw Error(LS_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                          " bytes failed: " + hipGetErrorString(e));
      }
    }
    n = count;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Converts exceptions into status codes at the C boundary.
template <class F> int guarded(F &&f) {
  try {
    f();
    return LS_SUCCESS;
  } catch (const Error &e) {
    if (logging_enabled()) log_line("error %d: %s", e.code, e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    return LS_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    if (logging_enabled()) log_line("internal error: %s", e.what());
    return LS_INTERNAL_ERROR;
  }
}

} // namespace spined
