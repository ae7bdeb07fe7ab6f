// common.hpp -- error handling, logging and the per-device HIP context of
// libspined.  Everything here is host code shared by the C-ABI objects and
// the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
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

// Pool of large device blocks (at least kPoolMinBytes).  The eigensolver's
// vectors (6.9 GB each at D = 8.6e8, dozens per solve) and the entry blocks
// of dropped gather caches go back to the pool instead of to hipFree, which
// unmaps them page by page, and the next allocation of the same size on the
// device takes them again.  Every library kernel runs on the device's
// library stream, so a reused block is only touched after the work that
// used it before.  An allocation that the driver refuses drops the gather
// caches not in use (into the pool), returns the pooled blocks of the
// device and retries once.
constexpr size_t kPoolMinBytes = size_t(256) << 20;
void *device_alloc(size_t bytes, bool pooled); // throws Error(LS_OUT_OF_MEMORY)
void device_free(void *p, size_t bytes, bool pooled, int device);
void pool_release_device();                    // hipFree the current device's pooled blocks
size_t pool_bytes(size_t bytes);               // pooled bytes in blocks of this size (current device)

// Device buffer with RAII release (any thread).  allocate_pooled(count,
// bytes) takes a block of max(count * sizeof(T), bytes) bytes from the pool.
template <class T> struct DeviceBuffer {
  T *ptr = nullptr;
  size_t n = 0;
  size_t cap = 0;      // bytes of the block
  bool pooled = false;
  int dev = 0;         // device of a pooled block
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t count) { allocate(count); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept
      : ptr(o.ptr), n(o.n), cap(o.cap), pooled(o.pooled), dev(o.dev) {
    o.ptr = nullptr;
    o.n = 0;
    o.cap = 0;
  }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      n = o.n;
      cap = o.cap;
      pooled = o.pooled;
      dev = o.dev;
      o.ptr = nullptr;
      o.n = 0;
      o.cap = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }
  void allocate(size_t count) { take(count, 0, false); }
  void allocate_pooled(size_t count, size_t bytes = 0) { take(count, bytes, true); }
  void release() {
    if (ptr) device_free(ptr, cap, pooled, dev);
    ptr = nullptr;
    n = 0;
    cap = 0;
  }
  size_t bytes() const { return n * sizeof(T); }

private:
  void take(size_t count, size_t bytes, bool pool) {
    release();
    const size_t b = std::max(count * sizeof(T), bytes);
    if (b) {
      pooled = pool && b >= kPoolMinBytes;
      if (pooled) dev = current_device();
      ptr = static_cast<T *>(device_alloc(b, pooled));
      cap = b;
    }
    n = count;
  }
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
