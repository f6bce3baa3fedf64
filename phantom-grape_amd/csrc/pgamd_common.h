// Shared host/device helpers for the MI355X (gfx950) Phantom-GRAPE kernels.
//
// Error behaviour follows the reference libraries: API misuse (too many
// i-particles for the pipelines, j-memory overflow) is fatal there through
// assert() (PG5/newton/libpg5/phantom_g5.c:77,105,121,246,
// PG5/table/libpg5/phantom_g5.c:88).  Here the same conditions print a
// message naming the call and abort, unless the caller switched the library to
// "record" mode with pgamd_set_abort_on_error(0), in which case the call does
// nothing and the error is kept for pgamd_last_error().  HIP runtime failures
// always abort: there is no CPU fallback.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <mutex>

#define PGAMD_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      std::fprintf(stderr, "pgamd: HIP error '%s' (%d) at %s:%d in %s\n",      \
                   hipGetErrorString(_e), (int)_e, __FILE__, __LINE__, #expr); \
      std::abort();                                                            \
    }                                                                          \
  } while (0)

namespace pgamd {

// Error state shared by every entry point of one library.
struct ErrorState {
  int abort_on_error = 1;
  int code = 0;
  char msg[512] = {0};
};
ErrorState& error_state();

// Report API misuse.  Returns true when the caller must return immediately.
bool api_error(int code, const char* fmt, ...);

// Select and check the GPU once per process.  Aborts with a clear message if
// no HIP device is visible: the library has no CPU path.
void ensure_device();

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t want) {
    if (want <= bytes) return;
    if (ptr) PGAMD_HIP_CHECK(hipFree(ptr));
    size_t b = want < 4096 ? 4096 : want;
    PGAMD_HIP_CHECK(hipMalloc(&ptr, b));
    bytes = b;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
  void release() {
    if (ptr) PGAMD_HIP_CHECK(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
  }
};

// Grow-only pinned host buffer (allocated, never registered).
struct HostBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t want, bool preserve = false) {
    if (want <= bytes) return;
    size_t b = want < 4096 ? 4096 : want;
    b = b + b / 2;
    void* p = nullptr;
    PGAMD_HIP_CHECK(hipHostMalloc(&p, b, hipHostMallocDefault));
    if (preserve && ptr) std::memcpy(p, ptr, bytes);
    if (ptr) PGAMD_HIP_CHECK(hipHostFree(ptr));
    ptr = p;
    bytes = b;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
  void release() {
    if (ptr) PGAMD_HIP_CHECK(hipHostFree(ptr));
    ptr = nullptr;
    bytes = 0;
  }
};

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Number of compute units of the current device (256 on MI355X).
int device_cu_count();

}  // namespace pgamd
