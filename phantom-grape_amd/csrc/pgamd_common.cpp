// Shared host helpers: error state, device selection.  Compiled into each of
// libpg5.so, libpg5_table.so and libpg6.so with hidden visibility, so every
// library keeps its own copy.
#include "pgamd_common.h"
#include "pgamd_api.h"

#include <cstdarg>

namespace pgamd {

ErrorState& error_state() {
  static ErrorState st;
  return st;
}

// Every read or write of the shared error state holds this lock: the MC
// entry points are called concurrently from one thread per devid.
static std::mutex g_err_mutex;

bool api_error(int code, const char* fmt, ...) {
  ErrorState& st = error_state();
  char msg[sizeof(st.msg)];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(msg, sizeof(msg), fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(g_err_mutex);
  std::memcpy(st.msg, msg, sizeof(msg));
  st.code = code;
  if (st.abort_on_error) {
    std::fprintf(stderr, "pgamd: %s\n", st.msg);
    std::abort();
  }
  return true;
}

static std::once_flag g_dev_once;
static int g_cu_count = 0;

void ensure_device() {
  std::call_once(g_dev_once, [] {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
      std::fprintf(stderr,
                   "pgamd: no HIP device is visible (%s); this library runs "
                   "only on a GPU and has no CPU path\n",
                   e == hipSuccess ? "device count 0" : hipGetErrorString(e));
      std::abort();
    }
    int dev = 0;
    PGAMD_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PGAMD_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    g_cu_count = prop.multiProcessorCount;
  });
}

int device_cu_count() {
  ensure_device();
  return g_cu_count;
}

}  // namespace pgamd

PGAMD_API void pgamd_set_abort_on_error(int flag) {
  std::lock_guard<std::mutex> lk(pgamd::g_err_mutex);
  pgamd::error_state().abort_on_error = flag;
}

PGAMD_API int pgamd_last_error(char* buf, int buflen) {
  std::lock_guard<std::mutex> lk(pgamd::g_err_mutex);
  pgamd::ErrorState& st = pgamd::error_state();
  if (buf && buflen > 0) {
    std::strncpy(buf, st.msg, (size_t)buflen - 1);
    buf[buflen - 1] = 0;
  }
  int c = st.code;
  st.code = 0;
  st.msg[0] = 0;
  return c;
}
