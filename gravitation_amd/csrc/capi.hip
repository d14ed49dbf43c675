// Library-wide entry points of libgravhip.so: version, error channel, device count.
#include <cstdarg>
#include <cstdio>

#include "grav_common.h"

namespace grav {

namespace {
thread_local char g_last_error[1024] = "";
}

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_last_error[0] = '\0'; }

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(GRAV_EHIP, "%s launch: %s", what, hipGetErrorString(e));
  return GRAV_OK;
}

}  // namespace grav

extern "C" {

int grav_version(void) { return 100; }  // 0.1.0

const char *grav_last_error(void) { return grav::g_last_error; }

int grav_device_count(int *count) {
  GRAV_REQUIRE(count != nullptr, "grav_device_count: count is NULL");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  (void)hipGetLastError();  // a missing device is not an error for this query
  *count = c;
  return GRAV_OK;
}

}  // extern "C"
