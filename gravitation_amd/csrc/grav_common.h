// Shared helpers of libgravhip.so: error channel, argument checks, launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gravitation_hip.h"

namespace grav {

constexpr int kWave = 64;           // CDNA wavefront width
constexpr int64_t kMaxBodies = int64_t(1) << 29;  // keeps 3*n + 64*8 (SoA offsets) < 2^31

// Records `fmt` as the calling thread's last error and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Clears the calling thread's last error.
void clear_error();

// Checks the launch that was just enqueued; returns GRAV_OK or GRAV_EHIP.
int check_launch(const char *what);

inline bool aligned(const void *p, uintptr_t a) {
  return (reinterpret_cast<uintptr_t>(p) % a) == 0;
}

}  // namespace grav

#define GRAV_REQUIRE(cond, ...)                         \
  do {                                                  \
    if (!(cond)) return ::grav::fail(GRAV_EINVAL, __VA_ARGS__); \
  } while (0)

#define GRAV_HIP_CHECK(expr)                                                    \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return ::grav::fail(GRAV_EHIP, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)
