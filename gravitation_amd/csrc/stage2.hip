// Stage 2: semi-implicit Euler, one body per lane, HBM-bandwidth bound.
//
// Order of operations follows the reference exactly (src/gravitation/kernel/_base_.py:58-62,
// vectorised in src/gravitation/kernel/np2.py:110-115):
//   v <- a*T + v        (the product rounded, then the sum)
//   r <- v*T + r        (with the new v)
// FP contraction is switched off in every kernel body, so each product is rounded before the
// sum (-ffp-contract=fast would fuse them into an FMA): stage 2 is then bit-identical to the
// same two NumPy operations in the same precision.
#include "grav_common.h"

namespace grav {
namespace {

constexpr int kBlock = 256;

// mirror != NULL (the *_mirror entry points): the new state is also stored straight into a
// mapped, coherent host buffer, laid out as the device state of the mode (below), so the host
// sees it when the stream has drained with no copy command.
__global__ __launch_bounds__(kBlock) void k_stage2_f32(float4 *__restrict__ posm,
                                                       float *__restrict__ vel,
                                                       const float *__restrict__ acc, int n,
                                                       float dt, float *__restrict__ mirror) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 p = posm[i];
  const float vx = acc[i] * dt + vel[i];
  const float vy = acc[n + i] * dt + vel[n + i];
  const float vz = acc[2 * n + i] * dt + vel[2 * n + i];
  p.x = vx * dt + p.x;
  p.y = vy * dt + p.y;
  p.z = vz * dt + p.z;
  vel[i] = vx;
  vel[n + i] = vy;
  vel[2 * n + i] = vz;
  posm[i] = p;
  if (mirror != nullptr) {   // [posm (n, 4) | vel (3, n)]
    reinterpret_cast<float4 *>(mirror)[i] = p;
    float *mv = mirror + 4 * static_cast<size_t>(n);
    mv[i] = vx;
    mv[n + i] = vy;
    mv[2 * n + i] = vz;
  }
}

// FP64 masters advanced with FP32 accelerations; the FP32 stage-1 copy gets x, y, z only
// (a 12-byte store, so m is neither read nor written).
__global__ __launch_bounds__(kBlock) void k_stage2_f64m(float *__restrict__ posm,
                                                        double *__restrict__ r64,
                                                        double *__restrict__ v64,
                                                        const float *__restrict__ acc, int n,
                                                        double dt, double *__restrict__ mirror) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double r[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double a = static_cast<double>(acc[c * n + i]);
    v[c] = a * dt + v64[c * n + i];
    r[c] = v[c] * dt + r64[c * n + i];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v64[c * n + i] = v[c];
    r64[c * n + i] = r[c];
  }
  float *p = posm + 4 * static_cast<size_t>(i);
  p[0] = static_cast<float>(r[0]);
  p[1] = static_cast<float>(r[1]);
  p[2] = static_cast<float>(r[2]);
  if (mirror != nullptr) {   // [r64 (3, n) | v64 (3, n)]
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      mirror[c * n + i] = r[c];
      mirror[(3 + c) * static_cast<size_t>(n) + i] = v[c];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_stage2_f64(double *__restrict__ posm64,
                                                       double *__restrict__ v64,
                                                       const double *__restrict__ acc, int n,
                                                       double dt, double *__restrict__ mirror) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double *p = posm64 + 4 * static_cast<size_t>(i);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = acc[c * n + i] * dt + v64[c * n + i];
    const double r = v * dt + p[c];
    v64[c * n + i] = v;
    p[c] = r;
    if (mirror != nullptr) {   // [posm64 (n, 4) | v64 (3, n)]
      mirror[4 * static_cast<size_t>(i) + c] = r;
      mirror[(4 + c) * static_cast<size_t>(n) + i] = v;
    }
  }
  if (mirror != nullptr) mirror[4 * static_cast<size_t>(i) + 3] = p[3];
}

int validate(const char *fn, const void *a, const void *b, const void *c, int64_t n, double dt) {
  GRAV_REQUIRE(a != nullptr && b != nullptr && c != nullptr, "%s: NULL buffer", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "%s: dt must be finite", fn);
  return GRAV_OK;
}

}  // namespace
}  // namespace grav

using namespace grav;

extern "C" {

namespace {
int stage2_f32(const char *fn, void *posm, void *vel, const void *acc, int64_t n, double dt,
               void *mirror, void *stream) {
  int rc = validate(fn, posm, vel, acc, n, dt);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(aligned(posm, 16), "%s: posm must be 16-byte aligned", fn);
  GRAV_REQUIRE(aligned(vel, 4) && aligned(acc, 4), "%s: vel/acc misaligned", fn);
  const int nn = static_cast<int>(n);
  hipLaunchKernelGGL(k_stage2_f32, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<float4 *>(posm),
                     static_cast<float *>(vel), static_cast<const float *>(acc), nn,
                     static_cast<float>(dt), static_cast<float *>(mirror));
  return check_launch("k_stage2_f32");
}

int stage2_f64m(const char *fn, void *posm, void *r64, void *v64, const void *acc, int64_t n,
                double dt, void *mirror, void *stream) {
  int rc = validate(fn, posm, r64, v64, n, dt);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(acc != nullptr, "%s: acc is NULL", fn);
  GRAV_REQUIRE(aligned(posm, 16), "%s: posm must be 16-byte aligned", fn);
  GRAV_REQUIRE(aligned(r64, 8) && aligned(v64, 8) && aligned(acc, 4),
               "%s: r64/v64/acc misaligned", fn);
  const int nn = static_cast<int>(n);
  hipLaunchKernelGGL(k_stage2_f64m, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<float *>(posm),
                     static_cast<double *>(r64), static_cast<double *>(v64),
                     static_cast<const float *>(acc), nn, dt, static_cast<double *>(mirror));
  return check_launch("k_stage2_f64m");
}

int stage2_f64(const char *fn, void *posm64, void *v64, const void *acc64, int64_t n, double dt,
               void *mirror, void *stream) {
  int rc = validate(fn, posm64, v64, acc64, n, dt);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(aligned(posm64, 32), "%s: posm64 must be 32-byte aligned", fn);
  GRAV_REQUIRE(aligned(v64, 8) && aligned(acc64, 8), "%s: v64/acc64 misaligned", fn);
  const int nn = static_cast<int>(n);
  hipLaunchKernelGGL(k_stage2_f64, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<double *>(posm64),
                     static_cast<double *>(v64), static_cast<const double *>(acc64), nn, dt,
                     static_cast<double *>(mirror));
  return check_launch("k_stage2_f64");
}

int mirror_ok(const char *fn, const void *mirror, uintptr_t align) {
  GRAV_REQUIRE(mirror != nullptr && aligned(mirror, align),
               "%s: mirror must be a %d-byte aligned device address of a host mirror", fn,
               static_cast<int>(align));
  return GRAV_OK;
}
}  // namespace

int grav_stage2_f32(void *posm, void *vel, const void *acc, int64_t n, double dt,
                    void *stream) {
  clear_error();
  return stage2_f32("grav_stage2_f32", posm, vel, acc, n, dt, nullptr, stream);
}

int grav_stage2_f64m(void *posm, void *r64, void *v64, const void *acc, int64_t n, double dt,
                     void *stream) {
  clear_error();
  return stage2_f64m("grav_stage2_f64m", posm, r64, v64, acc, n, dt, nullptr, stream);
}

int grav_stage2_f64(void *posm64, void *v64, const void *acc64, int64_t n, double dt,
                    void *stream) {
  clear_error();
  return stage2_f64("grav_stage2_f64", posm64, v64, acc64, n, dt, nullptr, stream);
}

int grav_stage2_f32_mirror(void *posm, void *vel, const void *acc, int64_t n, double dt,
                           void *mirror, void *stream) {
  clear_error();
  const int rc = mirror_ok("grav_stage2_f32_mirror", mirror, 16);
  if (rc != GRAV_OK) return rc;
  return stage2_f32("grav_stage2_f32_mirror", posm, vel, acc, n, dt, mirror, stream);
}

int grav_stage2_f64m_mirror(void *posm, void *r64, void *v64, const void *acc, int64_t n,
                            double dt, void *mirror, void *stream) {
  clear_error();
  const int rc = mirror_ok("grav_stage2_f64m_mirror", mirror, 8);
  if (rc != GRAV_OK) return rc;
  return stage2_f64m("grav_stage2_f64m_mirror", posm, r64, v64, acc, n, dt, mirror, stream);
}

int grav_stage2_f64_mirror(void *posm64, void *v64, const void *acc64, int64_t n, double dt,
                           void *mirror, void *stream) {
  clear_error();
  const int rc = mirror_ok("grav_stage2_f64_mirror", mirror, 32);
  if (rc != GRAV_OK) return rc;
  return stage2_f64("grav_stage2_f64_mirror", posm64, v64, acc64, n, dt, mirror, stream);
}

int grav_host_mirror_alloc(int64_t bytes, void **host, void **device) {
  clear_error();
  GRAV_REQUIRE(host != nullptr && device != nullptr, "grav_host_mirror_alloc: NULL output");
  GRAV_REQUIRE(bytes >= 1 && bytes <= (int64_t(1) << 40),
               "grav_host_mirror_alloc: bytes must be in [1, 2^40] (got %lld)", (long long)bytes);
  *host = nullptr;
  *device = nullptr;
  void *h = nullptr;
  GRAV_HIP_CHECK(hipHostMalloc(&h, static_cast<size_t>(bytes),
                               hipHostMallocMapped | hipHostMallocCoherent));
  void *d = nullptr;
  const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
  if (e != hipSuccess || d == nullptr) {
    (void)hipHostFree(h);
    return fail(GRAV_EHIP, "grav_host_mirror_alloc: hipHostGetDevicePointer: %s",
                hipGetErrorString(e));
  }
  *host = h;
  *device = d;
  return GRAV_OK;
}

int grav_host_mirror_free(void *host) {
  clear_error();
  if (host == nullptr) return GRAV_OK;
  GRAV_HIP_CHECK(hipHostFree(host));
  return GRAV_OK;
}

}  // extern "C"
