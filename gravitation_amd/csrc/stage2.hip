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

__global__ __launch_bounds__(kBlock) void k_stage2_f32(float4 *__restrict__ posm,
                                                       float *__restrict__ vel,
                                                       const float *__restrict__ acc, int n,
                                                       float dt) {
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
}

// FP64 masters advanced with FP32 accelerations; the FP32 stage-1 copy gets x, y, z only
// (a 12-byte store, so m is neither read nor written).
__global__ __launch_bounds__(kBlock) void k_stage2_f64m(float *__restrict__ posm,
                                                        double *__restrict__ r64,
                                                        double *__restrict__ v64,
                                                        const float *__restrict__ acc, int n,
                                                        double dt) {
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
}

__global__ __launch_bounds__(kBlock) void k_stage2_f64(double *__restrict__ posm64,
                                                       double *__restrict__ v64,
                                                       const double *__restrict__ acc, int n,
                                                       double dt) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double *p = posm64 + 4 * static_cast<size_t>(i);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = acc[c * n + i] * dt + v64[c * n + i];
    v64[c * n + i] = v;
    p[c] = v * dt + p[c];
  }
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

int grav_stage2_f32(void *posm, void *vel, const void *acc, int64_t n, double dt,
                    void *stream) {
  clear_error();
  int rc = validate("grav_stage2_f32", posm, vel, acc, n, dt);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(aligned(posm, 16), "grav_stage2_f32: posm must be 16-byte aligned");
  GRAV_REQUIRE(aligned(vel, 4) && aligned(acc, 4), "grav_stage2_f32: vel/acc misaligned");
  const int nn = static_cast<int>(n);
  hipLaunchKernelGGL(k_stage2_f32, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<float4 *>(posm),
                     static_cast<float *>(vel), static_cast<const float *>(acc), nn,
                     static_cast<float>(dt));
  return check_launch("k_stage2_f32");
}

int grav_stage2_f64m(void *posm, void *r64, void *v64, const void *acc, int64_t n, double dt,
                     void *stream) {
  clear_error();
  int rc = validate("grav_stage2_f64m", posm, r64, v64, n, dt);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(acc != nullptr, "grav_stage2_f64m: acc is NULL");
  GRAV_REQUIRE(aligned(posm, 16), "grav_stage2_f64m: posm must be 16-byte aligned");
  GRAV_REQUIRE(aligned(r64, 8) && aligned(v64, 8) && aligned(acc, 4),
               "grav_stage2_f64m: r64/v64/acc misaligned");
  const int nn = static_cast<int>(n);
  hipLaunchKernelGGL(k_stage2_f64m, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<float *>(posm),
                     static_cast<double *>(r64), static_cast<double *>(v64),
                     static_cast<const float *>(acc), nn, dt);
  return check_launch("k_stage2_f64m");
}

int grav_stage2_f64(void *posm64, void *v64, const void *acc64, int64_t n, double dt,
                    void *stream) {
  clear_error();
  int rc = validate("grav_stage2_f64", posm64, v64, acc64, n, dt);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(aligned(posm64, 32), "grav_stage2_f64: posm64 must be 32-byte aligned");
  GRAV_REQUIRE(aligned(v64, 8) && aligned(acc64, 8), "grav_stage2_f64: v64/acc64 misaligned");
  const int nn = static_cast<int>(n);
  hipLaunchKernelGGL(k_stage2_f64, dim3((nn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), static_cast<double *>(posm64),
                     static_cast<double *>(v64), static_cast<const double *>(acc64), nn, dt);
  return check_launch("k_stage2_f64");
}

}  // extern "C"
