// Allocation-free variant of the j-split FP32 stage 1 (grav_stage1_f32_split): the partial
// sums of the j-ranges live in a file-scope device array of the library instead of a
// workspace the caller allocates, so a call allocates no memory at all.  For callers that
// cannot allocate: code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.
//
// Each GPU holds its own copy of `scratch` (kScratchBytes) from the moment the library's code
// object is loaded on it until the process ends.  Host code never passes the array's name to
// a kernel (on the host it is only a stand-in for the symbol): the first call on a device, with
// that device current, asks the runtime for the array's device address with
// hipGetSymbolAddress, and every launch takes that address.  The partials of a call are
// written in full before they are read, so nothing needs clearing between calls; calls that
// overlap on one GPU (other streams or threads) share the array and race.
#include <mutex>

#include "grav_common.h"

namespace grav {
namespace {

constexpr int64_t kScratchBytes = GRAV_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size must keep 256-byte alignment");

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(void **out) {
  int dev = -1;
  GRAV_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(GRAV_ENOTSUP, "device ordinal %d beyond the %d the library tracks", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(GRAV_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GRAV_OK;
}

}  // namespace
}  // namespace grav

using namespace grav;

extern "C" {

int grav_stage1_f32_split_noalloc(const void *posm, void *acc, int64_t n, double G, double eps,
                                  int i_per_lane, int j_waves, int splits, void *stream) {
  clear_error();
  const char *fn = "grav_stage1_f32_split_noalloc";
  GRAV_REQUIRE(posm != nullptr, "%s: posm is NULL", fn);
  GRAV_REQUIRE(acc != nullptr, "%s: acc is NULL", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(aligned(posm, 16), "%s: posm must be 16-byte aligned", fn);
  GRAV_REQUIRE(aligned(acc, 4), "%s: acc must be 4-byte aligned", fn);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "%s: eps must be finite and >= 0 (got %g)", fn, eps);
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) split_default_cfg(n, false, K, S, P);
  GRAV_REQUIRE(split_supported(K, S, P),
               "%s: (i_per_lane, j_waves, splits) must be in {1, 2} x {1, 2, 4, 8} x [1, %d], "
               "or all 0 (got %d, %d, %d)", fn, kMaxSplits, K, S, P);
  const int64_t need = split_workspace_bytes(n, P);
  GRAV_REQUIRE(need <= kScratchBytes,
               "%s: %lld bodies in %d j-ranges need %lld bytes of partials, more than the "
               "library's %lld-byte array (32 * splits * n <= %lld)", fn, (long long)n, P,
               (long long)need, (long long)kScratchBytes, (long long)kScratchBytes);
  void *ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc != GRAV_OK) return rc;
  return stage1_f32_split_launch(static_cast<const float4 *>(posm), static_cast<float *>(acc),
                                 static_cast<double *>(ws), static_cast<int>(n),
                                 static_cast<float>(G), static_cast<float>(eps * eps),
                                 eps == 0.0 ? 1 : 0, K, S, P, static_cast<hipStream_t>(stream));
}

}  // extern "C"
