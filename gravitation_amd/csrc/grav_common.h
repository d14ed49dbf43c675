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

// One block of IB bodies of a universe for the batched symmetric stage 1: bodies
// [start, end), the universe's first block-pair item and block count, and the block's index.
struct SymBlock {
  int start, end, item_base, nb, beta, pad[3];
};

// Batched symmetric FP32 stage 1 (stage1_sym.hip): diag_items {i0, i1, j0, j1} (tiles of 256
// bodies against their own block) into diag, pair_items {beta start, gamma start, beta end,
// gamma end} (block pairs) into the partials pi / pj (3 * IB floats per item each), then
// acc = G * (diag + partials) per body; IB = 64*K*W, K in {4,8}, W in {1,2,4}.
int sym_batched_launch(const float4 *posm, float *acc, int n, float G, float eps2, int mask,
                       int K, int W, const int4 *pair_items, int npairs, const int4 *diag_items,
                       int ndiag, const SymBlock *blocks, int nblocks, float *diag, float *pi,
                       float *pj, hipStream_t s);
bool sym_batched_supported(int K, int W);

// 1/sqrt(x) in FP64: the v_rsq_f64 seed refined by one Newton-Raphson step
// y <- y + (y/2)(1 - x y^2).  Measured on gfx950 over 2^20 arguments spanning 2^-60..2^61,
// the seed's relative error is at most 5.2e-8 (2^-24.2) and the refined value's at most
// 4.1e-15 (1.5 eps^2, always below the true value), 1.7e-16 on average, so a second step
// (4 more FP64 ops per pair) would buy under 2e-14 of the 1e-12 error budget.
__device__ __forceinline__ double rsqrt_nr64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-(x * y), y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}

// FP64 stage 1 of one universe of n bodies, or, with diag_block > 0 (a multiple of 128), of
// its diagonal blocks of diag_block bodies only.  (stage1.hip)
int stage1_f64_launch(const double4 *posm, double *acc, int n, double G, double eps2, int mask,
                      int diag_block, hipStream_t stream);

// FP32 stage 1 over an item list (items[b] = {i0, i1, j0, j1}: bodies [i0, i1), at most
// 64*K of them, against j in [j0, j1)) or, with items == NULL, over one universe of n bodies
// (nitems == 0) or over the diagonal blocks of nitems bodies (a multiple of 64*K) only.
// K in {1,2,4,8} bodies per lane, S in {1,2,4,8} wavefronts splitting j (or 16 with K <= 2).
// (stage1.hip)
int stage1_f32_launch(const float4 *posm, float *acc, int n, float G, float eps2, int mask,
                      int K, int S, int jblock, const int4 *items, int nitems,
                      hipStream_t stream);

// Unscaled FP32 sums of the bodies [i0, i0 + m) against their own blocks of diag_block bodies
// (i0 and diag_block multiples of 256), into acc (3, m): the block-diagonal pass of the
// symmetric form for a range of blocks, (K, S) = (4, 4).  (stage1.hip)
int stage1_f32_diag_launch(const float4 *posm, float *acc, int n, int i0, int m, float eps2,
                           int mask, int diag_block, hipStream_t stream);

// mu = nu + h rho (FP64, into nu) and its FP32 copy scaled by 2^-E per universe, E in
// mu4.w; one workgroup per universe of offsets (device, n_universes + 1 entries), or one
// universe of n bodies when offsets is NULL and n_universes == 1. (adjoint.hip)
int adjoint_mu_launch(const double *rho, double *nu, float4 *mu4, int n, double dt,
                      const int64_t *offsets, int n_universes, hipStream_t stream);

// VJP of stage 1 over an item list with the same convention (K in {1,2,4}). (adjoint.hip)
int stage1_vjp_launch(const float4 *posm, const float4 *mu4, double *rho, double *gamma, int n,
                      double hG, float eps2, int K, int S, const int4 *items, int nitems,
                      hipStream_t stream);

// j-split form of the owner-computes stage 1 and of its VJP, for universes too small to fill
// the GPU with i-tiles: a (tiles x splits) grid, each workgroup one i-tile of 64*K bodies
// against one of `splits` equal j-ranges, writes its FP64 partial to part[split][c][i]
// (3 components for stage 1, 4 for the VJP), and one small kernel sums the ranges in order.
constexpr int kMaxSplits = 64;
inline bool split_supported(int K, int S, int splits) {
  return (K == 1 || K == 2) && (S == 1 || S == 2 || S == 4 || S == 8) && splits >= 1 &&
         splits <= kMaxSplits;
}
// Workspace bytes for either form: splits * 4 * n doubles.
inline int64_t split_workspace_bytes(int64_t n, int splits) {
  return static_cast<int64_t>(splits) * 4 * n * 8;
}
// FP64 stage 1 in the j-split form: K in {1,2}, S in {4,8}.
inline bool split_supported_f64(int K, int S, int splits) {
  return (K == 1 || K == 2) && (S == 4 || S == 8) && splits >= 1 && splits <= kMaxSplits;
}
// Default (K, S, splits) by n for FP32 stage 1 or the VJP, the FP64 stage 1's, and the
// largest of their splits (the default workspace covers all three).  (stage1.hip)
void split_default_cfg(int64_t n, bool vjp, int &K, int &S, int &splits);
void split_default_cfg_f64(int64_t n, int &K, int &S, int &splits);
int split_default_splits(int64_t n);
int stage1_f64_split_launch(const double4 *posm, double *acc, double *part, int n, double G,
                            double eps2, int mask, int K, int S, int splits, hipStream_t stream);
int stage1_f32_split_launch(const float4 *posm, float *acc, double *part, int n, float G,
                            float eps2, int mask, int K, int S, int splits, hipStream_t stream);
// FP32 stage 1 of the bodies [i0, i0 + m) of n against all n bodies, into acc (3, m):
// owner-computes with splits == 1 (K, S as stage1_f32_launch), else the j-split form with
// part[splits][3][m]; its rows equal the whole-universe launch's of the same shape bit for
// bit.  shard_default_cfg: the default shape for m of n bodies.  (stage1.hip)
void shard_default_cfg(int64_t n, int64_t m, int &K, int &S, int &splits);
void shard_vjp_default_cfg(int64_t n, int64_t m, int &K, int &S, int &splits);
// The VJP of the bodies [i0, i0 + m) of n into rho (3, m) and gamma (m): owner-computes with
// splits == 1, else the j-split form with part[splits][4][m].  (adjoint.hip)
int stage1_vjp_shard_launch(const float4 *posm, const float4 *mu4, double *rho, double *gamma,
                            double *part, int n, int i0, int m, double hG, float eps2, int K,
                            int S, int splits, hipStream_t stream);
int stage1_f32_shard_launch(const float4 *posm, float *acc, double *part, int n, int i0, int m,
                            float G, float eps2, int mask, int K, int S, int splits,
                            hipStream_t stream);
int stage1_vjp_split_launch(const float4 *posm, const float4 *mu4, double *rho, double *gamma,
                            double *part, int n, double hG, float eps2, int K, int S, int splits,
                            hipStream_t stream);  // (adjoint.hip)
int stage1_vjp_split_chain_launch(const float4 *posm, float4 *mu4, double *rho, double *nu,
                                  double *gamma, double *part, int n, double hG, double dt,
                                  float eps2, int K, int S, int splits, int next_mu,
                                  unsigned *eslot, hipStream_t stream);  // (adjoint.hip)

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
