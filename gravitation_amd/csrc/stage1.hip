// Stage 1: all-pairs accelerations, owner-computes (every body i is summed by exactly one
// lane, so there are no atomics and results are deterministic).
//
// Physics follows the reference's canonical pair formula (src/gravitation/kernel/py1.py:57-64)
// in the all-pairs form of src/gravitation/kernel/pc2.py:60-90 (self term skipped, G applied
// once at the end), plus optional Plummer softening eps:
//   a_i = G * sum_{j != i} m_j (r_j - r_i) * (|r_j - r_i|^2 + eps^2)^(-3/2)
//
// FP32 kernel layout (gfx950, wave64):
//   * a workgroup owns 64*K consecutive i-bodies (lane l holds i = base + k*64 + l, k < K) and
//     runs S wavefronts that split the j-range S ways; the S partial sums are folded through
//     LDS at the end in a fixed order;
//   * j-bodies are read as wave-uniform float4 (x, y, z, m) loads, which the compiler turns
//     into scalar (SMEM) loads: every j operand of the pair arithmetic is an SGPR, so a pair
//     costs 12 VALU ops + 1 v_rsq_f32 and no LDS or VGPR traffic for j;
//   * compensated accumulation: j is processed in blocks of `jblock` bodies whose FP32 partial
//     sums are folded into an FP32 (hi, lo) pair with Knuth's TwoSum once per block.
//   * eps == 0 excludes the self term: only the j-segment overlapping the workgroup's own
//     i-range runs the masked pair routine, the rest runs unmasked.
#include <algorithm>

#include "grav_common.h"
#include "stage1_f32.h"

namespace grav {
namespace {

// Workgroup (blockIdx.x, blockIdx.y) of stage1_f32_block.
template <int K, int S, int JB>
__global__ __launch_bounds__(kWave *S) void k_stage1_f32(const float4 *__restrict__ posm,
                                                          float *__restrict__ acc, int n,
                                                          int i0, int m, float G, float eps2,
                                                          int mask_self, int jblock,
                                                          const int4 *__restrict__ items,
                                                          int diag_block,
                                                          double *__restrict__ part) {
  extern __shared__ float red[];  // (S - 1) * 6 * (64 * K) floats
  stage1_f32_block<K, S, JB>(posm, acc, n, i0, m, G, eps2, mask_self, jblock, items,
                             diag_block, part, static_cast<int>(blockIdx.x),
                             static_cast<int>(blockIdx.y), static_cast<int>(gridDim.y), red);
}

// acc[c][i] = G * split_sum (T = float or double).
template <typename T>
__global__ __launch_bounds__(256) void k_split_reduce(const double *__restrict__ part,
                                                      T *__restrict__ acc, int n, int splits,
                                                      double G) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  acc[static_cast<size_t>(c) * n + i] = static_cast<T>(G * split_sum(part, c, i, n, splits));
}

// split_step_body for every body (out may be in).
__global__ __launch_bounds__(256) void k_split_step_f64m(const double *__restrict__ part,
                                                         const float4 *in, float4 *out,
                                                         double *__restrict__ r64,
                                                         double *__restrict__ v64, int n,
                                                         int splits, double G, double dt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) split_step_body(part, in, out, r64, v64, n, splits, G, dt, i);
}

// ---------------------------------------------------------------------------------------------
// FP64 (rsqrt_nr64 is in grav_common.h).

template <int K, bool MASK>
__device__ __forceinline__ void pair_f64(const double4 q, const int j, const double (&xi)[K],
                                         const double (&yi)[K], const double (&zi)[K],
                                         const int (&ii)[K], const double eps2,
                                         double (&ax)[K], double (&ay)[K], double (&az)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double dx = q.x - xi[k];
    const double dy = q.y - yi[k];
    const double dz = q.z - zi[k];
    double r2 = __builtin_fma(dx, dx, eps2);
    r2 = __builtin_fma(dy, dy, r2);
    r2 = __builtin_fma(dz, dz, r2);
    double inv = rsqrt_nr64(r2);
    if (MASK) inv = (j == ii[k]) ? 0.0 : inv;
    const double s = (q.w * inv) * (inv * inv);
    ax[k] = __builtin_fma(dx, s, ax[k]);
    ay[k] = __builtin_fma(dy, s, ay[k]);
    az[k] = __builtin_fma(dz, s, az[k]);
  }
}

// j in [j0, j1) in order, one j-body per half-step, each loaded one half-step ahead in two
// SGPR sets (as the FP32 loop of one body per lane: wait for every scalar load, issue the
// next one, then compute).  MI355X, Plummer, ms: 2^18 39.1-39.2 against 41.7, 2^16 2.56-2.66
// against 3.34-3.50, the j-split form at 8192 bodies 0.054-0.055 against 0.056-0.059; same
// bits.
template <int K, bool MASK>
__device__ __forceinline__ void pair_range_f64(const double4 *__restrict__ posm, int j0, int j1,
                                               const double (&xi)[K], const double (&yi)[K],
                                               const double (&zi)[K], const int (&ii)[K],
                                               const double eps2, double (&ax)[K],
                                               double (&ay)[K], double (&az)[K]) {
  int j = j0;
  if (j < j1) {
    double4 qa = posm[j], qb;
    for (; j + 2 <= j1; j += 2) {
      __builtin_amdgcn_s_waitcnt(kWaitScalar);
      qb = posm[j + 1];
      __builtin_amdgcn_sched_barrier(0);
      pair_f64<K, MASK>(qa, j, xi, yi, zi, ii, eps2, ax, ay, az);
      // the next body of set a, or (past the end) a body in range that is not used
      const int jn = min(j + 2, j1 - 1);
      __builtin_amdgcn_s_waitcnt(kWaitScalar);
      qa = posm[jn];
      __builtin_amdgcn_sched_barrier(0);
      pair_f64<K, MASK>(qb, j + 1, xi, yi, zi, ii, eps2, ax, ay, az);
    }
    if (j < j1) pair_f64<K, MASK>(qa, j, xi, yi, zi, ii, eps2, ax, ay, az);
  }
}

// FP64: workgroup owns 64*K i-bodies; S wavefronts split the j-range and their FP64 sums are
// added through LDS in a fixed order (deterministic).  diag_block > 0: each i-tile only
// against its own block of diag_block bodies (a multiple of 64*K).  part != NULL: the
// j-split form, as k_stage1_f32's (the unscaled sums go to part[blockIdx.y][c][i]).
template <int K, int S>
__global__ __launch_bounds__(kWave *S) void k_stage1_f64(const double4 *__restrict__ posm,
                                                          double *__restrict__ acc, int n,
                                                          double G, double eps2,
                                                          int mask_self, int diag_block,
                                                          double *__restrict__ part) {
  extern __shared__ double red64[];  // (S - 1) * 3 * (64 * K)
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int ibase = blockIdx.x * (kWave * K);
  const int iend = min(n, ibase + kWave * K);
  double xi[K], yi[K], zi[K], ax[K], ay[K], az[K];
  int ii[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = ibase + k * kWave + lane;
    ii[k] = i;
    const double4 p = posm[min(i, n - 1)];
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
    ax[k] = ay[k] = az[k] = 0.0;
  }
  int jlo = diag_block > 0 ? (ibase / diag_block) * diag_block : 0;
  int jhi = diag_block > 0 ? min(n, jlo + diag_block) : n;
  if (part != nullptr) {
    const int ns = static_cast<int>(gridDim.y);
    const int span = (n + ns - 1) / ns;
    jlo = min(n, static_cast<int>(blockIdx.y) * span);
    jhi = min(n, jlo + span);
  }
  const int chunk = (jhi - jlo + S - 1) / S;
  const int j0 = min(jhi, jlo + wave * chunk);
  const int j1 = min(jhi, j0 + chunk);
  if (mask_self) {
    const int a = min(j1, max(j0, ibase));
    const int b = max(a, min(j1, iend));
    pair_range_f64<K, false>(posm, j0, a, xi, yi, zi, ii, eps2, ax, ay, az);
    pair_range_f64<K, true>(posm, a, b, xi, yi, zi, ii, eps2, ax, ay, az);
    pair_range_f64<K, false>(posm, b, j1, xi, yi, zi, ii, eps2, ax, ay, az);
  } else {
    pair_range_f64<K, false>(posm, j0, j1, xi, yi, zi, ii, eps2, ax, ay, az);
  }
  if constexpr (S > 1) {
    constexpr int L = kWave * K;
    if (wave > 0) {
      double *r = red64 + (wave - 1) * 3 * L;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        r[0 * L + k * kWave + lane] = ax[k];
        r[1 * L + k * kWave + lane] = ay[k];
        r[2 * L + k * kWave + lane] = az[k];
      }
    }
    __syncthreads();
    if (wave != 0) return;
    for (int w = 1; w < S; ++w) {
      const double *r = red64 + (w - 1) * 3 * L;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        ax[k] += r[0 * L + k * kWave + lane];
        ay[k] += r[1 * L + k * kWave + lane];
        az[k] += r[2 * L + k * kWave + lane];
      }
    }
  }
  if (part != nullptr) {
    double *o = part + static_cast<size_t>(blockIdx.y) * 3 * n;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = ii[k];
      if (i < n) {
        o[i] = ax[k];
        o[n + i] = ay[k];
        o[2 * n + i] = az[k];
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = ii[k];
    if (i < n) {
      acc[i] = G * ax[k];
      acc[n + i] = G * ay[k];
      acc[2 * n + i] = G * az[k];
    }
  }
}

// i-tiles of the shard [i0, i0 + m) (items == NULL) or the nitems items, each against all n
// bodies; see stage1_f32_block.
template <int K, int S>
int launch_f32(const float4 *posm, float *acc, int n, int i0, int m, float G, float eps2,
               int mask, int jblock, const int4 *items, int nitems, hipStream_t stream) {
  const int per_block = kWave * K;
  const int grid = items ? nitems : (m + per_block - 1) / per_block;
  const size_t lds = (S > 1) ? sizeof(float) * (S - 1) * 6 * per_block : 0;
  if (grid == 0) return GRAV_OK;
  constexpr int kNarrow = K >= 4 ? 8 : 2;   // pair_loop_width: eight from K = 4 on
  if (pair_loop_width(grid, K) == 8)
    hipLaunchKernelGGL((k_stage1_f32<K, S, 8>), dim3(grid), dim3(kWave * S), lds, stream, posm,
                       acc, n, i0, m, G, eps2, mask, jblock, items, items ? 0 : nitems,
                       static_cast<double *>(nullptr));
  else
    hipLaunchKernelGGL((k_stage1_f32<K, S, kNarrow>), dim3(grid), dim3(kWave * S), lds, stream,
                       posm, acc, n, i0, m, G, eps2, mask, jblock, items, items ? 0 : nitems,
                       static_cast<double *>(nullptr));
  return check_launch("k_stage1_f32");
}

// The j-split partials part[splits][3][m] of the shard [i0, i0 + m).
template <int K, int S>
int launch_f32_split(const float4 *posm, double *part, int n, int i0, int m, float eps2,
                     int mask, int splits, hipStream_t stream) {
  const int per_block = kWave * K;
  const int tiles = (m + per_block - 1) / per_block;
  const size_t lds = (S > 1) ? sizeof(float) * (S - 1) * 6 * per_block : 0;
  constexpr int kNarrow = K >= 4 ? 8 : 2;   // pair_loop_width: eight from K = 4 on
  if (pair_loop_width(int64_t(tiles) * splits, K) == 8)
    hipLaunchKernelGGL((k_stage1_f32<K, S, 8>), dim3(tiles, splits), dim3(kWave * S), lds,
                       stream, posm, static_cast<float *>(nullptr), n, i0, m, 1.0f, eps2, mask,
                       256, static_cast<const int4 *>(nullptr), 0, part);
  else
    hipLaunchKernelGGL((k_stage1_f32<K, S, kNarrow>), dim3(tiles, splits), dim3(kWave * S),
                       lds, stream, posm, static_cast<float *>(nullptr), n, i0, m, 1.0f, eps2,
                       mask, 256, static_cast<const int4 *>(nullptr), 0, part);
  return check_launch("k_stage1_f32 (j-split)");
}

template <int K>
int launch_f32_s(int S, const float4 *posm, float *acc, int n, int i0, int m, float G,
                 float eps2, int mask, int jblock, const int4 *items, int nitems,
                 hipStream_t stream) {
#define GRAV_F32(SS)   launch_f32<K, SS>(posm, acc, n, i0, m, G, eps2, mask, jblock, items, nitems, stream)
  switch (S) {
    case 1: return GRAV_F32(1);
    case 2: return GRAV_F32(2);
    case 4: return GRAV_F32(4);
    case 8: return GRAV_F32(8);
    case 16:
      // 16 wavefronts fold 15 partials through LDS: 23 KB at K = 1, 46 KB at K = 2
      if constexpr (K <= 2) return GRAV_F32(16);
      break;
  }
#undef GRAV_F32
  return fail(GRAV_EINVAL,
              "j_waves must be 1, 2, 4 or 8, or 16 with i_per_lane 1 or 2 (got %d with %d)", S,
              K);
}

// Default launch shape: one body per lane and sixteen wavefronts per workgroup from 8192
// bodies, eight below.  MI355X, Plummer spheres, ms, with the pair loop taking two j-bodies
// per step: 2^14 (1, 16) 0.104 against 0.140 for (1, 8) and 0.180 for (2, 16); 2^16 1.14
// against 1.24 and 1.21; 2^18 17.0 against 17.2 and 17.9 ((2, 8) 18.4, (4, 4) 21.6); 2^20
// 282 against 273 and 279.  (With eight j-bodies per step, (2, 16) had been the fastest.)
void default_cfg_f32(int64_t n, int &K, int &S) {
  K = 1;
  S = n >= 8192 ? 16 : 8;
}

int validate_common(const char *fn, const void *posm, const void *acc, int64_t n, double G,
                    double eps, uintptr_t align) {
  GRAV_REQUIRE(posm != nullptr, "%s: posm is NULL", fn);
  GRAV_REQUIRE(acc != nullptr, "%s: acc is NULL", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(aligned(posm, align), "%s: posm must be %d-byte aligned", fn, (int)align);
  GRAV_REQUIRE(aligned(acc, 4), "%s: acc must be 4-byte aligned", fn);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "%s: eps must be finite and >= 0 (got %g)", fn, eps);
  return GRAV_OK;
}

}  // namespace

int stage1_f64_launch(const double4 *posm, double *acc, int n, double G, double eps2, int mask,
                      int diag_block, hipStream_t stream) {
  // Fill the GPU: 64*K bodies per workgroup, S wavefronts over j.
  if (n >= (1 << 16)) {
    const int grid = (n + 127) / 128;
    hipLaunchKernelGGL((k_stage1_f64<2, 4>), dim3(grid), dim3(256),
                       sizeof(double) * 3 * 3 * 128, stream, posm, acc, n, G, eps2, mask,
                       diag_block, static_cast<double *>(nullptr));
  } else {
    const int grid = (n + 63) / 64;
    hipLaunchKernelGGL((k_stage1_f64<1, 8>), dim3(grid), dim3(512),
                       sizeof(double) * 7 * 3 * 64, stream, posm, acc, n, G, eps2, mask,
                       diag_block, static_cast<double *>(nullptr));
  }
  return check_launch("k_stage1_f64");
}

namespace {
template <int K, int S>
int launch_f64_split(const double4 *posm, double *part, int n, double eps2, int mask,
                     int splits, hipStream_t stream) {
  const int per_block = kWave * K;
  const int tiles = (n + per_block - 1) / per_block;
  const size_t lds = sizeof(double) * (S - 1) * 3 * per_block;
  hipLaunchKernelGGL((k_stage1_f64<K, S>), dim3(tiles, splits), dim3(kWave * S), lds, stream,
                     posm, static_cast<double *>(nullptr), n, 1.0, eps2, mask, 0, part);
  return check_launch("k_stage1_f64 (j-split)");
}
}  // namespace

int stage1_f64_split_launch(const double4 *posm, double *acc, double *part, int n, double G,
                            double eps2, int mask, int K, int S, int splits,
                            hipStream_t stream) {
  if (!split_supported_f64(K, S, splits))
    return fail(GRAV_EINVAL, "FP64 j-split shape (%d, %d, %d) not supported", K, S, splits);
  int rc = GRAV_OK;
  if (K == 1 && S == 4) rc = launch_f64_split<1, 4>(posm, part, n, eps2, mask, splits, stream);
  if (K == 1 && S == 8) rc = launch_f64_split<1, 8>(posm, part, n, eps2, mask, splits, stream);
  if (K == 2 && S == 4) rc = launch_f64_split<2, 4>(posm, part, n, eps2, mask, splits, stream);
  if (K == 2 && S == 8) rc = launch_f64_split<2, 8>(posm, part, n, eps2, mask, splits, stream);
  if (rc != GRAV_OK) return rc;
  hipLaunchKernelGGL((k_split_reduce<double>), dim3((n + 255) / 256, 3), dim3(256), 0, stream,
                     part, acc, n, splits, G);
  return check_launch("k_split_reduce<double>");
}

namespace {
int launch_f32_ks(int K, int S, const float4 *posm, float *acc, int n, int i0, int m, float G,
                  float eps2, int mask, int jblock, const int4 *items, int nitems,
                  hipStream_t stream) {
#define GRAV_F32(KK)   launch_f32_s<KK>(S, posm, acc, n, i0, m, G, eps2, mask, jblock, items, nitems, stream)
  switch (K) {
    case 1: return GRAV_F32(1);
    case 2: return GRAV_F32(2);
    case 4: return GRAV_F32(4);
    case 8: return GRAV_F32(8);
  }
#undef GRAV_F32
  return fail(GRAV_EINVAL, "i_per_lane must be 1, 2, 4 or 8 (got %d)", K);
}
}  // namespace

int stage1_f32_launch(const float4 *posm, float *acc, int n, float G, float eps2, int mask,
                      int K, int S, int jblock, const int4 *items, int nitems,
                      hipStream_t stream) {
  return launch_f32_ks(K, S, posm, acc, n, 0, n, G, eps2, mask, jblock, items, nitems, stream);
}

int stage1_f32_diag_launch(const float4 *posm, float *acc, int n, int i0, int m, float eps2,
                           int mask, int diag_block, hipStream_t stream) {
  return launch_f32_ks(4, 4, posm, acc, n, i0, m, 1.0f, eps2, mask, 256, nullptr, diag_block,
                       stream);
}

void split_default_cfg(int64_t n, bool vjp, int &K, int &S, int &splits) {
  // Shortest kernel durations (rocprofv3 kernel trace) of a sweep over {1,2} x {2,4,8} x
  // {2..32} on an MI355X, Plummer spheres of 2048 to 16384 bodies: eight wavefronts per
  // workgroup, 8 or 16 j-ranges.  The VJP alone is shortest with 16 ranges from 3072 bodies,
  // but in the backward pass it shares the GPU with the recomputation of the next segment,
  // and there 8 ranges win up to 6144 bodies (4096 bodies, T = 200: forward + adjoint 7.95 ms
  // against 8.17 ms, medians of six interleaved rounds).
  // Stage 1 with the pair loop taking two j-bodies per step (HIP events, two launches, µs):
  // 8192 (1, 8, 16) 29.9 against 30.8 for (1, 8, 8); 16384 80.7 against 83.4 for (2, 8, 16).
  // With that loop's loads one step ahead (one body per lane only), one body per lane wins at
  // every size: 32768 (1, 8, 16) 327 against 350 for (2, 8, 16), 40960 479 against 505,
  // 2^18 16.0 against 17.9 ms.
  S = 8;
  K = 1;
  if (vjp)
    splits = n < 6144 ? 8 : 16;
  else
    splits = n < 8192 ? 8 : 16;
}

// Default shape for a share of m of n bodies.  MI355X, Plummer spheres of 2^18 bodies, shares
// of 2^14 to 2^17 bodies, with the pair loop taking two j-bodies per step: the j-split form
// with one body per lane and eight wavefronts, 32 j-ranges up to 1024 tiles of 64 bodies and 16
// beyond, within 1.6 % of the fastest of {1, 2} x {4, 8} x {8..64} and the owner-computes
// shapes at every size (2^15 of 2^18: 2.11 ms against 2.48 for owner-computes (1, 16), whose
// 512 i-tiles give two workgroups per CU).  From 4096 tiles on, owner-computes (1, 16), the
// whole universe's default.  The j-ranges keep at least 4096 bodies each.
void shard_default_cfg(int64_t n, int64_t m, int &K, int &S, int &splits) {
  const int64_t tiles = (m + 63) / 64;
  K = 1;
  if (tiles >= 4096) {
    S = 16;
    splits = 1;
    return;
  }
  S = 8;
  const int64_t want = tiles <= 1024 ? 32 : 16;
  splits = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(want, n / 4096)));
}

// Default VJP shape for a share of m of n bodies: the j-split form with one body per lane and
// eight wavefronts, and about 16384 workgroups (64-body tiles x ranges; ranges of at least 4096
// bodies, at least two of them).  MI355X, Plummer 2^18, shares of 2^14, 2^15 and 2^16 bodies:
// within 0.3 % of the fastest of {1, 2} x {4, 8} x {4..64} and the owner-computes shapes, and
// within 1.2 % of the whole universe's VJP time / shares (2^15: 4.43 ms against 35.32 / 8 =
// 4.42; owner-computes (1, 8) 5.72).
void shard_vjp_default_cfg(int64_t n, int64_t m, int &K, int &S, int &splits) {
  const int64_t tiles = (m + 63) / 64;
  K = 1;
  S = 8;
  const int64_t want = std::min<int64_t>(64, std::max<int64_t>(2, 16384 / tiles));
  splits = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(want, n / 4096)));
}

void split_default_cfg_f64(int64_t n, int &K, int &S, int &splits) {
  // Within 5 % of the shortest kernel of a sweep over {1,2} x {4,8} x {2..32} at every n
  // from 512 to 16384 (Plummer, MI355X); FP64 pairs are long enough that 16 j-ranges pay
  // at any n where the form is used.
  (void)n;
  K = 1;
  S = 8;
  splits = 16;
}

int split_default_splits(int64_t n) {
  int K, S, a, b, c;
  split_default_cfg(n, false, K, S, a);
  split_default_cfg(n, true, K, S, b);
  split_default_cfg_f64(n, K, S, c);
  return std::max(a, std::max(b, c));
}

int split_step_f64m_launch(const float4 *in, float4 *out, double *r64, double *v64,
                           double *part, int n, float G, float eps2, int mask, int K, int S,
                           int splits, double dt, hipStream_t stream);

int stage1_f32_split_launch(const float4 *posm, float *acc, double *part, int n, float G,
                            float eps2, int mask, int K, int S, int splits,
                            hipStream_t stream) {
  if (!split_supported(K, S, splits))
    return fail(GRAV_EINVAL, "j-split shape (%d, %d, %d) not supported", K, S, splits);
  int rc = GRAV_OK;
#define GRAV_CASE(KK, SS) \
  if (K == KK && S == SS)   \
    rc = launch_f32_split<KK, SS>(posm, part, n, 0, n, eps2, mask, splits, stream);
  GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
  GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4) GRAV_CASE(2, 8)
#undef GRAV_CASE
  if (rc != GRAV_OK) return rc;
  hipLaunchKernelGGL((k_split_reduce<float>), dim3((n + 255) / 256, 3), dim3(256), 0, stream,
                     part, acc, n, splits, static_cast<double>(G));
  return check_launch("k_split_reduce<float>");
}

// Stage 1 of the bodies [i0, i0 + m) of n (one GPU's share when the bodies are divided among
// GPUs by i-range, every GPU holding all n positions): owner-computes with splits == 1, else
// the j-split form with its partials part[splits][3][m] reduced over the m rows.  Every form
// splits [0, n) among its wavefronts and j-ranges as for the whole universe, so the rows equal
// those of the whole-universe launch of the same shape bit for bit.
int stage1_f32_shard_launch(const float4 *posm, float *acc, double *part, int n, int i0, int m,
                            float G, float eps2, int mask, int K, int S, int splits,
                            hipStream_t stream) {
  if (splits == 1)
    return launch_f32_ks(K, S, posm, acc, n, i0, m, G, eps2, mask, 256, nullptr, 0, stream);
  if (!split_supported(K, S, splits))
    return fail(GRAV_EINVAL, "j-split shape (%d, %d, %d) not supported", K, S, splits);
  int rc = GRAV_OK;
#define GRAV_CASE(KK, SS) \
  if (K == KK && S == SS)   \
    rc = launch_f32_split<KK, SS>(posm, part, n, i0, m, eps2, mask, splits, stream);
  GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
  GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4) GRAV_CASE(2, 8)
#undef GRAV_CASE
  if (rc != GRAV_OK) return rc;
  hipLaunchKernelGGL((k_split_reduce<float>), dim3((m + 255) / 256, 3), dim3(256), 0, stream,
                     part, acc, m, splits, static_cast<double>(G));
  return check_launch("k_split_reduce<float>");
}

int split_step_f64m_launch(const float4 *in, float4 *out, double *r64, double *v64,
                           double *part, int n, float G, float eps2, int mask, int K, int S,
                           int splits, double dt, hipStream_t stream) {
  if (!split_supported(K, S, splits))
    return fail(GRAV_EINVAL, "j-split shape (%d, %d, %d) not supported", K, S, splits);
  int rc = GRAV_OK;
#define GRAV_CASE(KK, SS) \
  if (K == KK && S == SS)   \
    rc = launch_f32_split<KK, SS>(in, part, n, 0, n, eps2, mask, splits, stream);
  GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
  GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4) GRAV_CASE(2, 8)
#undef GRAV_CASE
  if (rc != GRAV_OK) return rc;
  // the step kernel writes out (which may be in) after the partials kernel, launched before
  // it on the same stream, has read every position
  hipLaunchKernelGGL(k_split_step_f64m, dim3((n + 255) / 256), dim3(256), 0, stream, part, in,
                     out, r64, v64, n, splits, static_cast<double>(G), dt);
  return check_launch("k_split_step_f64m");
}

}  // namespace grav

using namespace grav;

namespace {
int step_split_f64m(const char *fn, const void *posm_in, void *posm_out, void *r64, void *v64,
                    int64_t n, double G, double eps, double dt, int i_per_lane, int j_waves,
                    int splits, void *workspace, int64_t workspace_bytes, void *stream) {
  GRAV_REQUIRE(posm_in != nullptr && posm_out != nullptr && r64 != nullptr && v64 != nullptr,
               "%s: NULL buffer", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(aligned(posm_in, 16) && aligned(posm_out, 16),
               "%s: positions must be 16-byte aligned", fn);
  GRAV_REQUIRE(aligned(r64, 8) && aligned(v64, 8), "%s: r64/v64 misaligned", fn);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "%s: eps must be >= 0", fn);
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "%s: dt must be finite", fn);
  const char *a = static_cast<const char *>(posm_in);
  const char *b = static_cast<const char *>(posm_out);
  GRAV_REQUIRE(a == b || a + 16 * n <= b || b + 16 * n <= a,
               "%s: posm_out must be posm_in or not overlap it", fn);
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) split_default_cfg(n, false, K, S, P);
  GRAV_REQUIRE(split_supported(K, S, P),
               "%s: (i_per_lane, j_waves, splits) must be in {1, 2} x {1, 2, 4, 8} x [1, %d], "
               "or all 0 (got %d, %d, %d)", fn, kMaxSplits, K, S, P);
  const int64_t need = split_workspace_bytes(n, P);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
               "%s: workspace must be 256-byte aligned and hold %lld bytes", fn,
               (long long)need);
  return split_step_f64m_launch(static_cast<const float4 *>(posm_in),
                                static_cast<float4 *>(posm_out), static_cast<double *>(r64),
                                static_cast<double *>(v64), static_cast<double *>(workspace),
                                static_cast<int>(n), static_cast<float>(G),
                                static_cast<float>(eps * eps), eps == 0.0 ? 1 : 0, K, S, P, dt,
                                static_cast<hipStream_t>(stream));
}
}  // namespace

extern "C" {

int grav_stage1_f32_cfg(const void *posm, void *acc, int64_t n, double G, double eps, int accum,
                        int i_per_lane, int j_waves, int j_block, void *stream) {
  clear_error();
  int rc = validate_common("grav_stage1_f32", posm, acc, n, G, eps, 16);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(accum == GRAV_ACCUM_PLAIN || accum == GRAV_ACCUM_KAHAN,
               "grav_stage1_f32: accum must be GRAV_ACCUM_PLAIN or GRAV_ACCUM_KAHAN (got %d)",
               accum);
  GRAV_REQUIRE(j_block >= 8 && j_block % 8 == 0,
               "grav_stage1_f32: j_block must be a positive multiple of 8 (got %d)", j_block);
  const float G32 = static_cast<float>(G);
  const float eps2 = static_cast<float>(eps * eps);
  const int mask = (eps == 0.0) ? 1 : 0;
  const int nn = static_cast<int>(n);
  // Plain accumulation is the compensated loop with a single block: hi = sum, lo = 0.
  const int jb = (accum == GRAV_ACCUM_PLAIN) ? 0x7ffffff8 : j_block;
  const float4 *p = static_cast<const float4 *>(posm);
  float *a = static_cast<float *>(acc);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return stage1_f32_launch(p, a, nn, G32, eps2, mask, i_per_lane, j_waves, jb, nullptr, 0, s);
}

int grav_stage1_f32(const void *posm, void *acc, int64_t n, double G, double eps, int accum,
                    void *stream) {
  int K = 4, S = 4;
  default_cfg_f32(n, K, S);
  return grav_stage1_f32_cfg(posm, acc, n, G, eps, accum, K, S, 256, stream);
}

int grav_stage1_split_workspace(int64_t n, int splits, int64_t *bytes) {
  clear_error();
  GRAV_REQUIRE(bytes != nullptr, "grav_stage1_split_workspace: bytes is NULL");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_stage1_split_workspace: n must be in [1, 2^29] (got %lld)", (long long)n);
  if (splits == 0) splits = split_default_splits(n);
  GRAV_REQUIRE(splits >= 1 && splits <= kMaxSplits,
               "grav_stage1_split_workspace: splits must be in [1, %d] or 0 (got %d)", kMaxSplits,
               splits);
  *bytes = split_workspace_bytes(n, splits);
  return GRAV_OK;
}

int grav_stage1_f32_split(const void *posm, void *acc, int64_t n, double G, double eps,
                          int i_per_lane, int j_waves, int splits, void *workspace,
                          int64_t workspace_bytes, void *stream) {
  clear_error();
  int rc = validate_common("grav_stage1_f32_split", posm, acc, n, G, eps, 16);
  if (rc != GRAV_OK) return rc;
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) split_default_cfg(n, false, K, S, P);
  GRAV_REQUIRE(split_supported(K, S, P),
               "grav_stage1_f32_split: (i_per_lane, j_waves, splits) must be in {1, 2} x "
               "{1, 2, 4, 8} x [1, %d], or all 0 (got %d, %d, %d)", kMaxSplits, K, S, P);
  const int64_t need = split_workspace_bytes(n, P);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
               "grav_stage1_f32_split: workspace must be 256-byte aligned and hold %lld bytes",
               (long long)need);
  return stage1_f32_split_launch(static_cast<const float4 *>(posm), static_cast<float *>(acc),
                                 static_cast<double *>(workspace), static_cast<int>(n),
                                 static_cast<float>(G), static_cast<float>(eps * eps),
                                 eps == 0.0 ? 1 : 0, K, S, P, static_cast<hipStream_t>(stream));
}


int grav_stage1_f32_shard(const void *posm, void *acc, int64_t n, int64_t i0, int64_t m,
                          double G, double eps, int i_per_lane, int j_waves, int splits,
                          void *workspace, int64_t workspace_bytes, void *stream) {
  clear_error();
  const char *fn = "grav_stage1_f32_shard";
  int rc = validate_common(fn, posm, acc, n, G, eps, 16);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(i0 >= 0 && m >= 1 && i0 <= n - m,
               "%s: the shard [i0, i0 + m) must be a non-empty part of [0, n) (got i0 = %lld, "
               "m = %lld, n = %lld)", fn, (long long)i0, (long long)m, (long long)n);
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) shard_default_cfg(n, m, K, S, P);
  if (P == 1) {
    GRAV_REQUIRE((K == 1 || K == 2 || K == 4 || K == 8) &&
                     (S == 1 || S == 2 || S == 4 || S == 8 || (S == 16 && K <= 2)),
                 "%s: with splits = 1, (i_per_lane, j_waves) must be in {1, 2, 4, 8} x "
                 "{1, 2, 4, 8}, or j_waves 16 with i_per_lane 1 or 2 (got %d, %d)", fn, K, S);
  } else {
    GRAV_REQUIRE(split_supported(K, S, P),
                 "%s: (i_per_lane, j_waves, splits) must be in {1, 2, 4, 8} x {1, 2, 4, 8, 16} "
                 "x {1}, {1, 2} x {1, 2, 4, 8} x [2, %d], or all 0 (got %d, %d, %d)", fn,
                 kMaxSplits, K, S, P);
    const int64_t need = split_workspace_bytes(m, P);
    GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
                 "%s: workspace must be 256-byte aligned and hold %lld bytes", fn,
                 (long long)need);
  }
  return stage1_f32_shard_launch(static_cast<const float4 *>(posm), static_cast<float *>(acc),
                                 static_cast<double *>(workspace), static_cast<int>(n),
                                 static_cast<int>(i0), static_cast<int>(m), static_cast<float>(G),
                                 static_cast<float>(eps * eps), eps == 0.0 ? 1 : 0, K, S, P,
                                 static_cast<hipStream_t>(stream));
}

int grav_stage1_shard_cfg(int64_t n, int64_t m, int *i_per_lane, int *j_waves, int *splits) {
  clear_error();
  GRAV_REQUIRE(i_per_lane != nullptr && j_waves != nullptr && splits != nullptr,
               "grav_stage1_shard_cfg: NULL output");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies && m >= 1 && m <= n,
               "grav_stage1_shard_cfg: need 1 <= m <= n <= 2^29 (got m = %lld, n = %lld)",
               (long long)m, (long long)n);
  shard_default_cfg(n, m, *i_per_lane, *j_waves, *splits);
  return GRAV_OK;
}

int grav_step_split_f64m(void *posm, void *r64, void *v64, int64_t n, double G, double eps,
                         double dt, int i_per_lane, int j_waves, int splits, void *workspace,
                         int64_t workspace_bytes, void *stream) {
  clear_error();
  return step_split_f64m("grav_step_split_f64m", posm, posm, r64, v64, n, G, eps, dt,
                         i_per_lane, j_waves, splits, workspace, workspace_bytes, stream);
}

int grav_step_split_f64m_io(const void *posm_in, void *posm_out, void *r64, void *v64,
                            int64_t n, double G, double eps, double dt, int i_per_lane,
                            int j_waves, int splits, void *workspace, int64_t workspace_bytes,
                            void *stream) {
  clear_error();
  return step_split_f64m("grav_step_split_f64m_io", posm_in, posm_out, r64, v64, n, G, eps, dt,
                         i_per_lane, j_waves, splits, workspace, workspace_bytes, stream);
}

int grav_stage1_f64_split(const void *posm64, void *acc64, int64_t n, double G, double eps,
                          int i_per_lane, int j_waves, int splits, void *workspace,
                          int64_t workspace_bytes, void *stream) {
  clear_error();
  int rc = validate_common("grav_stage1_f64_split", posm64, acc64, n, G, eps, 32);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(aligned(acc64, 8), "grav_stage1_f64_split: acc64 must be 8-byte aligned");
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) split_default_cfg_f64(n, K, S, P);
  GRAV_REQUIRE(split_supported_f64(K, S, P),
               "grav_stage1_f64_split: (i_per_lane, j_waves, splits) must be in {1, 2} x "
               "{4, 8} x [1, %d], or all 0 (got %d, %d, %d)", kMaxSplits, K, S, P);
  const int64_t need = split_workspace_bytes(n, P);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
               "grav_stage1_f64_split: workspace must be 256-byte aligned and hold %lld bytes",
               (long long)need);
  return stage1_f64_split_launch(static_cast<const double4 *>(posm64),
                                 static_cast<double *>(acc64), static_cast<double *>(workspace),
                                 static_cast<int>(n), G, eps * eps, eps == 0.0 ? 1 : 0, K, S, P,
                                 static_cast<hipStream_t>(stream));
}

int grav_stage1_f64(const void *posm64, void *acc64, int64_t n, double G, double eps,
                    void *stream) {
  clear_error();
  int rc = validate_common("grav_stage1_f64", posm64, acc64, n, G, eps, 32);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(aligned(acc64, 8), "grav_stage1_f64: acc64 must be 8-byte aligned");
  const int nn = static_cast<int>(n);
  const double4 *p = static_cast<const double4 *>(posm64);
  double *a = static_cast<double *>(acc64);
  hipStream_t s = static_cast<hipStream_t>(stream);
  return stage1_f64_launch(p, a, nn, G, eps * eps, eps == 0.0 ? 1 : 0, 0, s);
}

}  // extern "C"
