// Stage 1, symmetric form: every unordered pair is evaluated once and updates both bodies
// (Newton's third law), as the reference's CPU kernels do (src/gravitation/kernel/py1.py:66-69,
// np1.py:81-108, _lib1_/lib.c:72-123), but organised for wave64 without atomics:
//
//   * bodies are cut into blocks of IB = 64*K*W (a workgroup of W wavefronts, K bodies per
//     lane); the strict upper triangle of block pairs (beta < gamma) are the work items, one
//     workgroup each; pairs inside a block are left to the owner-computes kernel
//     (k_stage1_f32 in block-diagonal mode);
//   * per 64-body chunk of block gamma, every lane holds K bodies of block beta in registers and
//     one j-body (x, y, z, m and its three accumulators) that moves one lane per step with a
//     DPP wave rotate; after 64 steps each lane has met every j-body of the chunk and the
//     j-bodies are back in their home lanes.  A pair costs 16 VALU ops + 1 v_rsq_f32 for two
//     directed interactions, plus 7 rotations per K pairs;
//   * i-side sums fold into FP32 (hi, lo) TwoSum pairs once per chunk; the W wavefronts' j-side
//     sums of a chunk are added through double-buffered LDS and written as one partial;
//   * a reduction kernel adds, for every body, its diagonal-block sum and all its partials in a
//     fixed order in FP64: the result is deterministic.
#include "grav_common.h"

namespace grav {
namespace {

constexpr float kFar = 1e30f;   // padding bodies: m = 0 at a point whose r^2 overflows to inf

__device__ __forceinline__ void two_sum_sym(float &hi, float &lo, float p) {
  const float s = hi + p;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (p - bp);
  hi = s;
  lo += e;
}

__device__ __forceinline__ float rot1(float v) {
  // wave_ror:1 (DPP control 0x13C): every lane takes its neighbour's value, wave-wide.  Every
  // lane has a valid source, so bound_ctrl only spares the compiler an initial move.
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x13C, 0xF, 0xF,
                                                     true));
}

__device__ __forceinline__ float4 load_or_far(const float4 *__restrict__ posm, int b, int n) {
  return b < n ? posm[b] : make_float4(kFar, kFar, kFar, 0.0f);
}

// Item index -> block pair (beta, gamma), beta < gamma < nb, row-major over beta.
__device__ __forceinline__ void decode_item(int item, int nb, int &beta, int &gamma) {
  int b = 0, rem = item;
  while (rem >= nb - 1 - b) {
    rem -= nb - 1 - b;
    ++b;
  }
  beta = b;
  gamma = b + 1 + rem;
}

// Work item blockIdx.x: block pair (beta, gamma) of one universe of n bodies (decoded from
// the item index, items == NULL), or, for many universes, items[blockIdx.x] = {first body of
// block beta, first body of block gamma, end of block beta, end of block gamma}.  Block beta
// is never its universe's last block, so only block gamma can be short.
template <int K, int W>
__global__ __launch_bounds__(kWave *W) void k_sym_pairs(const float4 *__restrict__ posm, int n,
                                                         int nb, float eps2,
                                                         float *__restrict__ p_i,
                                                         float *__restrict__ p_j,
                                                         const int4 *__restrict__ items) {
  constexpr int IB = kWave * K * W;
  __shared__ float red[2][W][3][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int ib0, jb0, iend, jend;
  if (items != nullptr) {
    const int4 it = items[blockIdx.x];
    ib0 = it.x;
    jb0 = it.y;
    iend = it.z;
    jend = it.w;
  } else {
    int beta, gamma;
    decode_item(blockIdx.x, nb, beta, gamma);
    ib0 = beta * IB;
    jb0 = gamma * IB;
    iend = n;
    jend = n;
  }
  const size_t base = static_cast<size_t>(blockIdx.x) * 3 * IB;

  float xi[K], yi[K], zi[K], mi[K];
  float hx[K], hy[K], hz[K], lx[K], ly[K], lz[K];
  const int i0 = ib0 + wave * (K * kWave);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float4 p = load_or_far(posm, i0 + k * kWave + lane, iend);
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
    mi[k] = p.w;
    hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.0f;
  }

  for (int jc = 0; jc < K * W; ++jc) {
    const float4 q = load_or_far(posm, jb0 + jc * kWave + lane, jend);
    float xj = q.x, yj = q.y, zj = q.z, mj = q.w;
    float ajx = 0.0f, ajy = 0.0f, ajz = 0.0f;
    float px[K], py[K], pz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) px[k] = py[k] = pz[k] = 0.0f;
#pragma unroll 2
    for (int t = 0; t < kWave; ++t) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float dx = xj - xi[k];
        const float dy = yj - yi[k];
        const float dz = zj - zi[k];
        float r2 = __builtin_fmaf(dx, dx, eps2);
        r2 = __builtin_fmaf(dy, dy, r2);
        r2 = __builtin_fmaf(dz, dz, r2);
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv3 = inv * (inv * inv);
        const float si = mj * inv3;
        const float sj = mi[k] * inv3;
        px[k] = __builtin_fmaf(dx, si, px[k]);
        py[k] = __builtin_fmaf(dy, si, py[k]);
        pz[k] = __builtin_fmaf(dz, si, pz[k]);
        ajx = __builtin_fmaf(-dx, sj, ajx);
        ajy = __builtin_fmaf(-dy, sj, ajy);
        ajz = __builtin_fmaf(-dz, sj, ajz);
      }
      xj = rot1(xj);
      yj = rot1(yj);
      zj = rot1(zj);
      mj = rot1(mj);
      ajx = rot1(ajx);
      ajy = rot1(ajy);
      ajz = rot1(ajz);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      two_sum_sym(hx[k], lx[k], px[k]);
      two_sum_sym(hy[k], ly[k], py[k]);
      two_sum_sym(hz[k], lz[k], pz[k]);
    }
    // After 64 rotations every j-body and its sums are back in their home lane.
    float(*r)[3][kWave] = red[jc & 1];
    r[wave][0][lane] = ajx;
    r[wave][1][lane] = ajy;
    r[wave][2][lane] = ajz;
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * kWave; t += kWave * W) {
      const int c = t / kWave, l = t % kWave;
      float s = 0.0f;
#pragma unroll
      for (int w = 0; w < W; ++w) s += r[w][c][l];
      p_j[base + static_cast<size_t>(c) * IB + jc * kWave + l] = s;
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int bl = wave * (K * kWave) + k * kWave + lane;
    p_i[base + bl] = hx[k] + lx[k];
    p_i[base + IB + bl] = hy[k] + ly[k];
    p_i[base + 2 * IB + bl] = hz[k] + lz[k];
  }
}

// ---------------------------------------------------------------------------------------------
// FP64 form: the same block pairs and wave rotation, each double moved as two 32-bit DPP
// words; sums in FP64 (no compensation needed), 1/sqrt by rsqrt_nr64.  Padding bodies sit at
// 1e30 with m = 0: r^2 = 3e60 stays finite in FP64, and every term they add is 0 (m_j = 0) or
// lands in a padding body's own partial, which is never read.
__device__ __forceinline__ double rot1d(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), 0x13C, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), 0x13C, 0xF, 0xF, true);
  return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) |
                                        static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double4 load_or_far_d(const double4 *__restrict__ posm, int b, int n) {
  return b < n ? posm[b] : make_double4(1e30, 1e30, 1e30, 0.0);
}

template <int K, int W>
__global__ __launch_bounds__(kWave *W) void k_sym_pairs_f64(const double4 *__restrict__ posm,
                                                             int n, int nb, double eps2,
                                                             double *__restrict__ p_i,
                                                             double *__restrict__ p_j) {
  constexpr int IB = kWave * K * W;
  __shared__ double red[2][W][3][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int beta, gamma;
  decode_item(blockIdx.x, nb, beta, gamma);
  const size_t base = static_cast<size_t>(blockIdx.x) * 3 * IB;

  double xi[K], yi[K], zi[K], mi[K], px[K], py[K], pz[K];
  const int i0 = beta * IB + wave * (K * kWave);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double4 p = load_or_far_d(posm, i0 + k * kWave + lane, n);
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
    mi[k] = p.w;
    px[k] = py[k] = pz[k] = 0.0;
  }

  const int j0 = gamma * IB;
  for (int jc = 0; jc < K * W; ++jc) {
    const double4 q = load_or_far_d(posm, j0 + jc * kWave + lane, n);
    double xj = q.x, yj = q.y, zj = q.z, mj = q.w;
    double ajx = 0.0, ajy = 0.0, ajz = 0.0;
    for (int t = 0; t < kWave; ++t) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double dx = xj - xi[k];
        const double dy = yj - yi[k];
        const double dz = zj - zi[k];
        double r2 = __builtin_fma(dx, dx, eps2);
        r2 = __builtin_fma(dy, dy, r2);
        r2 = __builtin_fma(dz, dz, r2);
        const double inv = rsqrt_nr64(r2);
        const double inv3 = inv * (inv * inv);
        const double si = mj * inv3;
        const double sj = mi[k] * inv3;
        px[k] = __builtin_fma(dx, si, px[k]);
        py[k] = __builtin_fma(dy, si, py[k]);
        pz[k] = __builtin_fma(dz, si, pz[k]);
        ajx = __builtin_fma(-dx, sj, ajx);
        ajy = __builtin_fma(-dy, sj, ajy);
        ajz = __builtin_fma(-dz, sj, ajz);
      }
      xj = rot1d(xj);
      yj = rot1d(yj);
      zj = rot1d(zj);
      mj = rot1d(mj);
      ajx = rot1d(ajx);
      ajy = rot1d(ajy);
      ajz = rot1d(ajz);
    }
    // After 64 rotations every j-body and its sums are back in their home lane.
    double(*r)[3][kWave] = red[jc & 1];
    r[wave][0][lane] = ajx;
    r[wave][1][lane] = ajy;
    r[wave][2][lane] = ajz;
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * kWave; t += kWave * W) {
      const int c = t / kWave, l = t % kWave;
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < W; ++w) s += r[w][c][l];
      p_j[base + static_cast<size_t>(c) * IB + jc * kWave + l] = s;
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int bl = wave * (K * kWave) + k * kWave + lane;
    p_i[base + bl] = px[k];
    p_i[base + IB + bl] = py[k];
    p_i[base + 2 * IB + bl] = pz[k];
  }
}

// acc[c][b] = G * (diag[c][b] + sum_{gamma > beta} p_i[(beta, gamma)][c][b'] +
//                  sum_{alpha < beta} p_j[(alpha, beta)][c][b']),  in FP64, fixed order.
// One thread per (body, component), blockIdx.y = component.  The loads of a body's partials
// are issued eight at a time ahead of their (in-order) additions, so a wave keeps eight
// coalesced 256-byte reads in flight instead of one.
template <typename T>
__global__ __launch_bounds__(256) void k_sym_reduce(const T *__restrict__ diag,
                                                    const T *__restrict__ p_i,
                                                    const T *__restrict__ p_j,
                                                    T *__restrict__ acc, int n, int nb, int ib,
                                                    double G) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (b >= n) return;
  const int beta = b / ib;
  const int bl = b - beta * ib;
  const size_t comp = static_cast<size_t>(c) * ib + bl;
  const size_t item_stride = static_cast<size_t>(3) * ib;
  double s = diag[static_cast<size_t>(c) * n + b];
  // Row beta's items (beta, gamma), gamma = beta+1..nb-1, are consecutive from
  // row(beta) = sum_{q < beta} (nb - 1 - q).
  const int64_t r0 = int64_t(beta) * (nb - 1) - (int64_t(beta) * (beta - 1)) / 2;
  const T *q = p_i + static_cast<size_t>(r0) * item_stride + comp;
  const int cnt = nb - 1 - beta;
  int g = 0;
  for (; g + 8 <= cnt; g += 8) {
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = q[static_cast<size_t>(g + k) * item_stride];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; g < cnt; ++g) s += q[static_cast<size_t>(g) * item_stride];
  // Items (alpha, beta), alpha = 0..beta-1: index row(alpha) + beta - alpha - 1, which
  // advances by nb - 2 - alpha from alpha to alpha + 1.
  int64_t idx = beta - 1;
  int a = 0;
  for (; a + 8 <= beta; a += 8) {
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[k] = p_j[static_cast<size_t>(idx) * item_stride + comp];
      idx += nb - 2 - (a + k);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; a < beta; ++a) {
    s += p_j[static_cast<size_t>(idx) * item_stride + comp];
    idx += nb - 2 - a;
  }
  acc[static_cast<size_t>(c) * n + b] = static_cast<T>(G * s);
}

// Batched reduce: blockIdx.x = block b of some universe, blocks[b] = {first body, end,
// universe's first item, nb of the universe, beta}; the universe's items are its block pairs
// in row-major order from the first item.  blockIdx.y = component.
__global__ __launch_bounds__(256) void k_sym_reduce_batched(const float *__restrict__ diag,
                                                            const float *__restrict__ p_i,
                                                            const float *__restrict__ p_j,
                                                            float *__restrict__ acc,
                                                            const SymBlock *__restrict__ blocks,
                                                            int n, int ib, float G) {
  const SymBlock blk = blocks[blockIdx.x];
  const int c = blockIdx.y;
  const int nb = blk.nb, beta = blk.beta;
  const size_t item_stride = static_cast<size_t>(3) * ib;
  const int64_t r0 = blk.item_base + int64_t(beta) * (nb - 1) - (int64_t(beta) * (beta - 1)) / 2;
  for (int b = blk.start + static_cast<int>(threadIdx.x); b < blk.end; b += 256) {
    const size_t comp = static_cast<size_t>(c) * ib + (b - blk.start);
    double s = diag[static_cast<size_t>(c) * n + b];
    const float *q = p_i + static_cast<size_t>(r0) * item_stride + comp;
    for (int g = 0; g < nb - 1 - beta; ++g) s += q[static_cast<size_t>(g) * item_stride];
    int64_t idx = blk.item_base + beta - 1;
    for (int a = 0; a < beta; ++a) {
      s += p_j[static_cast<size_t>(idx) * item_stride + comp];
      idx += nb - 2 - a;
    }
    acc[static_cast<size_t>(c) * n + b] = static_cast<float>(static_cast<double>(G) * s);
  }
}

struct SymLayout {
  int ib, nb;
  int64_t items, diag_bytes, part_bytes, total;
};

SymLayout sym_layout(int64_t n, int K, int W, int64_t elem = 4) {
  SymLayout L{};
  L.ib = kWave * K * W;
  L.nb = static_cast<int>((n + L.ib - 1) / L.ib);
  L.items = static_cast<int64_t>(L.nb) * (L.nb - 1) / 2;
  auto up = [](int64_t x) { return (x + 255) & ~int64_t(255); };
  L.diag_bytes = up(int64_t(3) * n * elem);
  L.part_bytes = up(L.items * 3 * L.ib * elem);
  L.total = L.diag_bytes + 2 * L.part_bytes;
  return L;
}

bool sym_supported(int K, int W) {
  return (K == 4 || K == 8) && (W == 2 || W == 4);
}

// (0, 0) selects the shape by n, from MI355X measurements on Plummer spheres: blocks of 512
// (K 4, W 2) below 57344 bodies, where blocks of 1024 leave too few block pairs to fill 256
// CUs (at 49152 bodies (4, 2) takes 0.60 ms against 0.77), and blocks of 1024 (K 8, W 2)
// above (1.34x over owner-computes at 2^18, whose own default is (2, 8)).
void sym_shape(int64_t n, int &K, int &W) {
  if (K == 0 && W == 0) {
    K = n < 57344 ? 4 : 8;
    W = 2;
  }
}

template <int K, int W>
int launch_pairs(const float4 *posm, int n, int nb, float eps2, float *pi, float *pj,
                 int64_t items, hipStream_t s) {
  if (items == 0) return GRAV_OK;
  hipLaunchKernelGGL((k_sym_pairs<K, W>), dim3(static_cast<unsigned>(items)), dim3(kWave * W),
                     0, s, posm, n, nb, eps2, pi, pj, static_cast<const int4 *>(nullptr));
  return check_launch("k_sym_pairs");
}

bool sym_supported_f64(int K, int W) {
  return (K == 2 || K == 4) && (W == 2 || W == 4);
}

// FP64 (0, 0), from MI355X measurements on Plummer spheres: blocks of 256 (K 2, W 2) below
// 24576 bodies, of 512 (K 4, W 2) below 65536, of 1024 (K 4, W 4) from there (1.37x over
// grav_stage1_f64 at 2^12, 1.7-2.1x at 2^15..2^18).
void sym_shape_f64(int64_t n, int &K, int &W) {
  if (K == 0 && W == 0) {
    K = n < 24576 ? 2 : 4;
    W = n < 65536 ? 2 : 4;
  }
}

template <int K, int W>
int launch_pairs_f64(const double4 *posm, int n, int nb, double eps2, double *pi, double *pj,
                     int64_t items, hipStream_t s) {
  if (items == 0) return GRAV_OK;
  hipLaunchKernelGGL((k_sym_pairs_f64<K, W>), dim3(static_cast<unsigned>(items)),
                     dim3(kWave * W), 0, s, posm, n, nb, eps2, pi, pj);
  return check_launch("k_sym_pairs_f64");
}

template <int K, int W>
int launch_pairs_items(const float4 *posm, float eps2, float *pi, float *pj, const int4 *items,
                       int nitems, hipStream_t s) {
  if (nitems == 0) return GRAV_OK;
  hipLaunchKernelGGL((k_sym_pairs<K, W>), dim3(static_cast<unsigned>(nitems)), dim3(kWave * W),
                     0, s, posm, 0, 0, eps2, pi, pj, items);
  return check_launch("k_sym_pairs");
}

}  // namespace

bool sym_batched_supported(int K, int W) {
  return (K == 4 || K == 8) && (W == 1 || W == 2 || W == 4);
}

int sym_batched_launch(const float4 *posm, float *acc, int n, float G, float eps2, int mask,
                       int K, int W, const int4 *pair_items, int npairs, const int4 *diag_items,
                       int ndiag, const SymBlock *blocks, int nblocks, float *diag, float *pi,
                       float *pj, hipStream_t s) {
  if (!sym_batched_supported(K, W))
    return fail(GRAV_EINVAL, "symmetric batched shape (%d, %d) not supported", K, W);
  // 1. pairs inside each block: owner-computes over tiles of 256 bodies against their block
  int rc = stage1_f32_launch(posm, diag, n, 1.0f, eps2, mask, 4, 4, 256, diag_items, ndiag, s);
  if (rc != GRAV_OK) return rc;
  // 2. pairs between the blocks of each universe, each once
  if (K == 4 && W == 1) rc = launch_pairs_items<4, 1>(posm, eps2, pi, pj, pair_items, npairs, s);
  if (K == 4 && W == 2) rc = launch_pairs_items<4, 2>(posm, eps2, pi, pj, pair_items, npairs, s);
  if (K == 4 && W == 4) rc = launch_pairs_items<4, 4>(posm, eps2, pi, pj, pair_items, npairs, s);
  if (K == 8 && W == 1) rc = launch_pairs_items<8, 1>(posm, eps2, pi, pj, pair_items, npairs, s);
  if (K == 8 && W == 2) rc = launch_pairs_items<8, 2>(posm, eps2, pi, pj, pair_items, npairs, s);
  if (K == 8 && W == 4) rc = launch_pairs_items<8, 4>(posm, eps2, pi, pj, pair_items, npairs, s);
  if (rc != GRAV_OK) return rc;
  // 3. per body: diagonal + partials of its universe, FP64, fixed order
  if (nblocks == 0) return GRAV_OK;
  hipLaunchKernelGGL(k_sym_reduce_batched, dim3(static_cast<unsigned>(nblocks), 3), dim3(256), 0,
                     s, diag, pi, pj, acc, blocks, n, kWave * K * W, G);
  return check_launch("k_sym_reduce_batched");
}

}  // namespace grav

using namespace grav;

extern "C" {

int grav_stage1_sym_workspace(int64_t n, int i_chunks_per_wave, int waves, int64_t *bytes) {
  clear_error();
  GRAV_REQUIRE(bytes != nullptr, "grav_stage1_sym_workspace: bytes is NULL");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_stage1_sym_workspace: n must be in [1, 2^29] (got %lld)", (long long)n);
  int K = i_chunks_per_wave, W = waves;
  sym_shape(n, K, W);
  GRAV_REQUIRE(sym_supported(K, W),
               "grav_stage1_sym_workspace: (i_chunks_per_wave, waves) must be (0, 0) or in "
               "{4,8} x {2,4}");
  *bytes = sym_layout(n, K, W).total;
  return GRAV_OK;
}

int grav_stage1_sym_f32(const void *posm, void *acc, int64_t n, double G, double eps,
                        int i_chunks_per_wave, int waves, void *workspace,
                        int64_t workspace_bytes, void *stream) {
  clear_error();
  GRAV_REQUIRE(posm != nullptr && acc != nullptr, "grav_stage1_sym_f32: NULL buffer");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_stage1_sym_f32: n must be in [1, 2^29] (got %lld)", (long long)n);
  GRAV_REQUIRE(aligned(posm, 16) && aligned(acc, 4), "grav_stage1_sym_f32: misaligned");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "grav_stage1_sym_f32: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_stage1_sym_f32: eps must be >= 0");
  int K = i_chunks_per_wave, W = waves;
  sym_shape(n, K, W);
  GRAV_REQUIRE(sym_supported(K, W),
               "grav_stage1_sym_f32: (i_chunks_per_wave, waves) must be (0, 0) or in "
               "{4,8} x {2,4}");
  const SymLayout L = sym_layout(n, K, W);
  GRAV_REQUIRE(L.items <= 0x7fffffff, "grav_stage1_sym_f32: n too large for this block size");
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= L.total,
               "grav_stage1_sym_f32: workspace must be 256-byte aligned and hold %lld bytes",
               (long long)L.total);
  char *ws = static_cast<char *>(workspace);
  float *diag = reinterpret_cast<float *>(ws);
  float *pi = reinterpret_cast<float *>(ws + L.diag_bytes);
  float *pj = reinterpret_cast<float *>(ws + L.diag_bytes + L.part_bytes);
  const int nn = static_cast<int>(n);
  const float eps2 = static_cast<float>(eps * eps);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float4 *p = static_cast<const float4 *>(posm);
  // 1. pairs inside each block: owner-computes, blocks of IB bodies (IB is a multiple of 256)
  int rc = stage1_f32_launch(p, diag, nn, 1.0f, eps2, eps == 0.0 ? 1 : 0, 4, 4, 256, nullptr,
                             L.ib, s);
  if (rc != GRAV_OK) return rc;
  // 2. pairs between blocks, each once
  if (K == 4 && W == 2) rc = launch_pairs<4, 2>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (K == 4 && W == 4) rc = launch_pairs<4, 4>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (K == 8 && W == 2) rc = launch_pairs<8, 2>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (K == 8 && W == 4) rc = launch_pairs<8, 4>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (rc != GRAV_OK) return rc;
  // 3. per body: diagonal + partials, FP64, fixed order
  hipLaunchKernelGGL((k_sym_reduce<float>), dim3((nn + 255) / 256, 3), dim3(256), 0, s, diag,
                     pi, pj, static_cast<float *>(acc), nn, L.nb, L.ib, G);
  return check_launch("k_sym_reduce");
}

int grav_stage1_sym_f64_workspace(int64_t n, int i_chunks_per_wave, int waves,
                                  int64_t *bytes) {
  clear_error();
  GRAV_REQUIRE(bytes != nullptr, "grav_stage1_sym_f64_workspace: bytes is NULL");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_stage1_sym_f64_workspace: n must be in [1, 2^29] (got %lld)",
               (long long)n);
  int K = i_chunks_per_wave, W = waves;
  sym_shape_f64(n, K, W);
  GRAV_REQUIRE(sym_supported_f64(K, W),
               "grav_stage1_sym_f64_workspace: (i_chunks_per_wave, waves) must be (0, 0) or in "
               "{2,4} x {2,4}");
  *bytes = sym_layout(n, K, W, 8).total;
  return GRAV_OK;
}

int grav_stage1_sym_f64(const void *posm64, void *acc64, int64_t n, double G, double eps,
                        int i_chunks_per_wave, int waves, void *workspace,
                        int64_t workspace_bytes, void *stream) {
  clear_error();
  GRAV_REQUIRE(posm64 != nullptr && acc64 != nullptr, "grav_stage1_sym_f64: NULL buffer");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_stage1_sym_f64: n must be in [1, 2^29] (got %lld)", (long long)n);
  GRAV_REQUIRE(aligned(posm64, 32) && aligned(acc64, 8), "grav_stage1_sym_f64: misaligned");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "grav_stage1_sym_f64: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_stage1_sym_f64: eps must be >= 0");
  int K = i_chunks_per_wave, W = waves;
  sym_shape_f64(n, K, W);
  GRAV_REQUIRE(sym_supported_f64(K, W),
               "grav_stage1_sym_f64: (i_chunks_per_wave, waves) must be (0, 0) or in "
               "{2,4} x {2,4}");
  const SymLayout L = sym_layout(n, K, W, 8);
  GRAV_REQUIRE(L.items <= 0x7fffffff, "grav_stage1_sym_f64: n too large for this block size");
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= L.total,
               "grav_stage1_sym_f64: workspace must be 256-byte aligned and hold %lld bytes",
               (long long)L.total);
  char *ws = static_cast<char *>(workspace);
  double *diag = reinterpret_cast<double *>(ws);
  double *pi = reinterpret_cast<double *>(ws + L.diag_bytes);
  double *pj = reinterpret_cast<double *>(ws + L.diag_bytes + L.part_bytes);
  const int nn = static_cast<int>(n);
  const double eps2 = eps * eps;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double4 *p = static_cast<const double4 *>(posm64);
  // 1. pairs inside each block (IB is a multiple of 256)
  int rc = stage1_f64_launch(p, diag, nn, 1.0, eps2, eps == 0.0 ? 1 : 0, L.ib, s);
  if (rc != GRAV_OK) return rc;
  // 2. pairs between blocks, each once
  if (K == 2 && W == 2) rc = launch_pairs_f64<2, 2>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (K == 2 && W == 4) rc = launch_pairs_f64<2, 4>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (K == 4 && W == 2) rc = launch_pairs_f64<4, 2>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (K == 4 && W == 4) rc = launch_pairs_f64<4, 4>(p, nn, L.nb, eps2, pi, pj, L.items, s);
  if (rc != GRAV_OK) return rc;
  // 3. per body: diagonal + partials, FP64, fixed order
  hipLaunchKernelGGL((k_sym_reduce<double>), dim3((nn + 255) / 256, 3), dim3(256), 0, s, diag,
                     pi, pj, static_cast<double *>(acc64), nn, L.nb, L.ib, G);
  return check_launch("k_sym_reduce");
}

}  // extern "C"
