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
//   * for one universe, each row of block pairs is cut into segments of consecutive gamma
//     blocks (one workgroup each, the i-side sums kept in registers along the segment), and
//     the rows are walked in slabs whose partials fit the caller's workspace: after every slab
//     a fold kernel adds its partials into each body's FP64 sum in a fixed order, so the
//     workspace is bounded (one slab down to one row) and the result is deterministic;
//   * batched universes list their block pairs as items, and one reduction kernel adds every
//     body's diagonal-block sum and partials in a fixed order in FP64.
#include <algorithm>
#include <utility>
#include <vector>

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

// One block pair (beta, gamma) for one workgroup: its wavefronts hold K block-beta bodies per
// lane (xi..mi) and take the K*W chunks of 64 block-gamma bodies from jb0 (bodies at or past
// jend load as far, massless padding).  The i-side sums of every chunk fold into the TwoSum
// pairs (hx, lx) .. (hz, lz), which the caller may carry over several gamma blocks; the j-side
// sums of a chunk, added over the W wavefronts through the double-buffered LDS `red` (buffer
// chunk & 1, chunk counting on across calls), go to pj[c * IB + jc * 64 + l], 3 * IB floats.
template <int K, int W>
__device__ __forceinline__ void sym_block_pair(const float4 *__restrict__ posm, int jb0,
                                               int jend, float eps2, const float (&xi)[K],
                                               const float (&yi)[K], const float (&zi)[K],
                                               const float (&mi)[K], float (&hx)[K],
                                               float (&hy)[K], float (&hz)[K], float (&lx)[K],
                                               float (&ly)[K], float (&lz)[K],
                                               float *__restrict__ pj,
                                               float (&red)[2][W][3][kWave], int &chunk,
                                               int lane, int wave) {
  constexpr int IB = kWave * K * W;
  for (int jc = 0; jc < K * W; ++jc, ++chunk) {
    const float4 q = load_or_far(posm, jb0 + jc * kWave + lane, jend);
    float xj = q.x, yj = q.y, zj = q.z, mj = q.w;
    float ajx = 0.0f, ajy = 0.0f, ajz = 0.0f;
    float px[K], py[K], pz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) px[k] = py[k] = pz[k] = 0.0f;
    // not unrolled: the 169-instruction body of one step ran 4 % faster than the 318 of two
    // steps (MI355X, 2^18: 13.36 against 13.96 ms; 2^20: 212 against 221 ms)
#pragma unroll 1
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
    // After 64 rotations every j-body and its sums are back in their home lane.  The next
    // chunk writes the other buffer, and a wave can only come back to this one after the
    // next chunk's barrier, which every wave reaches after it has read this one.
    float(*r)[3][kWave] = red[chunk & 1];
    r[wave][0][lane] = ajx;
    r[wave][1][lane] = ajy;
    r[wave][2][lane] = ajz;
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * kWave; t += kWave * W) {
      const int c = t / kWave, l = t % kWave;
      float s = 0.0f;
#pragma unroll
      for (int w = 0; w < W; ++w) s += r[w][c][l];
      pj[static_cast<size_t>(c) * IB + jc * kWave + l] = s;
    }
  }
}

// Batched work item blockIdx.x: items[blockIdx.x] = {first body of block beta, first body of
// block gamma, end of block beta, end of block gamma} of one universe.  Block beta is never
// its universe's last block, so only block gamma can be short.  Writes the i-side partial to
// p_i and the j-side partial to p_j at item * 3 * IB.
template <int K, int W>
__global__ __launch_bounds__(kWave *W) void k_sym_pairs(const float4 *__restrict__ posm,
                                                         float eps2, float *__restrict__ p_i,
                                                         float *__restrict__ p_j,
                                                         const int4 *__restrict__ items) {
  constexpr int IB = kWave * K * W;
  __shared__ float red[2][W][3][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int4 it = items[blockIdx.x];
  const size_t base = static_cast<size_t>(blockIdx.x) * 3 * IB;

  float xi[K], yi[K], zi[K], mi[K];
  float hx[K], hy[K], hz[K], lx[K], ly[K], lz[K];
  const int i0 = it.x + wave * (K * kWave);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float4 p = load_or_far(posm, i0 + k * kWave + lane, it.z);
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
    mi[k] = p.w;
    hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.0f;
  }
  int chunk = 0;
  sym_block_pair<K, W>(posm, it.y, it.w, eps2, xi, yi, zi, mi, hx, hy, hz, lx, ly, lz,
                       p_j + base, red, chunk, lane, wave);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int bl = wave * (K * kWave) + k * kWave + lane;
    p_i[base + bl] = hx[k] + lx[k];
    p_i[base + IB + bl] = hy[k] + ly[k];
    p_i[base + 2 * IB + bl] = hz[k] + lz[k];
  }
}

// ---------------------------------------------------------------------------------------------
// One universe: row segments, processed in slabs of rows.
//
// Block row beta (0 <= beta < nb - 1) pairs with gamma = beta+1 .. nb-1.  Its r = nb-1-beta
// block pairs are cut into segments of seglen = 2^lg consecutive gamma blocks (the last one
// shorter), numbered row-major over all rows.  A segment is one workgroup: the i-side sums of
// its block-beta bodies stay in registers over all of its gamma blocks and are written once
// (p_i, 3 * IB values per segment), the j-side sums go out per block pair (p_j, seglen slots of
// 3 * IB values per segment; the slots past a short last segment stay unused).  A slab is a
// range of consecutive rows whose segments fit the workspace at once; after each slab a fold
// kernel adds that slab's partials into every body's FP64 sum in a fixed order, so results are
// deterministic for a given workspace size.

// Segments of the rows of fewer than x block pairs: sum_{m=1}^{x} ceil(m / 2^lg).
__host__ __device__ inline int64_t seg_prefix(int64_t x, int lg) {
  const int64_t a = x >> lg, b = x - (a << lg);
  return (a * (a + 1) / 2 << lg) + b * (a + 1);
}
// First segment of row beta (rows 0 .. beta-1 hold nb-1 .. nb-beta block pairs).
__host__ __device__ inline int64_t seg_base(int beta, int nb, int lg) {
  return seg_prefix(nb - 1, lg) - seg_prefix(nb - 1 - beta, lg);
}
__host__ __device__ inline int seg_count(int beta, int nb, int lg) {
  return static_cast<int>((static_cast<int64_t>(nb - 1 - beta) + (1 << lg) - 1) >> lg);
}
// Row of global segment s, among rows [beta_lo, nb - 1): the last row whose first segment is
// at or before s (binary search; seg_base is increasing in beta).
__device__ inline int seg_row(int64_t s, int beta_lo, int nb, int lg) {
  int lo = beta_lo, hi = nb - 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg_base(mid, nb, lg) <= s)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The segments of one launch: n1 from segment first (rows from beta_lo on), then the rest from
// segment first2 (rows from beta_lo2 on); the slab-local index is the launch index.
struct SegSpan {
  int64_t first, first2;
  int beta_lo, beta_lo2, n1, pad;
};

__device__ __forceinline__ void span_segment(const SegSpan &sp, int b, int nb, int lg,
                                             int64_t &seg, int &beta) {
  const bool second = b >= sp.n1;
  seg = second ? sp.first2 + (b - sp.n1) : sp.first + b;
  beta = seg_row(seg, second ? sp.beta_lo2 : sp.beta_lo, nb, lg);
}

// blockIdx.x: the blockIdx.x-th segment of the span; its partials go to p_i / p_j at the
// slab-local index blockIdx.x.
template <int K, int W>
__global__ __launch_bounds__(kWave *W) void k_sym_segments(const float4 *__restrict__ posm,
                                                            int n, int nb, int lg,
                                                            SegSpan span, float eps2,
                                                            float *__restrict__ p_i,
                                                            float *__restrict__ p_j) {
  constexpr int IB = kWave * K * W;
  __shared__ float red[2][W][3][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int64_t seg;
  int beta;
  span_segment(span, static_cast<int>(blockIdx.x), nb, lg, seg, beta);
  const int g0 = beta + 1 + (static_cast<int>(seg - seg_base(beta, nb, lg)) << lg);
  const int g1 = min(nb, g0 + (1 << lg));
  const size_t local = blockIdx.x;

  float xi[K], yi[K], zi[K], mi[K];
  float hx[K], hy[K], hz[K], lx[K], ly[K], lz[K];
  const int i0 = beta * IB + wave * (K * kWave);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float4 p = load_or_far(posm, i0 + k * kWave + lane, n);
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
    mi[k] = p.w;
    hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.0f;
  }
  int chunk = 0;
  for (int gamma = g0; gamma < g1; ++gamma) {
    float *pj = p_j + ((local << lg) + (gamma - g0)) * (3 * IB);
    sym_block_pair<K, W>(posm, gamma * IB, n, eps2, xi, yi, zi, mi, hx, hy, hz, lx, ly, lz,
                         pj, red, chunk, lane, wave);
  }
  const size_t base = local * 3 * IB;
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

// A segment of the FP64 form (the row-segment scheme of k_sym_segments, FP64 sums).
template <int K, int W>
__global__ __launch_bounds__(kWave *W) void k_sym_segments_f64(const double4 *__restrict__ posm,
                                                                int n, int nb, int lg,
                                                                SegSpan span, double eps2,
                                                                double *__restrict__ p_i,
                                                                double *__restrict__ p_j) {
  constexpr int IB = kWave * K * W;
  __shared__ double red[2][W][3][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int64_t seg;
  int beta;
  span_segment(span, static_cast<int>(blockIdx.x), nb, lg, seg, beta);
  const int g0 = beta + 1 + (static_cast<int>(seg - seg_base(beta, nb, lg)) << lg);
  const int g1 = min(nb, g0 + (1 << lg));
  const size_t local = blockIdx.x;

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

  int chunk = 0;
  for (int gamma = g0; gamma < g1; ++gamma) {
    double *pj = p_j + ((local << lg) + (gamma - g0)) * (3 * IB);
    for (int jc = 0; jc < K * W; ++jc, ++chunk) {
      const double4 q = load_or_far_d(posm, gamma * IB + jc * kWave + lane, n);
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
      // After 64 rotations every j-body and its sums are back in their home lane; buffers
      // alternate with the chunk count as in sym_block_pair.
      double(*r)[3][kWave] = red[chunk & 1];
      r[wave][0][lane] = ajx;
      r[wave][1][lane] = ajy;
      r[wave][2][lane] = ajz;
      __syncthreads();
      for (int t = threadIdx.x; t < 3 * kWave; t += kWave * W) {
        const int c = t / kWave, l = t % kWave;
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w) sum += r[w][c][l];
        pj[static_cast<size_t>(c) * IB + jc * kWave + l] = sum;
      }
    }
  }
  const size_t base = local * 3 * IB;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int bl = wave * (K * kWave) + k * kWave + lane;
    p_i[base + bl] = px[k];
    p_i[base + IB + bl] = py[k];
    p_i[base + 2 * IB + bl] = pz[k];
  }
}

// Fold of one slab into every body's FP64 sum s, one thread per (body, component),
// blockIdx.y = component.  The slab is rows [beta0, beta1) (slab-local segments from 0) and
// rows [beta0b, beta1b) (slab-local segments from nseg1; empty for one universe):
//   s  = the first slab: the body's block-diagonal sum if its block lies in [d_lo, d_hi)
//        (diag holds blocks d_lo.. as rows of stride dstride), else 0; later slabs: acc64
//   s += the body's i-side partials, if its block's row lies in the slab (its segments in order)
//   s += its j-side partials from every slab row beta < its block, in row order
// and then acc = G * s (last slab) or acc64 = s.  The order is fixed, so the result is
// deterministic.  Loads are issued eight at a time ahead of their in-order additions, so a
// wave keeps eight coalesced reads in flight.
template <typename T>
__global__ __launch_bounds__(256) void k_sym_fold(const T *__restrict__ diag,
                                                  double *__restrict__ acc64,
                                                  T *__restrict__ acc,
                                                  const T *__restrict__ p_i,
                                                  const T *__restrict__ p_j, int n, int nb,
                                                  int ib, int lg, int beta0, int beta1,
                                                  int beta0b, int beta1b, int nseg1,
                                                  int first, int last, double G, int d_lo,
                                                  int d_hi, int dstride) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (b >= n) return;
  const int row = b / ib;
  const size_t comp = static_cast<size_t>(c) * ib + (b - row * ib);
  const size_t stride = static_cast<size_t>(3) * ib;
  const size_t at = static_cast<size_t>(c) * n + b;
  double s;
  if (first)
    s = (row >= d_lo && row < d_hi)
            ? static_cast<double>(diag[static_cast<size_t>(c) * dstride + (b - d_lo * ib)])
            : 0.0;
  else
    s = acc64[at];
  int64_t ibase = -1;   // slab-local first segment of the body's row, if the slab holds it
  if (row >= beta0 && row < beta1)
    ibase = seg_base(row, nb, lg) - seg_base(beta0, nb, lg);
  else if (row >= beta0b && row < beta1b)
    ibase = nseg1 + seg_base(row, nb, lg) - seg_base(beta0b, nb, lg);
  if (ibase >= 0) {
    const T *q = p_i + static_cast<size_t>(ibase) * stride + comp;
    const int cnt = seg_count(row, nb, lg);
    int g = 0;
    for (; g + 8 <= cnt; g += 8) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = q[static_cast<size_t>(g + k) * stride];
#pragma unroll
      for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; g < cnt; ++g) s += q[static_cast<size_t>(g) * stride];
  }
  // rows beta < row of each range: pair (beta, row) sits in slot (row - beta - 1) of row
  // beta's segments
  for (int r = 0; r < 2; ++r) {
    const int top = min(r == 0 ? beta1 : beta1b, row);
    int64_t sb = r == 0 ? 0 : nseg1;   // slab-local first segment of row beta
    int beta = r == 0 ? beta0 : beta0b;
    for (; beta + 8 <= top; beta += 8) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int d = row - (beta + k) - 1;
        const int64_t slot = ((sb + (d >> lg)) << lg) + (d & ((1 << lg) - 1));
        v[k] = p_j[static_cast<size_t>(slot) * stride + comp];
        sb += seg_count(beta + k, nb, lg);
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; beta < top; ++beta) {
      const int d = row - beta - 1;
      const int64_t slot = ((sb + (d >> lg)) << lg) + (d & ((1 << lg) - 1));
      s += p_j[static_cast<size_t>(slot) * stride + comp];
      sb += seg_count(beta, nb, lg);
    }
  }
  if (last)
    acc[at] = static_cast<T>(G * s);
  else
    acc64[at] = s;
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

// Segment length 2^lg: the longest, up to 8 gamma blocks, that still leaves kSegTarget
// segments (a segment is one workgroup of W wavefronts, about 2048 of which are resident on
// 256 CUs at once).  Longer segments write fewer i-side partials but leave a coarser last
// round of workgroups.  Stage-1 times on an MI355X (Plummer, default shapes, ms): 2^18 13.98
// (1 pair per segment) / 14.11 (2) / 14.93 (4); 2^19 55.4 (1) / 55.6 (2) / 57.7 (8); 2^20
// 221.4 (1) / 223.2 (8); FP64 2^18 24.4 (1) / 24.4 (2) / 25.2 (4).  So single pairs are kept
// up to 2^20 bodies, and segments grow only where their count runs into the millions and the
// workspace they save is in the tens of GB.
constexpr int64_t kSegTarget = int64_t(1) << 18;
int seg_lg(int nb) {
  const int64_t pairs = static_cast<int64_t>(nb) * (nb - 1) / 2;
  int lg = 3;
  while (lg > 0 && (pairs >> lg) < kSegTarget) --lg;
  return lg;
}

inline int64_t up256(int64_t x) { return (x + 255) & ~int64_t(255); }

// Sizes of the single-universe symmetric form for n bodies in blocks of ib = 64*K*W.  The
// workspace holds the block-diagonal sums (3 n values), the FP64 sums between slabs (3 n
// doubles, only when there is more than one slab) and one slab's partials: per segment 3 * ib
// values of p_i and seglen * 3 * ib of p_j.
struct SymPlan {
  int ib, nb, lg;
  int64_t elem, nseg, diag_bytes, acc64_bytes, seg_bytes, full_bytes, min_bytes;
};

SymPlan sym_plan(int64_t n, int K, int W, int64_t elem) {
  SymPlan P{};
  P.ib = kWave * K * W;
  P.nb = static_cast<int>((n + P.ib - 1) / P.ib);
  P.lg = seg_lg(P.nb);
  P.elem = elem;
  P.nseg = P.nb > 1 ? seg_base(P.nb - 1, P.nb, P.lg) : 0;
  const int64_t per = int64_t(3) * P.ib * elem;
  P.diag_bytes = up256(int64_t(3) * n * elem);
  P.acc64_bytes = up256(int64_t(3) * n * 8);
  P.seg_bytes = per * (1 + (int64_t(1) << P.lg));
  P.full_bytes = P.diag_bytes + P.nseg * P.seg_bytes;
  const int64_t row0 = P.nb > 1 ? seg_count(0, P.nb, P.lg) : 0;
  P.min_bytes = std::min(P.full_bytes, P.diag_bytes + P.acc64_bytes + row0 * P.seg_bytes);
  return P;
}

// Slabs of consecutive rows of [row_lo, row_hi) whose segments fit `maxseg`: rows
// [cut[k], cut[k+1]).
std::vector<int> sym_slabs(const SymPlan &P, int64_t maxseg, int row_lo, int row_hi) {
  std::vector<int> cut{row_lo};
  int64_t have = 0;
  for (int beta = row_lo; beta < row_hi; ++beta) {
    const int c = seg_count(beta, P.nb, P.lg);
    if (have + c > maxseg) {
      cut.push_back(beta);
      have = 0;
    }
    have += c;
  }
  cut.push_back(std::max(row_lo, row_hi));
  return cut;
}

// The part of the symmetric form one launch sequence computes: rows [row_lo, row_hi) and
// [row_lo2, row_hi2) of block pairs and the block-diagonal sums of blocks [d_lo, d_hi).  The
// whole universe is rows [0, nb - 1) and blocks [0, nb).
struct SymRange {
  int row_lo, row_hi, row_lo2, row_hi2, d_lo, d_hi;
};

// Segments of a range's rows, and of its longest row (the smallest-index one).
inline void sym_range_segments(const SymPlan &P, const SymRange &R, int64_t &nseg,
                               int64_t &row0) {
  nseg = 0;
  row0 = 0;
  const int lo[2] = {R.row_lo, R.row_lo2}, hi[2] = {R.row_hi, R.row_hi2};
  for (int r = 0; r < 2; ++r) {
    if (hi[r] <= lo[r]) continue;
    nseg += seg_base(hi[r], P.nb, P.lg) - seg_base(lo[r], P.nb, P.lg);
    row0 = std::max<int64_t>(row0, seg_count(lo[r], P.nb, P.lg));
  }
}

// The symmetric form's launches: the block-diagonal pass `diag_fn(diag)` over the range's
// blocks, then per slab the segments `seg_fn(span, nseg, p_i, p_j)` and the fold.
// sums == NULL: acc = G * (every sum), as T; else the range's unscaled FP64 sums go to sums
// (3 n doubles) and acc is not written.  Workspace: the range's diagonal sums, (without sums)
// the FP64 sums between slabs when there is more than one, and one slab's partials.
template <typename T, typename DiagFn, typename SegFn>
int sym_run(const SymPlan &P, const SymRange &R, int n, T *acc, double *sums, double G,
            void *workspace, int64_t workspace_bytes, hipStream_t s, DiagFn diag_fn,
            SegFn seg_fn, const char *fn) {
  const int64_t per = int64_t(3) * P.ib * P.elem;
  const int dstride = std::min(n, R.d_hi * P.ib) - R.d_lo * P.ib;
  const int64_t diag_bytes = up256(int64_t(3) * std::max(dstride, 0) * P.elem);
  int64_t nseg, row0;
  sym_range_segments(P, R, nseg, row0);
  char *ws = static_cast<char *>(workspace);
  T *diag = reinterpret_cast<T *>(ws);
  double *acc64 = sums;
  char *parts = ws + diag_bytes;
  int64_t maxseg = nseg;
  if (workspace_bytes < diag_bytes + nseg * P.seg_bytes) {
    int64_t room = workspace_bytes - diag_bytes;
    if (sums == nullptr) {
      acc64 = reinterpret_cast<double *>(parts);
      parts += P.acc64_bytes;
      room -= P.acc64_bytes;
    }
    // largest segment count whose partials fit (3 * ib * elem, hence every region, is a
    // multiple of 256 bytes: ib is a multiple of 256)
    maxseg = room / P.seg_bytes;
    if (maxseg < row0)
      return fail(GRAV_EINVAL, "%s: workspace of %lld bytes is below the minimum", fn,
                  (long long)workspace_bytes);
  }
  T *p_i = reinterpret_cast<T *>(parts);
  T *p_j = reinterpret_cast<T *>(parts + maxseg * per);
  int rc = dstride > 0 ? diag_fn(diag) : GRAV_OK;
  if (rc != GRAV_OK) return rc;
  // Slabs {rows [a, b), rows [a2, b2)}: both ranges at once when every segment fits, else the
  // slabs of the first range and then of the second, one range each.
  struct Slab {
    int a, b, a2, b2;
  };
  std::vector<Slab> slab;
  if (maxseg >= nseg) {
    slab.push_back({R.row_lo, R.row_hi, R.row_lo2, R.row_hi2});
  } else {
    for (int r = 0; r < 2; ++r) {
      const int lo = r == 0 ? R.row_lo : R.row_lo2, hi = r == 0 ? R.row_hi : R.row_hi2;
      if (r == 1 && hi <= lo) break;
      const std::vector<int> cut = sym_slabs(P, maxseg, lo, hi);
      for (size_t k = 0; k + 1 < cut.size(); ++k) slab.push_back({cut[k], cut[k + 1], 0, 0});
    }
  }
  const int slabs = static_cast<int>(slab.size());
  for (int k = 0; k < slabs; ++k) {
    const Slab &L = slab[k];
    SegSpan sp{};
    sp.first = seg_base(L.a, P.nb, P.lg);
    sp.first2 = seg_base(std::max(L.a2, 0), P.nb, P.lg);
    sp.beta_lo = L.a;
    sp.beta_lo2 = L.a2;
    const int64_t n1 = seg_base(L.b, P.nb, P.lg) - sp.first;
    const int64_t n2 = L.b2 > L.a2 ? seg_base(L.b2, P.nb, P.lg) - sp.first2 : 0;
    if (n1 + n2 > 0x7fffffff) return fail(GRAV_EINVAL, "%s: too many segments in one slab", fn);
    sp.n1 = static_cast<int>(n1);
    if (n1 + n2 > 0) {
      rc = seg_fn(sp, static_cast<int>(n1 + n2), p_i, p_j);
      if (rc != GRAV_OK) return rc;
    }
    const int last = (sums == nullptr && k == slabs - 1) ? 1 : 0;
    hipLaunchKernelGGL((k_sym_fold<T>), dim3((n + 255) / 256, 3), dim3(256), 0, s, diag, acc64,
                       acc, p_i, p_j, n, P.nb, P.ib, P.lg, L.a, L.b, L.a2, L.b2, sp.n1,
                       k == 0 ? 1 : 0, last, G, R.d_lo, R.d_hi, dstride);
    rc = check_launch("k_sym_fold");
    if (rc != GRAV_OK) return rc;
  }
  return GRAV_OK;
}

// Share `share` of `shares` when the symmetric form is divided among GPUs by rows of block
// pairs.  Row beta holds nb - 1 - beta block pairs, so rows beta and nb - 2 - beta together
// hold nb: these row couples (the middle row alone when the row count is odd) are the units,
// and share k takes the units [u0, u1) cut as np4 cuts its row batches (the first unit with
// at least k/shares of the pairs before it, src/gravitation/kernel/np4.py:101-118), that is
// rows [u0, u1) and their partners.  Equal units make equal shares whenever the unit count
// divides; contiguous rows would differ by up to one row.  The block-diagonal sums go by
// blocks, [k nb / shares, (k + 1) nb / shares).
SymRange sym_share_range(const SymPlan &P, int share, int shares) {
  const int rows = std::max(0, P.nb - 1);
  const int units = (rows + 1) / 2;
  const int64_t total = int64_t(P.nb) * (P.nb - 1) / 2;
  auto start = [&](int k) {
    if (k >= shares) return units;
    // units before u hold u * nb pairs (only the last unit can be a single row)
    const int64_t u = (int64_t(k) * total + int64_t(shares) * P.nb - 1) /
                      (int64_t(shares) * P.nb);
    return static_cast<int>(std::min<int64_t>(u, units));
  };
  const int u0 = start(share), u1 = start(share + 1);
  SymRange R;
  R.row_lo = u0;
  R.row_hi = u1;
  R.row_lo2 = std::max(u1, rows - u1);
  R.row_hi2 = rows - u0;
  if (R.row_hi2 < R.row_lo2) R.row_hi2 = R.row_lo2;
  R.d_lo = static_cast<int>(int64_t(share) * P.nb / shares);
  R.d_hi = static_cast<int>(int64_t(share + 1) * P.nb / shares);
  return R;
}

// Workspace of a share: its diagonal sums and its partials, all at once (full) or one row of
// segments at a time (min).
void sym_share_bytes(const SymPlan &P, const SymRange &R, int64_t n, int64_t &full,
                     int64_t &min) {
  const int64_t dstride = std::min<int64_t>(n, int64_t(R.d_hi) * P.ib) - int64_t(R.d_lo) * P.ib;
  const int64_t diag = up256(3 * std::max<int64_t>(dstride, 0) * P.elem);
  int64_t nseg, row0;
  sym_range_segments(P, R, nseg, row0);
  full = std::max<int64_t>(256, diag + nseg * P.seg_bytes);
  min = std::max<int64_t>(256, diag + row0 * P.seg_bytes);
}

// acc[c][i] = G * sum over the shares k, in order, of sums[k][c][i] (FP64), as FP32.
__global__ __launch_bounds__(256) void k_sym_combine(const double *__restrict__ sums,
                                                     int shares, int count, double G,
                                                     float *__restrict__ acc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= count) return;
  double s = 0.0;
  for (int k = 0; k < shares; ++k)
    s += sums[(static_cast<size_t>(k) * 3 + c) * count + i];
  acc[static_cast<size_t>(c) * count + i] = static_cast<float>(G * s);
}

bool sym_supported(int K, int W) {
  return (K == 4 || K == 8) && (W == 2 || W == 4);
}

// (0, 0) selects the shape by n, from MI355X measurements on Plummer spheres: blocks of 512
// (K 4, W 2) below 57344 bodies, where blocks of 1024 leave too few block pairs to fill 256
// CUs (at 49152 bodies (4, 2) takes 0.60 ms against 0.77), and blocks of 1024 (K 8, W 2)
// above (1.27x over owner-computes at 2^18 in its own default shape, (1, 16)).
void sym_shape(int64_t n, int &K, int &W) {
  if (K == 0 && W == 0) {
    K = n < 57344 ? 4 : 8;
    W = 2;
  }
}

template <int K, int W>
int launch_segments(const float4 *posm, int n, const SymPlan &P, const SegSpan &sp, int nseg,
                    float eps2, float *pi, float *pj, hipStream_t s) {
  hipLaunchKernelGGL((k_sym_segments<K, W>), dim3(static_cast<unsigned>(nseg)),
                     dim3(kWave * W), 0, s, posm, n, P.nb, P.lg, sp, eps2, pi, pj);
  return check_launch("k_sym_segments");
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
int launch_segments_f64(const double4 *posm, int n, const SymPlan &P, const SegSpan &sp,
                        int nseg, double eps2, double *pi, double *pj, hipStream_t s) {
  hipLaunchKernelGGL((k_sym_segments_f64<K, W>), dim3(static_cast<unsigned>(nseg)),
                     dim3(kWave * W), 0, s, posm, n, P.nb, P.lg, sp, eps2, pi, pj);
  return check_launch("k_sym_segments_f64");
}

template <int K, int W>
int launch_pairs_items(const float4 *posm, float eps2, float *pi, float *pj, const int4 *items,
                       int nitems, hipStream_t s) {
  if (nitems == 0) return GRAV_OK;
  hipLaunchKernelGGL((k_sym_pairs<K, W>), dim3(static_cast<unsigned>(nitems)), dim3(kWave * W),
                     0, s, posm, eps2, pi, pj, items);
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

namespace {
int sym_workspace_query(const char *fn, int64_t n, int K, int W, bool f64, bool minimum,
                        int64_t *bytes) {
  clear_error();
  GRAV_REQUIRE(bytes != nullptr, "%s: bytes is NULL", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  if (f64) {
    sym_shape_f64(n, K, W);
    GRAV_REQUIRE(sym_supported_f64(K, W),
                 "%s: (i_chunks_per_wave, waves) must be (0, 0) or in {2,4} x {2,4}", fn);
  } else {
    sym_shape(n, K, W);
    GRAV_REQUIRE(sym_supported(K, W),
                 "%s: (i_chunks_per_wave, waves) must be (0, 0) or in {4,8} x {2,4}", fn);
  }
  const SymPlan P = sym_plan(n, K, W, f64 ? 8 : 4);
  *bytes = minimum ? P.min_bytes : P.full_bytes;
  return GRAV_OK;
}

int sym_validate(const char *fn, const void *posm, const void *acc, int64_t n, double G,
                 double eps, uintptr_t align_p, uintptr_t align_a, const void *workspace) {
  GRAV_REQUIRE(posm != nullptr && acc != nullptr, "%s: NULL buffer", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(aligned(posm, align_p) && aligned(acc, align_a), "%s: misaligned", fn);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "%s: eps must be >= 0", fn);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256),
               "%s: workspace must be a 256-byte aligned device buffer", fn);
  return GRAV_OK;
}
}  // namespace

int grav_stage1_sym_workspace(int64_t n, int i_chunks_per_wave, int waves, int64_t *bytes) {
  return sym_workspace_query("grav_stage1_sym_workspace", n, i_chunks_per_wave, waves, false,
                             false, bytes);
}

int grav_stage1_sym_workspace_min(int64_t n, int i_chunks_per_wave, int waves,
                                  int64_t *bytes) {
  return sym_workspace_query("grav_stage1_sym_workspace_min", n, i_chunks_per_wave, waves,
                             false, true, bytes);
}

int grav_stage1_sym_f32(const void *posm, void *acc, int64_t n, double G, double eps,
                        int i_chunks_per_wave, int waves, void *workspace,
                        int64_t workspace_bytes, void *stream) {
  clear_error();
  const char *fn = "grav_stage1_sym_f32";
  int rc = sym_validate(fn, posm, acc, n, G, eps, 16, 4, workspace);
  if (rc != GRAV_OK) return rc;
  int K = i_chunks_per_wave, W = waves;
  sym_shape(n, K, W);
  GRAV_REQUIRE(sym_supported(K, W),
               "%s: (i_chunks_per_wave, waves) must be (0, 0) or in {4,8} x {2,4}", fn);
  const SymPlan P = sym_plan(n, K, W, 4);
  GRAV_REQUIRE(workspace_bytes >= P.min_bytes,
               "%s: the workspace must hold at least %lld bytes (%lld for one slab)", fn,
               (long long)P.min_bytes, (long long)P.full_bytes);
  const int nn = static_cast<int>(n);
  const float eps2 = static_cast<float>(eps * eps);
  const int mask = eps == 0.0 ? 1 : 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float4 *p = static_cast<const float4 *>(posm);
  // pairs inside each block: owner-computes over the blocks of IB bodies (a multiple of 256)
  auto diag_fn = [&](float *diag) {
    return stage1_f32_launch(p, diag, nn, 1.0f, eps2, mask, 4, 4, 256, nullptr, P.ib, s);
  };
  auto seg_fn = [&](const SegSpan &sp, int nseg, float *pi, float *pj) {
    if (K == 4 && W == 2) return launch_segments<4, 2>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    if (K == 4 && W == 4) return launch_segments<4, 4>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    if (K == 8 && W == 2) return launch_segments<8, 2>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    return launch_segments<8, 4>(p, nn, P, sp, nseg, eps2, pi, pj, s);
  };
  const SymRange R{0, P.nb - 1, 0, 0, 0, P.nb};
  return sym_run<float>(P, R, nn, static_cast<float *>(acc), nullptr, G, workspace,
                        workspace_bytes, s, diag_fn, seg_fn, fn);
}

int grav_stage1_sym_share_plan(int64_t n, int share, int shares, int i_chunks_per_wave,
                               int waves, int *rows, int64_t *bytes) {
  clear_error();
  const char *fn = "grav_stage1_sym_share_plan";
  GRAV_REQUIRE(rows != nullptr && bytes != nullptr, "%s: NULL output", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(shares >= 1 && shares <= 65536 && share >= 0 && share < shares,
               "%s: need 0 <= share < shares <= 65536 (got %d of %d)", fn, share, shares);
  int K = i_chunks_per_wave, W = waves;
  sym_shape(n, K, W);
  GRAV_REQUIRE(sym_supported(K, W),
               "%s: (i_chunks_per_wave, waves) must be (0, 0) or in {4,8} x {2,4}", fn);
  const SymPlan P = sym_plan(n, K, W, 4);
  const SymRange R = sym_share_range(P, share, shares);
  rows[0] = R.row_lo;
  rows[1] = R.row_hi;
  rows[2] = R.row_lo2;
  rows[3] = R.row_hi2;
  rows[4] = R.d_lo;
  rows[5] = R.d_hi;
  sym_share_bytes(P, R, n, bytes[0], bytes[1]);
  return GRAV_OK;
}

int grav_stage1_sym_f32_share(const void *posm, void *sums64, int64_t n, double eps, int share,
                              int shares, int i_chunks_per_wave, int waves, void *workspace,
                              int64_t workspace_bytes, void *stream) {
  clear_error();
  const char *fn = "grav_stage1_sym_f32_share";
  int rc = sym_validate(fn, posm, sums64, n, 1.0, eps, 16, 8, workspace);
  if (rc != GRAV_OK) return rc;
  GRAV_REQUIRE(shares >= 1 && shares <= 65536 && share >= 0 && share < shares,
               "%s: need 0 <= share < shares <= 65536 (got %d of %d)", fn, share, shares);
  int K = i_chunks_per_wave, W = waves;
  sym_shape(n, K, W);
  GRAV_REQUIRE(sym_supported(K, W),
               "%s: (i_chunks_per_wave, waves) must be (0, 0) or in {4,8} x {2,4}", fn);
  const SymPlan P = sym_plan(n, K, W, 4);
  const SymRange R = sym_share_range(P, share, shares);
  int64_t full = 0, min = 0;
  sym_share_bytes(P, R, n, full, min);
  GRAV_REQUIRE(workspace_bytes >= min,
               "%s: the workspace must hold at least %lld bytes (%lld for one slab)", fn,
               (long long)min, (long long)full);
  const int nn = static_cast<int>(n);
  const float eps2 = static_cast<float>(eps * eps);
  const int mask = eps == 0.0 ? 1 : 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float4 *p = static_cast<const float4 *>(posm);
  const int i0 = R.d_lo * P.ib;
  const int m = std::min(nn, R.d_hi * P.ib) - i0;
  auto diag_fn = [&](float *diag) {
    return stage1_f32_diag_launch(p, diag, nn, i0, m, eps2, mask, P.ib, s);
  };
  auto seg_fn = [&](const SegSpan &sp, int nseg, float *pi, float *pj) {
    if (K == 4 && W == 2) return launch_segments<4, 2>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    if (K == 4 && W == 4) return launch_segments<4, 4>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    if (K == 8 && W == 2) return launch_segments<8, 2>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    return launch_segments<8, 4>(p, nn, P, sp, nseg, eps2, pi, pj, s);
  };
  return sym_run<float>(P, R, nn, nullptr, static_cast<double *>(sums64), 1.0, workspace,
                        workspace_bytes, s, diag_fn, seg_fn, fn);
}

int grav_stage1_sym_shares_combine(const void *sums64, int shares, int64_t count, double G,
                                   void *acc, void *stream) {
  clear_error();
  const char *fn = "grav_stage1_sym_shares_combine";
  GRAV_REQUIRE(sums64 != nullptr && acc != nullptr, "%s: NULL buffer", fn);
  GRAV_REQUIRE(aligned(sums64, 8) && aligned(acc, 4), "%s: misaligned", fn);
  GRAV_REQUIRE(shares >= 1 && shares <= 65536, "%s: shares must be in [1, 65536] (got %d)", fn,
               shares);
  GRAV_REQUIRE(count >= 1 && count <= kMaxBodies, "%s: count must be in [1, 2^29] (got %lld)",
               fn, (long long)count);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  const int c = static_cast<int>(count);
  hipLaunchKernelGGL(k_sym_combine, dim3((c + 255) / 256, 3), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const double *>(sums64),
                     shares, c, G, static_cast<float *>(acc));
  return check_launch("k_sym_combine");
}

int grav_stage1_sym_f64_workspace(int64_t n, int i_chunks_per_wave, int waves,
                                  int64_t *bytes) {
  return sym_workspace_query("grav_stage1_sym_f64_workspace", n, i_chunks_per_wave, waves,
                             true, false, bytes);
}

int grav_stage1_sym_f64_workspace_min(int64_t n, int i_chunks_per_wave, int waves,
                                      int64_t *bytes) {
  return sym_workspace_query("grav_stage1_sym_f64_workspace_min", n, i_chunks_per_wave, waves,
                             true, true, bytes);
}

int grav_stage1_sym_f64(const void *posm64, void *acc64, int64_t n, double G, double eps,
                        int i_chunks_per_wave, int waves, void *workspace,
                        int64_t workspace_bytes, void *stream) {
  clear_error();
  const char *fn = "grav_stage1_sym_f64";
  int rc = sym_validate(fn, posm64, acc64, n, G, eps, 32, 8, workspace);
  if (rc != GRAV_OK) return rc;
  int K = i_chunks_per_wave, W = waves;
  sym_shape_f64(n, K, W);
  GRAV_REQUIRE(sym_supported_f64(K, W),
               "%s: (i_chunks_per_wave, waves) must be (0, 0) or in {2,4} x {2,4}", fn);
  const SymPlan P = sym_plan(n, K, W, 8);
  GRAV_REQUIRE(workspace_bytes >= P.min_bytes,
               "%s: the workspace must hold at least %lld bytes (%lld for one slab)", fn,
               (long long)P.min_bytes, (long long)P.full_bytes);
  const int nn = static_cast<int>(n);
  const double eps2 = eps * eps;
  const int mask = eps == 0.0 ? 1 : 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double4 *p = static_cast<const double4 *>(posm64);
  auto diag_fn = [&](double *diag) {
    return stage1_f64_launch(p, diag, nn, 1.0, eps2, mask, P.ib, s);
  };
  auto seg_fn = [&](const SegSpan &sp, int nseg, double *pi, double *pj) {
    if (K == 2 && W == 2) return launch_segments_f64<2, 2>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    if (K == 2 && W == 4) return launch_segments_f64<2, 4>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    if (K == 4 && W == 2) return launch_segments_f64<4, 2>(p, nn, P, sp, nseg, eps2, pi, pj, s);
    return launch_segments_f64<4, 4>(p, nn, P, sp, nseg, eps2, pi, pj, s);
  };
  const SymRange R{0, P.nb - 1, 0, 0, 0, P.nb};
  return sym_run<double>(P, R, nn, static_cast<double *>(acc64), nullptr, G, workspace,
                         workspace_bytes, s, diag_fn, seg_fn, fn);
}

}  // extern "C"
