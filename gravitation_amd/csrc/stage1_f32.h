// FP32 stage 1 of one workgroup (owner-computes, blocked TwoSum accumulation) and the
// j-split form's reduce + FP64-master stage 2 of one body, shared by the stage-1 kernels
// (stage1.hip) and the paired backward kernels (adjoint.hip).  The kernel layout is
// described at the top of stage1.hip.
#pragma once

#include "grav_common.h"

namespace grav {
namespace {

// Knuth's TwoSum: hi + p == s + e exactly; hi <- s, lo <- lo + e.
__device__ __forceinline__ void two_sum_fold(float &hi, float &lo, float p) {
  const float s = hi + p;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (p - bp);
  hi = s;
  lo += e;
}

template <int K, bool MASK>
__device__ __forceinline__ void pair_f32(const float4 q, const int j, const float (&xi)[K],
                                         const float (&yi)[K], const float (&zi)[K],
                                         const int (&ii)[K], const float eps2, float (&px)[K],
                                         float (&py)[K], float (&pz)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float dx = q.x - xi[k];
    const float dy = q.y - yi[k];
    const float dz = q.z - zi[k];
    float r2 = __builtin_fmaf(dx, dx, eps2);
    r2 = __builtin_fmaf(dy, dy, r2);
    r2 = __builtin_fmaf(dz, dz, r2);
    float inv = __builtin_amdgcn_rsqf(r2);
    if (MASK) inv = (j == ii[k]) ? 0.0f : inv;
    const float s = (q.w * inv) * (inv * inv);
    px[k] = __builtin_fmaf(dx, s, px[k]);
    py[k] = __builtin_fmaf(dy, s, py[k]);
    pz[k] = __builtin_fmaf(dz, s, pz[k]);
  }
}

// s_waitcnt immediate that waits for every scalar (and LDS) access and no other counter:
// vmcnt 63 (bits 3:0 and 15:14), expcnt 7 (6:4), lgkmcnt 0 (11:8).
constexpr unsigned kWaitScalar = 0xC07F;

// Accumulates j in [j0, j1) into the block partials px/py/pz, JB j-bodies per loop step (one
// s_load of JB float4).  The order of the pairs, hence the result, is the same for any JB.
// Two per step is the faster loop in launches of many workgroups, eight in smaller ones
// (pair_loop_width).  With two and one body per lane, the loads run one step ahead in two
// SGPR sets a and b: each half-step waits for the set it is about to use, issues the load of
// the other set's next step and only then computes, so the scalar load's latency hides
// behind a step of pairs (scalar loads can return out of order, so a wait is always for all
// of them; the explicit wait comes before the next load is issued).  MI355X, Plummer:
// 16.57-16.64 against 17.12-17.17 ms at 2^18 with (1, 16), the j-split form at 16384 bodies
// 0.079 against 0.081 ms, the batched per-step schedule on 4096 universes of 32-2048 bodies
// 0.548 against 0.563 ms per step.  It made
// the j-split form with two bodies per lane slower (2^18: 18.7 against 17.4 ms), and with
// eight j-bodies per step the second set pushed SGPRs into VGPRs (K = 2: 51 to 118).
template <int K, bool MASK, int JB>
__device__ __forceinline__ void pair_range_f32(const float4 *__restrict__ posm, int j0, int j1,
                                               const float (&xi)[K], const float (&yi)[K],
                                               const float (&zi)[K], const int (&ii)[K],
                                               const float eps2, float (&px)[K],
                                               float (&py)[K], float (&pz)[K]) {
  int j = j0;
  if constexpr (JB <= 2 && K == 1) {
    if (j + JB <= j1) {
      float4 a[JB], b[JB];
#pragma unroll
      for (int u = 0; u < JB; ++u) a[u] = posm[j + u];
      for (; j + 2 * JB <= j1; j += 2 * JB) {
        __builtin_amdgcn_s_waitcnt(kWaitScalar);
#pragma unroll
        for (int u = 0; u < JB; ++u) b[u] = posm[j + JB + u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < JB; ++u)
          pair_f32<K, MASK>(a[u], j + u, xi, yi, zi, ii, eps2, px, py, pz);
        // the next step of set a, or (past the end) a step in range that is not used
        const int jn = min(j + 2 * JB, j1 - JB);
        __builtin_amdgcn_s_waitcnt(kWaitScalar);
#pragma unroll
        for (int u = 0; u < JB; ++u) a[u] = posm[jn + u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < JB; ++u)
          pair_f32<K, MASK>(b[u], j + JB + u, xi, yi, zi, ii, eps2, px, py, pz);
      }
      if (j + JB <= j1) {
#pragma unroll
        for (int u = 0; u < JB; ++u)
          pair_f32<K, MASK>(a[u], j + u, xi, yi, zi, ii, eps2, px, py, pz);
        j += JB;
      }
    }
  } else {
    for (; j + JB <= j1; j += JB) {
      float4 q[JB];
#pragma unroll
      for (int u = 0; u < JB; ++u) q[u] = posm[j + u];
#pragma unroll
      for (int u = 0; u < JB; ++u)
        pair_f32<K, MASK>(q[u], j + u, xi, yi, zi, ii, eps2, px, py, pz);
    }
  }
  for (; j < j1; ++j) {
    const float4 q = posm[j];
    pair_f32<K, MASK>(q, j, xi, yi, zi, ii, eps2, px, py, pz);
  }
}

// Blocked compensated sum over j in [j0, j1).  MASK: the pairs with j in [ma, mb) (the
// workgroup's own i-range) run the self-masked pair routine.  That choice does not change a
// pair's value for j != i, and the blocks start at j0 whatever [ma, mb) is, so a body's sum
// does not depend on which i-tile holds it.
template <int K, bool MASK, int JB>
__device__ __forceinline__ void sum_segment_f32(const float4 *__restrict__ posm, int j0, int j1,
                                                int jblock, int ma, int mb,
                                                const float (&xi)[K],
                                                const float (&yi)[K], const float (&zi)[K],
                                                const int (&ii)[K], const float eps2,
                                                float (&hx)[K], float (&hy)[K], float (&hz)[K],
                                                float (&lx)[K], float (&ly)[K], float (&lz)[K]) {
  for (int jb = j0, je; jb < j1; jb = je) {
    je = (j1 - jb > jblock) ? jb + jblock : j1;  // no overflow for jblock near INT_MAX
    float px[K], py[K], pz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) px[k] = py[k] = pz[k] = 0.0f;
    if constexpr (MASK) {
      const int a = min(je, max(jb, ma));
      const int b = max(a, min(je, mb));
      pair_range_f32<K, false, JB>(posm, jb, a, xi, yi, zi, ii, eps2, px, py, pz);
      pair_range_f32<K, true, JB>(posm, a, b, xi, yi, zi, ii, eps2, px, py, pz);
      pair_range_f32<K, false, JB>(posm, b, je, xi, yi, zi, ii, eps2, px, py, pz);
    } else {
      pair_range_f32<K, false, JB>(posm, jb, je, xi, yi, zi, ii, eps2, px, py, pz);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      two_sum_fold(hx[k], lx[k], px[k]);
      two_sum_fold(hy[k], ly[k], py[k]);
      two_sum_fold(hz[k], lz[k], pz[k]);
    }
  }
}

// The workgroup at (bx, by) of an (x, ny) grid owns i in [items[bx].x, items[bx].y) (at most
// 64*K bodies) against j in [items[bx].z, items[bx].w); items == NULL means one universe:
// i-tile bx of the shard [i0, i0 + m) against all n bodies, or, with diag_block > 0, against
// its own block of diag_block bodies only, or, with part != NULL, against the by-th of ny
// equal j-ranges of [0, n) (the j-split form: the workgroup writes its unscaled compensated
// partial to part[by][c][i - i0] in FP64, and a reduce kernel sums the ranges).  Outputs are
// rows i - i0 of (3, m) arrays; a whole universe is i0 = 0, m = n, and with items i0 = 0 and
// m is the row stride.  The item list lets one launch cover many universes (one item per
// universe tile).  JB: j-bodies per step of the pair loop (pair_range_f32), 2 or 8; the
// result does not depend on it.  red: (S - 1) * 6 * (64 * K) floats of LDS.
template <int K, int S, int JB>
__device__ __forceinline__ void stage1_f32_block(const float4 *__restrict__ posm,
                                                 float *__restrict__ acc, int n, int i0, int m,
                                                 float G, float eps2, int mask_self,
                                                 int jblock,
                                                 const int4 *__restrict__ items,
                                                 int diag_block, double *__restrict__ part,
                                                 int bx, int by, int ny, float *red) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int ibase, iend, jlo, jhi;
  if (items != nullptr) {
    const int4 it = items[bx];
    ibase = it.x;
    iend = it.y;
    jlo = it.z;
    jhi = it.w;
  } else {
    ibase = i0 + bx * (kWave * K);
    iend = min(i0 + m, ibase + kWave * K);
    // diag_block > 0: only the i-tile's own block of diag_block bodies (a multiple of 64*K)
    jlo = diag_block > 0 ? (ibase / diag_block) * diag_block : 0;
    jhi = diag_block > 0 ? min(n, jlo + diag_block) : n;
    if (part != nullptr) {
      const int ns = ny;
      const int span = ((n + ns - 1) / ns + 7) & ~7;
      jlo = min(n, by * span);
      jhi = min(n, jlo + span);
    }
  }

  float xi[K], yi[K], zi[K];
  int ii[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = ibase + k * kWave + lane;
    ii[k] = i;
    const float4 p = posm[min(i, iend - 1)];
    xi[k] = p.x;
    yi[k] = p.y;
    zi[k] = p.z;
  }

  float hx[K], hy[K], hz[K], lx[K], ly[K], lz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.0f;

  // This wave's j-range; with mask_self the pairs with j in the workgroup's own i-range
  // [ibase, iend) are self-masked.
  const int chunk = ((jhi - jlo + S - 1) / S + 7) & ~7;
  const int j0 = min(jhi, jlo + wave * chunk);
  const int j1 = min(jhi, j0 + chunk);
  // eps2 in a VGPR, so that the first r2 FMA takes no SGPR operand (an SGPR operand adds
  // issue cycles to an FP32 instruction): 2^18 with (1, 16) 16.86 against 17.14 ms
  float eps2v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(eps2v) : "s"(eps2));
  if (mask_self)
    sum_segment_f32<K, true, JB>(posm, j0, j1, jblock, ibase, iend, xi, yi, zi, ii, eps2v, hx,
                                 hy, hz, lx, ly, lz);
  else
    sum_segment_f32<K, false, JB>(posm, j0, j1, jblock, 0, 0, xi, yi, zi, ii, eps2v, hx, hy,
                                  hz, lx, ly, lz);

  if constexpr (S > 1) {
    constexpr int L = kWave * K;
    if (wave > 0) {
      float *r = red + (wave - 1) * 6 * L;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int t = k * kWave + lane;
        r[0 * L + t] = hx[k];
        r[1 * L + t] = hy[k];
        r[2 * L + t] = hz[k];
        r[3 * L + t] = lx[k];
        r[4 * L + t] = ly[k];
        r[5 * L + t] = lz[k];
      }
    }
    __syncthreads();
    if (wave != 0) return;
    for (int w = 1; w < S; ++w) {
      const float *r = red + (w - 1) * 6 * L;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int t = k * kWave + lane;
        two_sum_fold(hx[k], lx[k], r[0 * L + t]);
        two_sum_fold(hy[k], ly[k], r[1 * L + t]);
        two_sum_fold(hz[k], lz[k], r[2 * L + t]);
        lx[k] += r[3 * L + t];
        ly[k] += r[4 * L + t];
        lz[k] += r[5 * L + t];
      }
    }
  }

  if (part != nullptr) {
    // hi + lo of two floats is exact in FP64
    double *o = part + static_cast<size_t>(by) * 3 * m;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = ii[k];
      if (i < iend) {
        o[i - i0] = static_cast<double>(hx[k]) + static_cast<double>(lx[k]);
        o[m + i - i0] = static_cast<double>(hy[k]) + static_cast<double>(ly[k]);
        o[2 * m + i - i0] = static_cast<double>(hz[k]) + static_cast<double>(lz[k]);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = ii[k];
    if (i < iend) {
      acc[i - i0] = G * (hx[k] + lx[k]);
      acc[m + i - i0] = G * (hy[k] + ly[k]);
      acc[2 * m + i - i0] = G * (hz[k] + lz[k]);
    }
  }
}

// The pair loop's j-bodies per step for a launch of `workgroups` workgroups of K bodies per
// lane.  MI355X, ms: 2^18 bodies, 4096 workgroups of (1, 16): 2 per step 17.90 against 18.43
// for 8; 4096 universes of 32-2048 bodies in the batched per-step schedule (thousands of
// workgroups) 0.56 against 0.62 per step; but one launch per universe (at most 32 workgroups)
// 44.0 against 33.9 per step of all 4096, and a 4096-body universe (512 workgroups) forward +
// adjoint 7.92 against 7.69.  From K = 4 eight per step always: the K registers of the
// i-bodies leave no room for the shorter loop's gain.  With one body per lane the loop of two
// loads its next step ahead, and then two per step wins from 256 workgroups on: the j-split
// form at 4096 bodies (512 workgroups) 0.0125 against 0.0133 ms, that universe's forward +
// adjoint 7.26-7.29 against 7.58-7.60 ms; one launch per universe of 32-2048 bodies (at
// most 32 workgroups) still 38.3 against 34.0 ms per step of all 4096.
inline int pair_loop_width(int64_t workgroups, int K) {
  if (K == 1) return workgroups < 256 ? 8 : 2;
  return (K >= 4 || workgroups < 1024) ? 8 : 2;
}

// sum over the j-ranges p of part[p][c][i], in FP64 and in range order, eight loads in
// flight.
__device__ __forceinline__ double split_sum(const double *__restrict__ part, int c, int i, int n,
                                            int splits) {
  const size_t stride = static_cast<size_t>(3) * n;
  const double *q = part + static_cast<size_t>(c) * n + i;
  double s = 0.0;
  int p = 0;
  for (; p + 8 <= splits; p += 8) {
    double v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = q[(p + t) * stride];
#pragma unroll
    for (int t = 0; t < 8; ++t) s += v[t];
  }
  for (; p < splits; ++p) s += q[p * stride];
  return s;
}

// One body's reduce of the FP32 j-split form fused with the FP64-master stage 2: the
// acceleration is rounded to FP32 exactly as k_split_reduce<float> stores it, then r, v
// advance as in k_stage2_f64m (v = a dt + v, r = v dt + r, no contraction), so a step equals
// the two kernels' bit for bit without writing acc.  The new FP32 position (x, y, z, m) of
// body i goes to out[i]; out may be in (body i's mass is read before its entry is written).
__device__ __forceinline__ void split_step_body(const double *__restrict__ part,
                                                const float4 *in, float4 *out,
                                                double *__restrict__ r64,
                                                double *__restrict__ v64, int n, int splits,
                                                double G, double dt, int i) {
#pragma clang fp contract(off)
  double r[3], v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float a32 = static_cast<float>(G * split_sum(part, c, i, n, splits));
    const double a = static_cast<double>(a32);
    v[c] = a * dt + v64[c * n + i];
    r[c] = v[c] * dt + r64[c * n + i];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v64[c * n + i] = v[c];
    r64[c * n + i] = r[c];
  }
  const float m = in[i].w;
  out[i] = make_float4(static_cast<float>(r[0]), static_cast<float>(r[1]),
                       static_cast<float>(r[2]), m);
}

}  // namespace
}  // namespace grav
