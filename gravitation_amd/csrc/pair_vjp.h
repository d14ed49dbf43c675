// Vector-Jacobian product of the pair acceleration, shared by the single-universe and the
// batched adjoint kernels.
//
// Forward (per owner k): a_k = G sum_j m_j s_kj d_kj with d_kj = r_j - r_k,
// q = |d|^2 + eps^2, s = q^(-3/2).  Its adjoint against a cotangent field mu is
//   (J_r^T mu)_k = G sum_{j != k} T_kj w_kj,   T_kj = s (I - 3 d d^T / q),
//                  w_kj = m_k mu_j - m_j mu_k
//   (J_m^T mu)_k = -G sum_{j != k} s_kj d_kj . mu_j
// One pair costs 29 VALU ops + 1 v_rsq_f32 in FP32.  The caller adds G (and the step) once.
// The self pair (j == k) must go through the MASK path for every eps: w is formed with an FMA,
// so for j == k it would be the rounding error of m_k mu_k rather than exactly 0.
#pragma once

#include <hip/hip_runtime.h>

namespace grav {

// Accumulates, for K owners per lane, the partials P += s (w - 3 (d.w)/q d) and
// Pg += s (d . mu_j) of one j-body (q = x, y, z, m; u = mu_j).
template <int K, bool MASK>
__device__ __forceinline__ void vjp_pair(const float4 q, const float4 u, const int j,
                                         const float (&xk)[K], const float (&yk)[K],
                                         const float (&zk)[K], const float (&mk)[K],
                                         const float (&uxk)[K], const float (&uyk)[K],
                                         const float (&uzk)[K], const int (&kk)[K],
                                         const float eps2, float (&px)[K], float (&py)[K],
                                         float (&pz)[K], float (&pg)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float dx = q.x - xk[k];
    const float dy = q.y - yk[k];
    const float dz = q.z - zk[k];
    float r2 = __builtin_fmaf(dx, dx, eps2);
    r2 = __builtin_fmaf(dy, dy, r2);
    r2 = __builtin_fmaf(dz, dz, r2);
    float inv = __builtin_amdgcn_rsqf(r2);
    if (MASK) inv = (j == kk[k]) ? 0.0f : inv;
    const float inv2 = inv * inv;
    const float s = inv * inv2;
    const float wx = __builtin_fmaf(mk[k], u.x, -(q.w * uxk[k]));
    const float wy = __builtin_fmaf(mk[k], u.y, -(q.w * uyk[k]));
    const float wz = __builtin_fmaf(mk[k], u.z, -(q.w * uzk[k]));
    float dw = dx * wx;
    dw = __builtin_fmaf(dy, wy, dw);
    dw = __builtin_fmaf(dz, wz, dw);
    const float c3 = (inv2 * dw) * 3.0f;
    px[k] = __builtin_fmaf(s, __builtin_fmaf(-c3, dx, wx), px[k]);
    py[k] = __builtin_fmaf(s, __builtin_fmaf(-c3, dy, wy), py[k]);
    pz[k] = __builtin_fmaf(s, __builtin_fmaf(-c3, dz, wz), pz[k]);
    float dm = dx * u.x;
    dm = __builtin_fmaf(dy, u.y, dm);
    dm = __builtin_fmaf(dz, u.z, dm);
    pg[k] = __builtin_fmaf(s, dm, pg[k]);
  }
}

// Knuth's TwoSum: hi + p == s + e exactly; hi <- s, lo <- lo + e.
__device__ __forceinline__ void two_sum_acc(float &hi, float &lo, float p) {
  const float s = hi + p;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (p - bp);
  hi = s;
  lo += e;
}

}  // namespace grav
