// Adjoint (reverse mode) of the time step, single universe.  Not present in the reference;
// the step it differentiates is the reference's stage 1 (src/gravitation/kernel/py1.py:57-64,
// pc2.py:60-90, with softening) followed by the semi-implicit Euler stage 2
// (src/gravitation/kernel/_base_.py:58-62):  v' = v + h a(r, m),  r' = r + h v'.
//
// With rho = dL/dr', nu = dL/dv' and mu = nu + h rho, one backward step is
//   rho <- rho + h J_r^T mu,   gamma <- gamma + h J_m^T mu,   nu <- mu
// (gamma = dL/dm).  grav_adjoint_mu does the element-wise part (mu, nu, and an FP32 copy of
// mu scaled per universe by a power of two); grav_stage1_vjp_f32 the all-pairs part, owner-computes like stage 1: lane l owns K bodies, j-bodies (x,y,z,m)
// and mu_j are wave-uniform loads (SMEM), S wavefronts split the j-range, FP32 pair terms with
// blocked TwoSum partials, FP64 adjoint state.
#include <algorithm>
#include <cfloat>
#include <climits>

#include "grav_common.h"
#include "pair_vjp.h"
#include "stage1_f32.h"

namespace grav {
namespace {

constexpr int kMuBlock = 1024;

// mu = nu + h rho for the bodies of universe u = blockIdx.x (bodies [offsets[u],
// offsets[u+1]), or all n bodies when offsets is NULL), stored to nu in FP64 and to mu4 in
// FP32 as mu * 2^-E, with the universe's scale exponent E in mu4.w (its bits, as an int).
// E is frexp's exponent of the universe's largest finite |mu| component (0 if it has none),
// so |mu4.xyz| < 1.  The VJP is linear in mu and its pair terms carry factors up to
// eps^-3 * m * N, so unscaled FP32 cotangents overflow once the adjoint has grown by ~1e28
// (chaotic trajectories reach that within a few hundred steps); scaled ones cannot, and the
// VJP kernels multiply their FP64 update by 2^E.  Scaling by a power of two is exact, so the
// result equals that of unscaled FP32 mu whenever the latter neither overflows nor
// underflows.  One workgroup per universe: the scale is a reduction over the universe.
__global__ __launch_bounds__(kMuBlock) void k_adjoint_mu(const double *__restrict__ rho,
                                                         double *__restrict__ nu,
                                                         float4 *__restrict__ mu4, int n,
                                                         double dt,
                                                         const int64_t *__restrict__ offsets) {
#pragma clang fp contract(off)
  __shared__ int wmax[kMuBlock / kWave];
  const int u = blockIdx.x;
  const int b0 = offsets ? static_cast<int>(offsets[u]) : 0;
  const int b1 = offsets ? static_cast<int>(offsets[u + 1]) : n;
  // Bodies are taken U at a time per thread with all their loads issued before any store
  // (nu is read and written, so the compiler would otherwise keep one body in flight).
  constexpr int U = 4;
  int e = INT_MIN;
  auto body = [&](int i, const double (&nv)[3], const double (&rv)[3]) {
    double a = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double mu = nv[c] + dt * rv[c];
      nu[c * n + i] = mu;
      a = fmax(a, fabs(mu));  // NaN components drop out here
    }
    if (a > 0.0 && a <= DBL_MAX) {
      int x;
      frexp(a, &x);
      e = max(e, x);
    }
  };
  int i0 = b0 + static_cast<int>(threadIdx.x);
  for (; i0 + (U - 1) * kMuBlock < b1; i0 += U * kMuBlock) {
    double nv[U][3], rv[U][3];
#pragma unroll
    for (int t = 0; t < U; ++t)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        nv[t][c] = nu[c * n + i0 + t * kMuBlock];
        rv[t][c] = rho[c * n + i0 + t * kMuBlock];
      }
#pragma unroll
    for (int t = 0; t < U; ++t) body(i0 + t * kMuBlock, nv[t], rv[t]);
  }
  for (; i0 < b1; i0 += kMuBlock) {
    double nv[3], rv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      nv[c] = nu[c * n + i0];
      rv[c] = rho[c * n + i0];
    }
    body(i0, nv, rv);
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) e = max(e, __shfl_xor(e, off));
  if ((threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x / kWave] = e;
  __syncthreads();
  int E = wmax[0];
#pragma unroll
  for (int w = 1; w < kMuBlock / kWave; ++w) E = max(E, wmax[w]);
  if (E == INT_MIN) E = 0;
  // Each thread reads back only the nu entries it wrote above.
  auto scaled = [&](int i, const double (&m)[3]) {
    mu4[i] = make_float4(static_cast<float>(ldexp(m[0], -E)), static_cast<float>(ldexp(m[1], -E)),
                         static_cast<float>(ldexp(m[2], -E)), __int_as_float(E));
  };
  i0 = b0 + static_cast<int>(threadIdx.x);
  for (; i0 + (U - 1) * kMuBlock < b1; i0 += U * kMuBlock) {
    double m[U][3];
#pragma unroll
    for (int t = 0; t < U; ++t)
#pragma unroll
      for (int c = 0; c < 3; ++c) m[t][c] = nu[c * n + i0 + t * kMuBlock];
#pragma unroll
    for (int t = 0; t < U; ++t) scaled(i0 + t * kMuBlock, m[t]);
  }
  for (; i0 < b1; i0 += kMuBlock) {
    double m[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] = nu[c * n + i0];
    scaled(i0, m);
  }
}

// An exponent biased into an order-preserving unsigned word (0 stands for INT_MIN: none).
__device__ __forceinline__ unsigned exp_bias(int e) {
  return static_cast<unsigned>(e) + 0x80000000u;
}

// mu = nu + h rho of body i into nu and m (FP64, no contraction, as k_adjoint_mu), returning
// frexp's exponent of its largest finite |mu| component (INT_MIN if none).
__device__ __forceinline__ int mu_body(const double *__restrict__ rho, double *__restrict__ nu,
                                       int n, int i, double dt, double (&m)[3]) {
#pragma clang fp contract(off)
  double nv[3], rv[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    nv[c] = nu[c * n + i];
    rv[c] = rho[c * n + i];
  }
  double a = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double mu = nv[c] + dt * rv[c];
    nu[c * n + i] = mu;
    m[c] = mu;
    a = fmax(a, fabs(mu));
  }
  int x = INT_MIN;
  if (a > 0.0 && a <= DBL_MAX) frexp(a, &x);
  return x;
}

constexpr int kMuTile = 1024;  // bodies per workgroup of the two-kernel form (256 threads)

// Two-kernel k_adjoint_mu for one universe: each workgroup does kMuTile bodies and writes
// their largest exponent to emax[blockIdx.x] ...
__global__ __launch_bounds__(256) void k_adjoint_mu_tiles(const double *__restrict__ rho,
                                                          double *__restrict__ nu,
                                                          int *__restrict__ emax, int n,
                                                          double dt) {
  __shared__ int wmax[256 / kWave];
  int e = INT_MIN;
#pragma unroll
  for (int t = 0; t < kMuTile / 256; ++t) {
    const int i = blockIdx.x * kMuTile + t * 256 + static_cast<int>(threadIdx.x);
    double m[3];
    if (i < n) e = max(e, mu_body(rho, nu, n, i, dt, m));
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) e = max(e, __shfl_xor(e, off));
  if ((threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x / kWave] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = wmax[0];
#pragma unroll
    for (int w = 1; w < 256 / kWave; ++w) m = max(m, wmax[w]);
    emax[blockIdx.x] = m;
  }
}

// ... and then every workgroup takes the universe's exponent E as the largest of emax[0..nt)
// (0 if none) and writes its bodies' mu4 = (mu * 2^-E, E), as k_adjoint_mu does.
__global__ __launch_bounds__(256) void k_adjoint_mu_scale(const double *__restrict__ nu,
                                                          float4 *__restrict__ mu4,
                                                          const int *__restrict__ emax, int n,
                                                          int nt) {
  __shared__ int wmax[256 / kWave];
  int e = INT_MIN;
  for (int t = static_cast<int>(threadIdx.x); t < nt; t += 256) e = max(e, emax[t]);
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) e = max(e, __shfl_xor(e, off));
  if ((threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x / kWave] = e;
  __syncthreads();
  int E = wmax[0];
#pragma unroll
  for (int w = 1; w < 256 / kWave; ++w) E = max(E, wmax[w]);
  if (E == INT_MIN) E = 0;
#pragma unroll
  for (int t = 0; t < kMuTile / 256; ++t) {
    const int i = blockIdx.x * kMuTile + t * 256 + static_cast<int>(threadIdx.x);
    if (i < n)
      mu4[i] = make_float4(static_cast<float>(ldexp(nu[i], -E)),
                           static_cast<float>(ldexp(nu[n + i], -E)),
                           static_cast<float>(ldexp(nu[2 * n + i], -E)), __int_as_float(E));
  }
}

template <int K, bool MASK>
__device__ __forceinline__ void vjp_segment(const float4 *__restrict__ posm,
                                            const float4 *__restrict__ mu4, int j0, int j1,
                                            int jblock, const float (&xk)[K],
                                            const float (&yk)[K], const float (&zk)[K],
                                            const float (&mk)[K], const float (&uxk)[K],
                                            const float (&uyk)[K], const float (&uzk)[K],
                                            const int (&kk)[K], float eps2, float (&hi)[4][K],
                                            float (&lo)[4][K]) {
  for (int jb = j0, je; jb < j1; jb = je) {
    je = (j1 - jb > jblock) ? jb + jblock : j1;
    float px[K], py[K], pz[K], pg[K];
#pragma unroll
    for (int k = 0; k < K; ++k) px[k] = py[k] = pz[k] = pg[k] = 0.0f;
    // j-bodies (position and cotangent) per loop step: two with one body per lane, four
    // otherwise.  Same order of pairs, same bits.  MI355X, Plummer, j-split VJP (1, 8, 16):
    // 2^18 35.6-35.7 against 36.7-37.1 ms, 16384 0.188 against 0.195 ms, a 4096-body
    // universe's forward + adjoint over 200 steps 7.48-7.50 against 7.56-7.59 ms; with two per
    // step the owner-computes VJP of two bodies per lane took 4.41 against 3.96 ms at 2^16.
    constexpr int JB = K == 1 ? 2 : 4;
    int j = jb;
    for (; j + JB <= je; j += JB) {
      float4 q[JB], u[JB];
#pragma unroll
      for (int t = 0; t < JB; ++t) {
        q[t] = posm[j + t];
        u[t] = mu4[j + t];
      }
#pragma unroll
      for (int t = 0; t < JB; ++t)
        vjp_pair<K, MASK>(q[t], u[t], j + t, xk, yk, zk, mk, uxk, uyk, uzk, kk, eps2, px, py,
                          pz, pg);
    }
    for (; j < je; ++j)
      vjp_pair<K, MASK>(posm[j], mu4[j], j, xk, yk, zk, mk, uxk, uyk, uzk, kk, eps2, px, py, pz,
                        pg);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      two_sum_acc(hi[0][k], lo[0][k], px[k]);
      two_sum_acc(hi[1][k], lo[1][k], py[k]);
      two_sum_acc(hi[2][k], lo[2][k], pz[k]);
      two_sum_acc(hi[3][k], lo[3][k], pg[k]);
    }
  }
}

// The workgroup at (bx, by) of an (x, ny) grid owns bodies [items[bx].x, items[bx].y) against
// j in [items[bx].z, items[bx].w), or, with items == NULL, i-tile bx of the share [i0, i0 + m)
// of one universe of n bodies against all of it, or, with part != NULL, against the by-th of
// ny equal j-ranges of [0, n) (the j-split form: the unscaled partials go to
// part[by][c][i - i0], c < 4, in FP64, and a reduce kernel applies them).  rho, gamma and part
// rows are i - i0 with stride m (a whole universe: i0 = 0, m = n; with items i0 = 0 and m is the
// stride).  The self pair always takes the masked path (see pair_vjp.h), whatever eps is.
// red: (S - 1) * 8 * (64 * K) floats of LDS.
// eslot != NULL (the chained backward step, j-split form only): the first thread of the grid
// moves the exponent that the previous chained reduce gathered, eslot[1], to eslot[0], where
// this step's reduce reads it (none gathered yet: mu4's own exponent), and clears eslot[1] for
// this step's reduce to gather into.  Nothing in this kernel reads eslot otherwise.
template <int K, int S>
__device__ __forceinline__ void stage1_vjp_f32_block(
    const float4 *__restrict__ posm, const float4 *__restrict__ mu4, double *__restrict__ rho,
    double *__restrict__ gamma, int n, int i0, int m, double hG, float eps2, int jblock,
    const int4 *__restrict__ items, double *__restrict__ part, unsigned *__restrict__ eslot,
    int bx, int by, int ny, float *red) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (eslot != nullptr && bx == 0 && by == 0 && threadIdx.x == 0) {
    const unsigned b = eslot[1];
    eslot[0] = b != 0u ? b : exp_bias(__float_as_int(mu4[0].w));
    eslot[1] = 0u;
  }
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
    jlo = 0;
    jhi = n;
    if (part != nullptr) {
      const int ns = ny;
      const int span = ((n + ns - 1) / ns + 3) & ~3;
      jlo = min(n, by * span);
      jhi = min(n, jlo + span);
    }
  }
  float xk[K], yk[K], zk[K], mk[K], uxk[K], uyk[K], uzk[K];
  int kk[K], ek[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = ibase + k * kWave + lane;
    kk[k] = i;
    const int ic = min(i, iend - 1);
    const float4 p = posm[ic];
    const float4 u = mu4[ic];
    ek[k] = __float_as_int(u.w);  // the universe's scale exponent (0: unscaled)
    xk[k] = p.x;
    yk[k] = p.y;
    zk[k] = p.z;
    mk[k] = p.w;
    uxk[k] = u.x;
    uyk[k] = u.y;
    uzk[k] = u.z;
  }
  float hi[4][K], lo[4][K];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int k = 0; k < K; ++k) hi[c][k] = lo[c][k] = 0.0f;

  const int chunk = ((jhi - jlo + S - 1) / S + 3) & ~3;
  const int j0 = min(jhi, jlo + wave * chunk);
  const int j1 = min(jhi, j0 + chunk);
  {
    // the masked segment is the workgroup's own bodies, so the compensation blocks depend on
    // where its i-tile starts: a share's rows equal the whole universe's bit for bit when the
    // share starts on a tile boundary (a multiple of 64 * K), and to rounding otherwise
    const int a = min(j1, max(j0, ibase));
    const int b = max(a, min(j1, iend));
    // eps2 in a VGPR, so that the first r2 FMA takes no SGPR operand (2^18 j-split VJP:
    // 34.85 against 35.27-35.37 ms)
    float eps2v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(eps2v) : "s"(eps2));
    vjp_segment<K, false>(posm, mu4, j0, a, jblock, xk, yk, zk, mk, uxk, uyk, uzk, kk, eps2v, hi,
                          lo);
    vjp_segment<K, true>(posm, mu4, a, b, jblock, xk, yk, zk, mk, uxk, uyk, uzk, kk, eps2v, hi,
                         lo);
    vjp_segment<K, false>(posm, mu4, b, j1, jblock, xk, yk, zk, mk, uxk, uyk, uzk, kk, eps2v, hi,
                          lo);
  }

  if constexpr (S > 1) {
    constexpr int L = kWave * K;
    if (wave > 0) {
      float *r = red + (wave - 1) * 8 * L;
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          r[c * L + k * kWave + lane] = hi[c][k];
          r[(4 + c) * L + k * kWave + lane] = lo[c][k];
        }
    }
    __syncthreads();
    if (wave != 0) return;
    for (int w = 1; w < S; ++w) {
      const float *r = red + (w - 1) * 8 * L;
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          two_sum_acc(hi[c][k], lo[c][k], r[c * L + k * kWave + lane]);
          lo[c][k] += r[(4 + c) * L + k * kWave + lane];
        }
    }
  }

  if (part != nullptr) {
    double *o = part + static_cast<size_t>(by) * 4 * m;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = kk[k];
      if (i < iend) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          o[c * m + i - i0] = static_cast<double>(hi[c][k]) + static_cast<double>(lo[c][k]);
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = kk[k];
    if (i < iend) {
      const double hs = ldexp(hG, ek[k]);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        rho[c * m + i - i0] +=
            hs * (static_cast<double>(hi[c][k]) + static_cast<double>(lo[c][k]));
      gamma[i - i0] -= hs * (static_cast<double>(hi[3][k]) + static_cast<double>(lo[3][k]));
    }
  }
}

// Workgroup (blockIdx.x, blockIdx.y) of stage1_vjp_f32_block.
template <int K, int S>
__global__ __launch_bounds__(kWave *S) void k_stage1_vjp_f32(
    const float4 *__restrict__ posm, const float4 *__restrict__ mu4, double *__restrict__ rho,
    double *__restrict__ gamma, int n, int i0, int m, double hG, float eps2, int jblock,
    const int4 *__restrict__ items, double *__restrict__ part, unsigned *__restrict__ eslot) {
  extern __shared__ float red[];  // (S - 1) * 8 * (64 * K)
  stage1_vjp_f32_block<K, S>(posm, mu4, rho, gamma, n, i0, m, hG, eps2, jblock, items, part,
                             eslot, static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y),
                             static_cast<int>(gridDim.y), red);
}

// rho[c][i] += hs * sum_p part[p][c][i], gamma[i] -= hs * sum_p part[p][3][i], hs = hG 2^E
// with the universe's exponent E from mu4[i].w; sums in FP64 in range order.
__device__ __forceinline__ void split_reduce_vjp_body(const double *__restrict__ part,
                                                      const float4 *mu4,
                                                      double *__restrict__ rho,
                                                      double *__restrict__ gamma, int n,
                                                      int splits, double hG, int i) {
  const size_t stride = static_cast<size_t>(4) * n;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int p = 0;
  for (; p + 2 <= splits; p += 2) {  // eight loads in flight, added in range order
    double v[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) v[t][c] = part[(p + t) * stride + static_cast<size_t>(c) * n + i];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) s[c] += v[t][c];
  }
  for (; p < splits; ++p) {
    const double *q = part + p * stride + i;
#pragma unroll
    for (int c = 0; c < 4; ++c) s[c] += q[static_cast<size_t>(c) * n];
  }
  const double hs = ldexp(hG, __float_as_int(mu4[i].w));
#pragma unroll
  for (int c = 0; c < 3; ++c) rho[static_cast<size_t>(c) * n + i] += hs * s[c];
  gamma[i] -= hs * s[3];
}

__global__ __launch_bounds__(256) void k_split_reduce_vjp(const double *__restrict__ part,
                                                          const float4 *__restrict__ mu4,
                                                          double *__restrict__ rho,
                                                          double *__restrict__ gamma, int n,
                                                          int splits, double hG) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) split_reduce_vjp_body(part, mu4, rho, gamma, n, splits, hG, i);
}

// The chained backward step's reduce: k_split_reduce_vjp, then, for the next backward step,
// mu = nu + h rho into nu (as k_adjoint_mu_tiles) and its FP32 copy mu * 2^-E into mu4 (each
// thread reads body i's current exponent before it overwrites body i's entry; no other thread
// touches it).  E is the exponent gathered for the previous mu (eslot[0], put there by the
// VJP kernel), so the scale lags the cotangent by one step: |mu4| may exceed 1 by the growth
// of one step, far inside the FP32 headroom of the pair terms, and since scaling by a power
// of two is exact the VJP's result is the same as with the current exponent unless a scaled
// component is subnormal.  The exponent of this mu goes into eslot[1] by an atomic max (the
// order of the workgroups does not change a max).
// Bodies [256 bx, 256 bx + 256) of the chained reduce (every thread of the workgroup calls
// it: the exponent's maximum goes through LDS).
__device__ __forceinline__ void split_reduce_vjp_chain_block(const double *__restrict__ part,
                                                             float4 *mu4,
                                                             double *__restrict__ rho,
                                                             double *__restrict__ nu,
                                                             double *__restrict__ gamma, int n,
                                                             int splits, double hG, double dt,
                                                             unsigned *__restrict__ eslot,
                                                             int bx) {
  __shared__ int wmax[256 / kWave];
  const int i = bx * 256 + static_cast<int>(threadIdx.x);
  int e = INT_MIN;
  if (i < n) {
    split_reduce_vjp_body(part, mu4, rho, gamma, n, splits, hG, i);
    double m[3];
    e = mu_body(rho, nu, n, i, dt, m);
    const int E = static_cast<int>(eslot[0] - 0x80000000u);
    mu4[i] = make_float4(static_cast<float>(ldexp(m[0], -E)), static_cast<float>(ldexp(m[1], -E)),
                         static_cast<float>(ldexp(m[2], -E)), __int_as_float(E));
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) e = max(e, __shfl_xor(e, off));
  if ((threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x / kWave] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    int b = wmax[0];
#pragma unroll
    for (int w = 1; w < 256 / kWave; ++w) b = max(b, wmax[w]);
    if (b != INT_MIN) atomicMax(eslot + 1, exp_bias(b));
  }
}

__global__ __launch_bounds__(256) void k_split_reduce_vjp_chain(const double *__restrict__ part,
                                                                float4 *mu4,
                                                                double *__restrict__ rho,
                                                                double *__restrict__ nu,
                                                                double *__restrict__ gamma,
                                                                int n, int splits, double hG,
                                                                double dt,
                                                                unsigned *__restrict__ eslot) {
  split_reduce_vjp_chain_block(part, mu4, rho, nu, gamma, n, splits, hG, dt, eslot,
                               static_cast<int>(blockIdx.x));
}

// The paired backward step: ONE launch holds a chained backward step's VJP partials (the
// first tiles_a * splits_a workgroups, at positions pa) and, independent of them, a recomputed
// forward step's stage-1 partials (the other tiles_r * splits_r workgroups, at positions pr),
// so the recomputation of the previous checkpoint segment fills the GPU beside the adjoint's
// critical path without a second stream.  Each half does exactly what k_stage1_vjp_f32 and
// k_stage1_f32 (j-split) do with the same shapes, so the bits are theirs.
template <int KA, int KR, int S>
__global__ __launch_bounds__(kWave *S) void k_backward_pair(
    const float4 *__restrict__ pa, const float4 *__restrict__ mu4, double *__restrict__ part_a,
    unsigned *__restrict__ eslot, int tiles_a, int splits_a, const float4 *__restrict__ pr,
    double *__restrict__ part_r, int mask, int tiles_r, int splits_r, int n, float eps2) {
  extern __shared__ float red[];  // max((S - 1) * 8 * 64 * KA, (S - 1) * 6 * 64 * KR)
  const int b = static_cast<int>(blockIdx.x);
  const int nba = tiles_a * splits_a;
  if (b < nba) {
    stage1_vjp_f32_block<KA, S>(pa, mu4, nullptr, nullptr, n, 0, n, 0.0, eps2, 256, nullptr,
                                part_a, eslot, b % tiles_a, b / tiles_a, splits_a, red);
  } else {
    const int c = b - nba;
    // j-bodies per pair-loop step as pair_loop_width takes them for the paired launches, which
    // serve universes of a few thousand bodies (hundreds of workgroups): two with one body
    // per lane (its loop loads the next step ahead), eight otherwise
    constexpr int JB = KR == 1 ? 2 : 8;
    stage1_f32_block<KR, S, JB>(pr, nullptr, n, 0, n, 1.0f, eps2, mask, 256, nullptr, 0, part_r,
                                c % tiles_r, c / tiles_r, splits_r, red);
  }
}

// ... and its reduce: the first nb workgroups the chained reduce of the backward step, the
// other nb the reduce + FP64-master stage 2 of the recomputed step (positions pr -> pr_out).
__global__ __launch_bounds__(256) void k_backward_pair_reduce(
    const double *__restrict__ part_a, float4 *mu4, double *__restrict__ rho,
    double *__restrict__ nu, double *__restrict__ gamma, int splits_a, double hG,
    unsigned *__restrict__ eslot, const double *__restrict__ part_r, const float4 *pr,
    float4 *pr_out, double *__restrict__ r64, double *__restrict__ v64, int splits_r, double G,
    double dt, int n, int nb) {
  const int b = static_cast<int>(blockIdx.x);
  if (b < nb) {
    split_reduce_vjp_chain_block(part_a, mu4, rho, nu, gamma, n, splits_a, hG, dt, eslot, b);
  } else {
    const int i = (b - nb) * 256 + static_cast<int>(threadIdx.x);
    if (i < n) split_step_body(part_r, pr, pr_out, r64, v64, n, splits_r, G, dt, i);
  }
}

// i-tiles of the share [i0, i0 + m) (items == NULL) or the nitems items.
template <int K, int S>
int launch_vjp(const float4 *posm, const float4 *mu4, double *rho, double *gamma, int n, int i0,
               int m, double hG, float eps2, const int4 *items, int nitems,
               hipStream_t stream) {
  const int per_block = kWave * K;
  const int grid = items ? nitems : (m + per_block - 1) / per_block;
  const size_t lds = (S > 1) ? sizeof(float) * (S - 1) * 8 * per_block : 0;
  if (grid == 0) return GRAV_OK;
  hipLaunchKernelGGL((k_stage1_vjp_f32<K, S>), dim3(grid), dim3(kWave * S), lds, stream, posm,
                     mu4, rho, gamma, n, i0, m, hG, eps2, 256, items,
                     static_cast<double *>(nullptr), static_cast<unsigned *>(nullptr));
  return check_launch("k_stage1_vjp_f32");
}

// The j-split partials part[splits][4][m] of the share [i0, i0 + m).
template <int K, int S>
int launch_vjp_split(const float4 *posm, const float4 *mu4, double *part, int n, int i0, int m,
                     float eps2, int splits, unsigned *eslot, hipStream_t stream) {
  const int per_block = kWave * K;
  const int tiles = (m + per_block - 1) / per_block;
  const size_t lds = (S > 1) ? sizeof(float) * (S - 1) * 8 * per_block : 0;
  hipLaunchKernelGGL((k_stage1_vjp_f32<K, S>), dim3(tiles, splits), dim3(kWave * S), lds,
                     stream, posm, mu4, static_cast<double *>(nullptr),
                     static_cast<double *>(nullptr), n, i0, m, 0.0, eps2, 256,
                     static_cast<const int4 *>(nullptr), part, eslot);
  return check_launch("k_stage1_vjp_f32 (j-split)");
}

}  // namespace

int adjoint_mu_launch(const double *rho, double *nu, float4 *mu4, int n, double dt,
                      const int64_t *offsets, int n_universes, hipStream_t stream) {
  hipLaunchKernelGGL(k_adjoint_mu, dim3(static_cast<unsigned>(n_universes)), dim3(kMuBlock), 0,
                     stream, rho, nu, mu4, n, dt, offsets);
  return check_launch("k_adjoint_mu");
}

int stage1_vjp_launch(const float4 *posm, const float4 *mu4, double *rho, double *gamma, int n,
                      double hG, float eps2, int K, int S, const int4 *items, int nitems,
                      hipStream_t stream) {
#define GRAV_CASE(KK, SS)                                                                     \
  if (K == KK && S == SS)                                                                     \
    return launch_vjp<KK, SS>(posm, mu4, rho, gamma, n, 0, n, hG, eps2, items, nitems, stream);
  GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
  GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4) GRAV_CASE(2, 8)
  GRAV_CASE(4, 1) GRAV_CASE(4, 2) GRAV_CASE(4, 4)
#undef GRAV_CASE
  return fail(GRAV_EINVAL, "VJP launch shape (%d, %d) not supported", K, S);
}

namespace {
int vjp_split_partials(const float4 *posm, const float4 *mu4, double *part, int n, int i0, int m,
                       float eps2, int K, int S, int splits, unsigned *eslot,
                       hipStream_t stream) {
  if (!split_supported(K, S, splits))
    return fail(GRAV_EINVAL, "j-split shape (%d, %d, %d) not supported", K, S, splits);
#define GRAV_CASE(KK, SS)   \
  if (K == KK && S == SS) \
    return launch_vjp_split<KK, SS>(posm, mu4, part, n, i0, m, eps2, splits, eslot, stream);
  GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
  GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4) GRAV_CASE(2, 8)
#undef GRAV_CASE
  return GRAV_OK;
}
}  // namespace

int stage1_vjp_split_launch(const float4 *posm, const float4 *mu4, double *rho, double *gamma,
                            double *part, int n, double hG, float eps2, int K, int S, int splits,
                            hipStream_t stream) {
  const int rc =
      vjp_split_partials(posm, mu4, part, n, 0, n, eps2, K, S, splits, nullptr, stream);
  if (rc != GRAV_OK) return rc;
  hipLaunchKernelGGL(k_split_reduce_vjp, dim3((n + 255) / 256), dim3(256), 0, stream, part, mu4,
                     rho, gamma, n, splits, hG);
  return check_launch("k_split_reduce_vjp");
}

// The VJP of the bodies [i0, i0 + m) of n (one GPU's share of the rows; mu4 holds all n
// cotangents, with one scale exponent for the universe): owner-computes with splits == 1, else
// the j-split form with part[splits][4][m]; rho (3, m) and gamma (m) are the share's rows.
int stage1_vjp_shard_launch(const float4 *posm, const float4 *mu4, double *rho, double *gamma,
                            double *part, int n, int i0, int m, double hG, float eps2, int K,
                            int S, int splits, hipStream_t stream) {
  if (splits == 1) {
#define GRAV_CASE(KK, SS)                                                                   \
  if (K == KK && S == SS)                                                                   \
    return launch_vjp<KK, SS>(posm, mu4, rho, gamma, n, i0, m, hG, eps2, nullptr, 0, stream);
    GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
    GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4) GRAV_CASE(2, 8)
    GRAV_CASE(4, 1) GRAV_CASE(4, 2) GRAV_CASE(4, 4)
#undef GRAV_CASE
    return fail(GRAV_EINVAL, "VJP launch shape (%d, %d) not supported", K, S);
  }
  const int rc = vjp_split_partials(posm, mu4, part, n, i0, m, eps2, K, S, splits, nullptr,
                                    stream);
  if (rc != GRAV_OK) return rc;
  hipLaunchKernelGGL(k_split_reduce_vjp, dim3((m + 255) / 256), dim3(256), 0, stream, part,
                     mu4 + i0, rho, gamma, m, splits, hG);
  return check_launch("k_split_reduce_vjp");
}

namespace {
template <int KA, int KR, int S>
int launch_backward_pair(const float4 *pa, float4 *mu4, double *rho, double *nu, double *gamma,
                         unsigned *eslot, double *part_a, int splits_a, const float4 *pr,
                         float4 *pr_out, double *r64, double *v64, double *part_r, int splits_r,
                         int mask, int n, double G, double hG, double dt, float eps2,
                         hipStream_t stream) {
  const int tiles_a = (n + kWave * KA - 1) / (kWave * KA);
  const int tiles_r = (n + kWave * KR - 1) / (kWave * KR);
  const size_t lds = sizeof(float) * (S - 1) * std::max(8 * kWave * KA, 6 * kWave * KR);
  const int grid = tiles_a * splits_a + tiles_r * splits_r;
  hipLaunchKernelGGL((k_backward_pair<KA, KR, S>), dim3(grid), dim3(kWave * S), lds, stream, pa,
                     mu4, part_a, eslot, tiles_a, splits_a, pr, part_r, mask, tiles_r, splits_r,
                     n, eps2);
  int rc = check_launch("k_backward_pair");
  if (rc != GRAV_OK) return rc;
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_backward_pair_reduce, dim3(2 * nb), dim3(256), 0, stream, part_a, mu4,
                     rho, nu, gamma, splits_a, hG, eslot, part_r, pr, pr_out, r64, v64, splits_r,
                     G, dt, n, nb);
  return check_launch("k_backward_pair_reduce");
}
}  // namespace

bool backward_pair_supported(int KA, int KR, int S) {
  return (KA == 1 || KA == 2) && (KR == 1 || KR == 2) && (S == 4 || S == 8);
}

int backward_pair_launch(const float4 *pa, float4 *mu4, double *rho, double *nu, double *gamma,
                         unsigned *eslot, double *part_a, int KA, int splits_a,
                         const float4 *pr, float4 *pr_out, double *r64, double *v64,
                         double *part_r, int KR, int splits_r, int S, int mask, int n, double G,
                         double hG, double dt, float eps2, hipStream_t stream) {
#define GRAV_CASE(A, R, SS)                                                                   \
  if (KA == A && KR == R && S == SS)                                                          \
    return launch_backward_pair<A, R, SS>(pa, mu4, rho, nu, gamma, eslot, part_a, splits_a,   \
                                          pr, pr_out, r64, v64, part_r, splits_r, mask, n, G, \
                                          hG, dt, eps2, stream);
  GRAV_CASE(1, 1, 8) GRAV_CASE(1, 2, 8) GRAV_CASE(2, 1, 8) GRAV_CASE(2, 2, 8)
  GRAV_CASE(1, 1, 4) GRAV_CASE(1, 2, 4) GRAV_CASE(2, 1, 4) GRAV_CASE(2, 2, 4)
#undef GRAV_CASE
  return fail(GRAV_EINVAL, "paired backward shape (%d, %d, %d) not supported", KA, KR, S);
}

// The chained backward step: the VJP's partials (rotating eslot), then the reduce with the
// next mu (next_mu) or the plain reduce (the chain's last step).
int stage1_vjp_split_chain_launch(const float4 *posm, float4 *mu4, double *rho, double *nu,
                                  double *gamma, double *part, int n, double hG, double dt,
                                  float eps2, int K, int S, int splits, int next_mu,
                                  unsigned *eslot, hipStream_t stream) {
  const int rc =
      vjp_split_partials(posm, mu4, part, n, 0, n, eps2, K, S, splits, eslot, stream);
  if (rc != GRAV_OK) return rc;
  if (!next_mu) {
    hipLaunchKernelGGL(k_split_reduce_vjp, dim3((n + 255) / 256), dim3(256), 0, stream, part,
                       mu4, rho, gamma, n, splits, hG);
    return check_launch("k_split_reduce_vjp");
  }
  hipLaunchKernelGGL(k_split_reduce_vjp_chain, dim3((n + 255) / 256), dim3(256), 0, stream, part,
                     mu4, rho, nu, gamma, n, splits, hG, dt, eslot);
  return check_launch("k_split_reduce_vjp_chain");
}

}  // namespace grav

using namespace grav;

extern "C" {

int grav_adjoint_mu(const void *rho64, void *nu64, void *mu4, int64_t n, double dt,
                    void *stream) {
  clear_error();
  GRAV_REQUIRE(rho64 != nullptr && nu64 != nullptr && mu4 != nullptr,
               "grav_adjoint_mu: NULL buffer");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "grav_adjoint_mu: n must be in [1, 2^29] (got %lld)",
               (long long)n);
  GRAV_REQUIRE(aligned(rho64, 8) && aligned(nu64, 8) && aligned(mu4, 16),
               "grav_adjoint_mu: misaligned buffer");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "grav_adjoint_mu: dt must be finite");
  return adjoint_mu_launch(static_cast<const double *>(rho64), static_cast<double *>(nu64),
                           static_cast<float4 *>(mu4), static_cast<int>(n), dt, nullptr, 1,
                           static_cast<hipStream_t>(stream));
}

int grav_adjoint_mu_ws(const void *rho64, void *nu64, void *mu4, int64_t n, double dt,
                       void *workspace, int64_t workspace_bytes, void *stream) {
  clear_error();
  GRAV_REQUIRE(rho64 != nullptr && nu64 != nullptr && mu4 != nullptr,
               "grav_adjoint_mu_ws: NULL buffer");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_adjoint_mu_ws: n must be in [1, 2^29] (got %lld)", (long long)n);
  GRAV_REQUIRE(aligned(rho64, 8) && aligned(nu64, 8) && aligned(mu4, 16),
               "grav_adjoint_mu_ws: misaligned buffer");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "grav_adjoint_mu_ws: dt must be finite");
  const int nt = static_cast<int>((n + kMuTile - 1) / kMuTile);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 4) &&
                   workspace_bytes >= int64_t(4) * nt,
               "grav_adjoint_mu_ws: workspace must be 4-byte aligned and hold %lld bytes",
               (long long)(int64_t(4) * nt));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nn = static_cast<int>(n);
  int *emax = static_cast<int *>(workspace);
  hipLaunchKernelGGL(k_adjoint_mu_tiles, dim3(nt), dim3(256), 0, s,
                     static_cast<const double *>(rho64), static_cast<double *>(nu64), emax, nn,
                     dt);
  int rc = check_launch("k_adjoint_mu_tiles");
  if (rc != GRAV_OK) return rc;
  hipLaunchKernelGGL(k_adjoint_mu_scale, dim3(nt), dim3(256), 0, s,
                     static_cast<const double *>(nu64), static_cast<float4 *>(mu4), emax, nn,
                     nt);
  return check_launch("k_adjoint_mu_scale");
}

int grav_stage1_vjp_f32(const void *posm, const void *mu4, void *rho64, void *gamma64,
                        int64_t n, double G, double eps, double dt, void *stream) {
  clear_error();
  GRAV_REQUIRE(posm != nullptr && mu4 != nullptr && rho64 != nullptr && gamma64 != nullptr,
               "grav_stage1_vjp_f32: NULL buffer");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_stage1_vjp_f32: n must be in [1, 2^29] (got %lld)", (long long)n);
  GRAV_REQUIRE(aligned(posm, 16) && aligned(mu4, 16) && aligned(rho64, 8) &&
                   aligned(gamma64, 8),
               "grav_stage1_vjp_f32: misaligned buffer");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "grav_stage1_vjp_f32: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_stage1_vjp_f32: eps must be >= 0");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "grav_stage1_vjp_f32: dt must be finite");
  const int nn = static_cast<int>(n);
  const float4 *p = static_cast<const float4 *>(posm);
  const float4 *u = static_cast<const float4 *>(mu4);
  double *rho = static_cast<double *>(rho64);
  double *gam = static_cast<double *>(gamma64);
  const float eps2 = static_cast<float>(eps * eps);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int K = n >= (1 << 16) ? 4 : (n >= (1 << 13) ? 2 : 1);
  const int S = n >= (1 << 13) ? 4 : 8;
  return stage1_vjp_launch(p, u, rho, gam, nn, dt * G, eps2, K, S, nullptr, 0, s);
}

int grav_stage1_vjp_f32_split(const void *posm, const void *mu4, void *rho64, void *gamma64,
                              int64_t n, double G, double eps, double dt, int i_per_lane,
                              int j_waves, int splits, void *workspace, int64_t workspace_bytes,
                              void *stream) {
  clear_error();
  GRAV_REQUIRE(posm != nullptr && mu4 != nullptr && rho64 != nullptr && gamma64 != nullptr,
               "grav_stage1_vjp_f32_split: NULL buffer");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_stage1_vjp_f32_split: n must be in [1, 2^29] (got %lld)", (long long)n);
  GRAV_REQUIRE(aligned(posm, 16) && aligned(mu4, 16) && aligned(rho64, 8) &&
                   aligned(gamma64, 8),
               "grav_stage1_vjp_f32_split: misaligned buffer");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300,
               "grav_stage1_vjp_f32_split: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_stage1_vjp_f32_split: eps must be >= 0");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300,
               "grav_stage1_vjp_f32_split: dt must be finite");
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) split_default_cfg(n, true, K, S, P);
  GRAV_REQUIRE(split_supported(K, S, P),
               "grav_stage1_vjp_f32_split: (i_per_lane, j_waves, splits) must be in {1, 2} x "
               "{1, 2, 4, 8} x [1, %d], or all 0 (got %d, %d, %d)", kMaxSplits, K, S, P);
  const int64_t need = split_workspace_bytes(n, P);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
               "grav_stage1_vjp_f32_split: workspace must be 256-byte aligned and hold %lld "
               "bytes", (long long)need);
  return stage1_vjp_split_launch(static_cast<const float4 *>(posm),
                                 static_cast<const float4 *>(mu4), static_cast<double *>(rho64),
                                 static_cast<double *>(gamma64),
                                 static_cast<double *>(workspace), static_cast<int>(n), dt * G,
                                 static_cast<float>(eps * eps), K, S, P,
                                 static_cast<hipStream_t>(stream));
}

int grav_stage1_vjp_f32_shard(const void *posm, const void *mu4, void *rho64, void *gamma64,
                              int64_t n, int64_t i0, int64_t m, double G, double eps, double dt,
                              int i_per_lane, int j_waves, int splits, void *workspace,
                              int64_t workspace_bytes, void *stream) {
  clear_error();
  const char *fn = "grav_stage1_vjp_f32_shard";
  GRAV_REQUIRE(posm != nullptr && mu4 != nullptr && rho64 != nullptr && gamma64 != nullptr,
               "%s: NULL buffer", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(i0 >= 0 && m >= 1 && i0 <= n - m,
               "%s: the shard [i0, i0 + m) must be a non-empty part of [0, n) (got i0 = %lld, "
               "m = %lld, n = %lld)", fn, (long long)i0, (long long)m, (long long)n);
  GRAV_REQUIRE(aligned(posm, 16) && aligned(mu4, 16) && aligned(rho64, 8) &&
                   aligned(gamma64, 8),
               "%s: misaligned buffer", fn);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "%s: eps must be >= 0", fn);
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "%s: dt must be finite", fn);
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) shard_vjp_default_cfg(n, m, K, S, P);
  if (P == 1) {
    GRAV_REQUIRE(((K == 1 || K == 2) && (S == 1 || S == 2 || S == 4 || S == 8)) ||
                     (K == 4 && (S == 1 || S == 2 || S == 4)),
                 "%s: with splits = 1, (i_per_lane, j_waves) must be in {1, 2} x {1, 2, 4, 8} "
                 "or {4} x {1, 2, 4} (got %d, %d)", fn, K, S);
  } else {
    GRAV_REQUIRE(split_supported(K, S, P),
                 "%s: (i_per_lane, j_waves, splits) must be in {1, 2} x {1, 2, 4, 8} x "
                 "[2, %d], splits 1 with the owner-computes shapes, or all 0 (got %d, %d, %d)",
                 fn, kMaxSplits, K, S, P);
    const int64_t need = split_workspace_bytes(m, P);
    GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
                 "%s: workspace must be 256-byte aligned and hold %lld bytes", fn,
                 (long long)need);
  }
  return stage1_vjp_shard_launch(static_cast<const float4 *>(posm),
                                 static_cast<const float4 *>(mu4), static_cast<double *>(rho64),
                                 static_cast<double *>(gamma64),
                                 static_cast<double *>(workspace), static_cast<int>(n),
                                 static_cast<int>(i0), static_cast<int>(m), dt * G,
                                 static_cast<float>(eps * eps), K, S, P,
                                 static_cast<hipStream_t>(stream));
}

int grav_stage1_vjp_shard_cfg(int64_t n, int64_t m, int *i_per_lane, int *j_waves,
                              int *splits) {
  clear_error();
  GRAV_REQUIRE(i_per_lane != nullptr && j_waves != nullptr && splits != nullptr,
               "grav_stage1_vjp_shard_cfg: NULL output");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies && m >= 1 && m <= n,
               "grav_stage1_vjp_shard_cfg: need 1 <= m <= n <= 2^29 (got m = %lld, n = %lld)",
               (long long)m, (long long)n);
  shard_vjp_default_cfg(n, m, *i_per_lane, *j_waves, *splits);
  return GRAV_OK;
}

int grav_stage1_vjp_f32_split_chain(const void *posm, void *mu4, void *rho64, void *nu64,
                                    void *gamma64, int64_t n, double G, double eps, double dt,
                                    int i_per_lane, int j_waves, int splits, int next_mu,
                                    void *workspace, int64_t workspace_bytes, void *eslot,
                                    void *stream) {
  clear_error();
  const char *fn = "grav_stage1_vjp_f32_split_chain";
  GRAV_REQUIRE(posm != nullptr && mu4 != nullptr && rho64 != nullptr && nu64 != nullptr &&
                   gamma64 != nullptr && eslot != nullptr,
               "%s: NULL buffer", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(aligned(posm, 16) && aligned(mu4, 16) && aligned(rho64, 8) && aligned(nu64, 8) &&
                   aligned(gamma64, 8) && aligned(eslot, 4),
               "%s: misaligned buffer", fn);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "%s: eps must be >= 0", fn);
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "%s: dt must be finite", fn);
  GRAV_REQUIRE(next_mu == 0 || next_mu == 1, "%s: next_mu must be 0 or 1 (got %d)", fn, next_mu);
  int K = i_per_lane, S = j_waves, P = splits;
  if (K == 0 && S == 0 && P == 0) split_default_cfg(n, true, K, S, P);
  GRAV_REQUIRE(split_supported(K, S, P),
               "%s: (i_per_lane, j_waves, splits) must be in {1, 2} x {1, 2, 4, 8} x [1, %d], "
               "or all 0 (got %d, %d, %d)", fn, kMaxSplits, K, S, P);
  const int64_t need = split_workspace_bytes(n, P);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
               "%s: workspace must be 256-byte aligned and hold %lld bytes", fn,
               (long long)need);
  return stage1_vjp_split_chain_launch(
      static_cast<const float4 *>(posm), static_cast<float4 *>(mu4),
      static_cast<double *>(rho64), static_cast<double *>(nu64), static_cast<double *>(gamma64),
      static_cast<double *>(workspace), static_cast<int>(n), dt * G, dt,
      static_cast<float>(eps * eps), K, S, P, next_mu, static_cast<unsigned *>(eslot),
      static_cast<hipStream_t>(stream));
}

int grav_backward_pair_split(const void *posm_adj, void *mu4, void *rho64, void *nu64,
                             void *gamma64, void *eslot, const void *posm_in, void *posm_out,
                             void *r64, void *v64, int64_t n, double G, double eps, double dt,
                             int vjp_i_per_lane, int vjp_splits, int step_i_per_lane,
                             int step_splits, int j_waves, void *workspace,
                             int64_t workspace_bytes, void *stream) {
  clear_error();
  const char *fn = "grav_backward_pair_split";
  GRAV_REQUIRE(posm_adj != nullptr && mu4 != nullptr && rho64 != nullptr && nu64 != nullptr &&
                   gamma64 != nullptr && eslot != nullptr && posm_in != nullptr &&
                   posm_out != nullptr && r64 != nullptr && v64 != nullptr,
               "%s: NULL buffer", fn);
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies, "%s: n must be in [1, 2^29] (got %lld)", fn,
               (long long)n);
  GRAV_REQUIRE(aligned(posm_adj, 16) && aligned(mu4, 16) && aligned(posm_in, 16) &&
                   aligned(posm_out, 16) && aligned(rho64, 8) && aligned(nu64, 8) &&
                   aligned(gamma64, 8) && aligned(r64, 8) && aligned(v64, 8) &&
                   aligned(eslot, 4),
               "%s: misaligned buffer", fn);
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "%s: G must be finite", fn);
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "%s: eps must be >= 0", fn);
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "%s: dt must be finite", fn);
  const char *a = static_cast<const char *>(posm_in);
  const char *b = static_cast<const char *>(posm_out);
  GRAV_REQUIRE(a == b || a + 16 * n <= b || b + 16 * n <= a,
               "%s: posm_out must be posm_in or not overlap it", fn);
  int KA = vjp_i_per_lane, PA = vjp_splits, KR = step_i_per_lane, PR = step_splits, S = j_waves;
  if (KA == 0 && PA == 0 && KR == 0 && PR == 0 && S == 0) {
    int SA, SR;
    split_default_cfg(n, true, KA, SA, PA);
    split_default_cfg(n, false, KR, SR, PR);
    S = SA == SR ? SA : 0;
  }
  GRAV_REQUIRE(backward_pair_supported(KA, KR, S) && PA >= 1 && PA <= kMaxSplits && PR >= 1 &&
                   PR <= kMaxSplits,
               "%s: (vjp_i_per_lane, vjp_splits, step_i_per_lane, step_splits, j_waves) must "
               "be in {1, 2} x [1, %d] x {1, 2} x [1, %d] x {4, 8}, or all 0 (got %d, %d, %d, "
               "%d, %d)", fn, kMaxSplits, kMaxSplits, KA, PA, KR, PR, S);
  const int64_t need = split_workspace_bytes(n, PA) + split_workspace_bytes(n, PR);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= need,
               "%s: workspace must be 256-byte aligned and hold %lld bytes", fn,
               (long long)need);
  double *part_a = static_cast<double *>(workspace);
  double *part_r = part_a + split_workspace_bytes(n, PA) / 8;
  return backward_pair_launch(
      static_cast<const float4 *>(posm_adj), static_cast<float4 *>(mu4),
      static_cast<double *>(rho64), static_cast<double *>(nu64), static_cast<double *>(gamma64),
      static_cast<unsigned *>(eslot), part_a, KA, PA, static_cast<const float4 *>(posm_in),
      static_cast<float4 *>(posm_out), static_cast<double *>(r64), static_cast<double *>(v64),
      part_r, KR, PR, S, eps == 0.0 ? 1 : 0, static_cast<int>(n),
      // the forward step scales its sums by G in FP32, as grav_step_split_f64m does; the
      // VJP by dt G in FP64, as grav_stage1_vjp_f32_split does
      static_cast<double>(static_cast<float>(G)), dt * G, dt, static_cast<float>(eps * eps),
      static_cast<hipStream_t>(stream));
}

int grav_backward_pair_workspace(int64_t n, int vjp_splits, int step_splits, int64_t *bytes) {
  clear_error();
  GRAV_REQUIRE(bytes != nullptr, "grav_backward_pair_workspace: bytes is NULL");
  GRAV_REQUIRE(n >= 1 && n <= kMaxBodies,
               "grav_backward_pair_workspace: n must be in [1, 2^29] (got %lld)", (long long)n);
  int K, S, PA = vjp_splits, PR = step_splits;
  if (PA == 0 && PR == 0) {
    split_default_cfg(n, true, K, S, PA);
    split_default_cfg(n, false, K, S, PR);
  }
  GRAV_REQUIRE(PA >= 1 && PA <= kMaxSplits && PR >= 1 && PR <= kMaxSplits,
               "grav_backward_pair_workspace: splits must be in [1, %d], or both 0 (got %d, %d)",
               kMaxSplits, PA, PR);
  *bytes = split_workspace_bytes(n, PA) + split_workspace_bytes(n, PR);
  return GRAV_OK;
}

}  // extern "C"
