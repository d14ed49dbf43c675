// Batched universes: thousands of independent universes of uneven size advanced for many
// steps in ONE launch.  Not present in the reference (whose kernels hold one universe,
// src/gravitation/kernel/_base_.py:64-93); the per-universe physics is the reference's stage 1
// (src/gravitation/kernel/py1.py:57-64, all-pairs form pc2.py:60-90, optional softening) and
// stage 2 (_base_.py:58-62) with FP64 position/velocity masters.
//
// Layout and schedule
//   * All universes' bodies are concatenated: universe u owns bodies [off[u], off[u+1]).
//     posm float4[n] (x, y, z, m; FP32 copy used by the pair terms), r64/v64 double[3*n] SoA.
//   * A plan (host-built once per set of sizes) bin-packs whole universes into G workgroup
//     "bins" by longest-processing-time-first on cost N_u^2; every bin's bodies fit in LDS.
//     Inside a bin, each universe is cut into tiles of 64*KT bodies and the tiles are
//     assigned to the bin's wavefronts (again LPT, cost N_u per tile), at most SMAX tiles per
//     wavefront.  A lane therefore owns at most KT*SMAX bodies for the whole launch.
//   * Per step, every wavefront computes the accelerations of its tiles against the j-bodies
//     of the tile's universe read from LDS (wave-uniform addresses: broadcast reads), folds
//     FP32 block partials with TwoSum, applies the semi-implicit Euler update to the tile's
//     FP64 masters (read from and written back to HBM by their owning lane, 96 bytes per body
//     per step, served by L2 / the Infinity Cache) right after the tile's pairs, and writes the
//     new float4 into the other LDS buffer.  One workgroup barrier per step.  Keeping the
//     masters of all of a lane's bodies in registers instead cost up to 200 VGPRs (2 waves per
//     SIMD) and made the schedule slower than the per-step one.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdio>
#include <new>
#include <numeric>
#include <vector>

#include "grav_common.h"
#include "pair_vjp.h"

namespace grav {

// One tile: a contiguous run of <= 64*KT bodies of one universe.
struct Tile {
  int gb;   // first body, global index
  int lb;   // first body, LDS index (within the bin)
  int ub;   // universe's first body, LDS index
  int ue;   // universe's end, LDS index
  int cnt;  // bodies in the tile
  int seg;  // the universe's index among the bin's segments
  int pad[2];
};

// A bin-load segment: copy n bodies from global g to LDS l.
struct Seg {
  int g, l, n, pad;
};

}  // namespace grav

struct grav_batched_plan {
  int device = 0;
  int64_t n_bodies = 0;
  int n_universes = 0;
  int groups = 0;
  int waves = 0;
  int kt = 0;
  int smax = 0;
  int max_bin_bodies = 0;
  int max_bin_segs = 0;             // universes in the fullest bin (fused adjoint LDS)
  int64_t max_bin_cost = 0;
  int64_t total_cost = 0;
  int max_universe = 0;
  // device arrays
  grav::Tile *tiles = nullptr;     // [groups * waves * smax]
  int *ntiles = nullptr;           // [groups * waves]
  grav::Seg *segs = nullptr;       // [total segments]
  int *seg_begin = nullptr;        // [groups + 1]
  int *bin_bodies = nullptr;       // [groups]
  int64_t *offsets = nullptr;      // [U + 1] (host copy kept below)
  int4 *items = nullptr;           // per-step path: {i0, i1, j0, j1} per i-tile, largest first
  int nitems = 0;
  bool fused_ok = false;           // the fused schedule exists (every universe fits a group)
  char fused_why[256] = "";
  std::vector<int64_t> h_offsets;
  // symmetric per-step stage 1 (grav_batched_sym_prepare): blocks of 64*sym_k*sym_w bodies
  int sym_k = 0, sym_w = 0;
  grav::SymBlock *sym_blocks = nullptr;
  int4 *sym_pairs = nullptr;  // {beta start, gamma start, beta end, gamma end}
  int4 *sym_diag = nullptr;   // {i0, i1, block start, block end}, tiles of 256
  int sym_nblocks = 0, sym_npairs = 0, sym_ndiag = 0;
};

namespace grav {
namespace {

constexpr int kLdsBytes = 160 * 1024;

__device__ __forceinline__ void two_sum_fold_b(float &hi, float &lo, float p) {
  const float s = hi + p;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (p - bp);
  hi = s;
  lo += e;
}

// Accelerations of KT bodies per lane (LDS indices li[k]) against j in [j0, j1) of `cur`.
template <int KT, bool MASK>
__device__ __forceinline__ void tile_pairs(const float4 *cur, int j0, int j1, const float (&xi)[KT],
                                           const float (&yi)[KT], const float (&zi)[KT],
                                           const int (&li)[KT], float eps2, int jblock,
                                           float (&hx)[KT], float (&hy)[KT], float (&hz)[KT],
                                           float (&lx)[KT], float (&ly)[KT], float (&lz)[KT]) {
  for (int jb = j0, je; jb < j1; jb = je) {
    je = (j1 - jb > jblock) ? jb + jblock : j1;
    float px[KT], py[KT], pz[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) px[k] = py[k] = pz[k] = 0.0f;
#pragma unroll 4
    for (int j = jb; j < je; ++j) {
      const float4 q = cur[j];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const float dx = q.x - xi[k];
        const float dy = q.y - yi[k];
        const float dz = q.z - zi[k];
        float r2 = __builtin_fmaf(dx, dx, eps2);
        r2 = __builtin_fmaf(dy, dy, r2);
        r2 = __builtin_fmaf(dz, dz, r2);
        float inv = __builtin_amdgcn_rsqf(r2);
        if (MASK) inv = (j == li[k]) ? 0.0f : inv;
        const float s = (q.w * inv) * (inv * inv);
        px[k] = __builtin_fmaf(dx, s, px[k]);
        py[k] = __builtin_fmaf(dy, s, py[k]);
        pz[k] = __builtin_fmaf(dz, s, pz[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      two_sum_fold_b(hx[k], lx[k], px[k]);
      two_sum_fold_b(hy[k], ly[k], py[k]);
      two_sum_fold_b(hz[k], lz[k], pz[k]);
    }
  }
}

template <int KT, int SMAX>
__global__ __launch_bounds__(512) void k_batched_fused(
    const Tile *__restrict__ tiles, const int *__restrict__ ntiles, const Seg *__restrict__ segs,
    const int *__restrict__ seg_begin, const int *__restrict__ bin_bodies,
    float4 *__restrict__ posm, double *__restrict__ r64, double *__restrict__ v64, int n,
    float G, float eps2, int mask_self, double dt, int steps, int jblock,
    float4 *__restrict__ traj) {
#pragma clang fp contract(off)
  extern __shared__ float4 lds[];
  const int nb = bin_bodies[blockIdx.x];
  float4 *buf0 = lds;
  float4 *buf1 = lds + nb;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int waves = blockDim.x / kWave;

  // Load the bin's universes into LDS buffer 0.
  for (int s = seg_begin[blockIdx.x]; s < seg_begin[blockIdx.x + 1]; ++s) {
    const Seg sg = segs[s];
    for (int t = threadIdx.x; t < sg.n; t += blockDim.x) buf0[sg.l + t] = posm[sg.g + t];
  }

  const int wid = blockIdx.x * waves + wave;
  const int nt = ntiles[wid];
  const Tile *mine = tiles + static_cast<size_t>(wid) * SMAX;
  __syncthreads();

  float4 *cur = buf0;
  float4 *nxt = buf1;
  for (int step = 0; step < steps; ++step) {
#pragma unroll 1
    for (int s = 0; s < nt; ++s) {
      // the tile (a wave-uniform record); lane slot k holds body gb + lane + 64 k when that
      // is below cnt, and otherwise aliases the tile's first LDS slot without writing
      const Tile tl = mine[s];
      int gi[KT], li[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int t = lane + k * kWave;
        const bool ok = t < tl.cnt;
        gi[k] = ok ? tl.gb + t : -1;
        li[k] = ok ? tl.lb + t : tl.lb;
      }
      if (traj != nullptr) {
        float4 *tr = traj + static_cast<size_t>(step) * n;
#pragma unroll
        for (int k = 0; k < KT; ++k)
          if (gi[k] >= 0) tr[gi[k]] = cur[li[k]];
      }
      // The FP64 masters of the tile's bodies: owned by this lane alone for the whole
      // launch, read from and written back to HBM (L2 / Infinity Cache resident) once per
      // step, issued here so that the pair loop hides the load.
      double rx[KT], ry[KT], rz[KT], vx[KT], vy[KT], vz[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int g = gi[k] >= 0 ? gi[k] : 0;
        rx[k] = r64[g];
        ry[k] = r64[n + g];
        rz[k] = r64[2 * n + g];
        vx[k] = v64[g];
        vy[k] = v64[n + g];
        vz[k] = v64[2 * n + g];
      }
      float xi[KT], yi[KT], zi[KT];
      float hx[KT], hy[KT], hz[KT], lx[KT], ly[KT], lz[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const float4 p = cur[li[k]];
        xi[k] = p.x;
        yi[k] = p.y;
        zi[k] = p.z;
        hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.0f;
      }
      const int ub = tl.ub, ue = tl.ue;
      if (mask_self) {
        const int a = tl.lb, b = tl.lb + tl.cnt;
        tile_pairs<KT, false>(cur, ub, a, xi, yi, zi, li, eps2, jblock, hx, hy, hz, lx, ly, lz);
        tile_pairs<KT, true>(cur, a, b, xi, yi, zi, li, eps2, jblock, hx, hy, hz, lx, ly, lz);
        tile_pairs<KT, false>(cur, b, ue, xi, yi, zi, li, eps2, jblock, hx, hy, hz, lx, ly, lz);
      } else {
        tile_pairs<KT, false>(cur, ub, ue, xi, yi, zi, li, eps2, jblock, hx, hy, hz, lx, ly, lz);
      }
      // Stage 2 of the tile right away: the new FP32 copy goes to `nxt`, which no wave reads
      // before the step's barrier.
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int g = gi[k];
        if (g >= 0) {
          const float ax = G * (hx[k] + lx[k]);
          const float ay = G * (hy[k] + ly[k]);
          const float az = G * (hz[k] + lz[k]);
          vx[k] = static_cast<double>(ax) * dt + vx[k];
          vy[k] = static_cast<double>(ay) * dt + vy[k];
          vz[k] = static_cast<double>(az) * dt + vz[k];
          rx[k] = vx[k] * dt + rx[k];
          ry[k] = vy[k] * dt + ry[k];
          rz[k] = vz[k] * dt + rz[k];
          r64[g] = rx[k];
          r64[n + g] = ry[k];
          r64[2 * n + g] = rz[k];
          v64[g] = vx[k];
          v64[n + g] = vy[k];
          v64[2 * n + g] = vz[k];
          const float m = cur[li[k]].w;
          nxt[li[k]] = make_float4(static_cast<float>(rx[k]), static_cast<float>(ry[k]),
                                   static_cast<float>(rz[k]), m);
        }
      }
    }
    __syncthreads();
    float4 *t = cur;
    cur = nxt;
    nxt = t;
  }

#pragma unroll 1
  for (int s = 0; s < nt; ++s) {
    const Tile tl = mine[s];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int t = lane + k * kWave;
      if (t < tl.cnt) posm[tl.gb + t] = cur[tl.lb + t];
    }
  }
}

// Backward over `steps` steps (t = steps-1 .. 0) of every universe of the bin, reading the
// positions each forward step used from traj[t].  Per step: trajectory positions -> LDS,
// owners form mu = nu + h rho (FP64, HBM) and take their universe's largest exponent
// (frexp of max |mu|) with an LDS atomic max; then they publish FP32 mu * 2^-E to LDS, and
// every tile adds 2^E h J_r^T mu to rho and 2^E h J_m^T mu to gamma (see pair_vjp.h).  The
// power-of-two scale keeps the FP32 pair terms in range however far a chaotic universe's
// adjoint grows, and is exact.  The exponents live in two LDS arrays used on alternate steps,
// so the array of the next step is cleared while this one is read.  Three barriers per step.
template <int KT, int SMAX>
__global__ __launch_bounds__(512) void k_batched_adjoint(
    const Tile *__restrict__ tiles, const int *__restrict__ ntiles, const Seg *__restrict__ segs,
    const int *__restrict__ seg_begin, const int *__restrict__ bin_bodies,
    const float4 *__restrict__ traj, double *__restrict__ rho, double *__restrict__ nu,
    double *__restrict__ gamma, int n, double hG, float eps2, int mask_self, double dt,
    int steps, int jblock) {
#pragma clang fp contract(off)
  extern __shared__ float4 lds[];
  const int nb = bin_bodies[blockIdx.x];
  const int sb = seg_begin[blockIdx.x];
  const int nseg = seg_begin[blockIdx.x + 1] - sb;
  float4 *pos = lds;
  float4 *mub = lds + nb;
  int *sexp = reinterpret_cast<int *>(lds + 2 * nb);  // [2][nseg] scale exponents
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int waves = blockDim.x / kWave;
  const int wid = blockIdx.x * waves + wave;
  const int nt = ntiles[wid];
  const Tile *mine = tiles + static_cast<size_t>(wid) * SMAX;
  // lane slot k of tile tl: body gb + lane + 64 k when below cnt (else gi = -1 and li the
  // tile's first LDS slot, never written)
  auto slots = [&](const Tile &tl, int (&gi)[KT], int (&li)[KT]) {
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int t = lane + k * kWave;
      const bool ok = t < tl.cnt;
      gi[k] = ok ? tl.gb + t : -1;
      li[k] = ok ? tl.lb + t : tl.lb;
    }
  };
  for (int t = threadIdx.x; t < 2 * nseg; t += blockDim.x) sexp[t] = INT_MIN;
  __syncthreads();
  for (int step = steps - 1; step >= 0; --step) {
    const float4 *tr = traj + static_cast<size_t>(step) * n;
    int *ecur = sexp + (step & 1) * nseg;
    int *enext = sexp + ((step & 1) ^ 1) * nseg;
    for (int s = sb; s < sb + nseg; ++s) {
      const Seg sg = segs[s];
      for (int t = threadIdx.x; t < sg.n; t += blockDim.x) pos[sg.l + t] = tr[sg.g + t];
    }
#pragma unroll 1
    for (int s = 0; s < nt; ++s) {
      const Tile tl = mine[s];
      int gi[KT], li[KT];
      slots(tl, gi, li);
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int g = gi[k];
        if (g >= 0) {
          double a = 0.0;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double m = nu[c * n + g] + dt * rho[c * n + g];
            nu[c * n + g] = m;
            a = fmax(a, fabs(m));  // NaN components drop out here
          }
          if (a > 0.0 && a <= DBL_MAX) {
            int x;
            frexp(a, &x);
            atomicMax(&ecur[tl.seg], x);
          }
        }
      }
    }
    __syncthreads();
    // Each owner reads back only the nu entries it wrote above.
#pragma unroll 1
    for (int s = 0; s < nt; ++s) {
      const Tile tl = mine[s];
      int gi[KT], li[KT];
      slots(tl, gi, li);
      const int e = ecur[tl.seg];
      const int E = (e == INT_MIN) ? 0 : e;
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int g = gi[k];
        if (g >= 0)
          mub[li[k]] = make_float4(static_cast<float>(ldexp(nu[g], -E)),
                                   static_cast<float>(ldexp(nu[n + g], -E)),
                                   static_cast<float>(ldexp(nu[2 * n + g], -E)), 0.0f);
      }
    }
    for (int t = threadIdx.x; t < nseg; t += blockDim.x) enext[t] = INT_MIN;
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < nt; ++s) {
      const Tile tl = mine[s];
      int gi[KT], li[KT];
      slots(tl, gi, li);
      float xk[KT], yk[KT], zk[KT], mk[KT], uxk[KT], uyk[KT], uzk[KT];
      float hi[4][KT], lo[4][KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const float4 p = pos[li[k]];
        const float4 u = mub[li[k]];
        xk[k] = p.x;
        yk[k] = p.y;
        zk[k] = p.z;
        mk[k] = p.w;
        uxk[k] = u.x;
        uyk[k] = u.y;
        uzk[k] = u.z;
#pragma unroll
        for (int c = 0; c < 4; ++c) hi[c][k] = lo[c][k] = 0.0f;
      }
      const int ub = tl.ub, ue = tl.ue;
      const int a = mask_self ? tl.lb : ue;
      const int b = mask_self ? tl.lb + tl.cnt : ue;
      // [ub, a) unmasked, [a, b) masked (own bodies), [b, ue) unmasked
      for (int part = 0; part < 3; ++part) {
        const int j0 = part == 0 ? ub : (part == 1 ? a : b);
        const int j1 = part == 0 ? a : (part == 1 ? b : ue);
        for (int jb = j0, je; jb < j1; jb = je) {
          je = (j1 - jb > jblock) ? jb + jblock : j1;
          float px[KT], py[KT], pz[KT], pg[KT];
#pragma unroll
          for (int k = 0; k < KT; ++k) px[k] = py[k] = pz[k] = pg[k] = 0.0f;
          if (part == 1) {
            for (int j = jb; j < je; ++j)
              vjp_pair<KT, true>(pos[j], mub[j], j, xk, yk, zk, mk, uxk, uyk, uzk, li, eps2, px,
                                 py, pz, pg);
          } else {
#pragma unroll 2
            for (int j = jb; j < je; ++j)
              vjp_pair<KT, false>(pos[j], mub[j], j, xk, yk, zk, mk, uxk, uyk, uzk, li, eps2,
                                  px, py, pz, pg);
          }
#pragma unroll
          for (int k = 0; k < KT; ++k) {
            two_sum_acc(hi[0][k], lo[0][k], px[k]);
            two_sum_acc(hi[1][k], lo[1][k], py[k]);
            two_sum_acc(hi[2][k], lo[2][k], pz[k]);
            two_sum_acc(hi[3][k], lo[3][k], pg[k]);
          }
        }
      }
      const int e = ecur[tl.seg];
      const double hs = ldexp(hG, e == INT_MIN ? 0 : e);
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int g = gi[k];
        if (g >= 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c)
            rho[c * n + g] += hs * (static_cast<double>(hi[c][k]) + static_cast<double>(lo[c][k]));
          gamma[g] -= hs * (static_cast<double>(hi[3][k]) + static_cast<double>(lo[3][k]));
        }
      }
    }
    __syncthreads();
  }
}

using AdjointFn = void (*)(const Tile *, const int *, const Seg *, const int *, const int *,
                           const float4 *, double *, double *, double *, int, double, float, int,
                           double, int, int);

AdjointFn pick_adjoint(int kt, int smax) {
#define GRAV_CASE(K, S) \
  if (kt == K && smax == S) return k_batched_adjoint<K, S>;
  GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
  GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4)
  GRAV_CASE(4, 1) GRAV_CASE(4, 2)
#undef GRAV_CASE
  return nullptr;
}

using FusedFn = void (*)(const Tile *, const int *, const Seg *, const int *, const int *,
                         float4 *, double *, double *, int, float, float, int, double, int, int,
                         float4 *);

FusedFn pick_fused(int kt, int smax) {
#define GRAV_CASE(K, S) \
  if (kt == K && smax == S) return k_batched_fused<K, S>;
  GRAV_CASE(1, 1) GRAV_CASE(1, 2) GRAV_CASE(1, 4) GRAV_CASE(1, 8)
  GRAV_CASE(2, 1) GRAV_CASE(2, 2) GRAV_CASE(2, 4)
  GRAV_CASE(4, 1) GRAV_CASE(4, 2)
#undef GRAV_CASE
  return nullptr;
}

void free_plan(grav_batched_plan *p) {
  if (!p) return;
  (void)hipFree(p->tiles);
  (void)hipFree(p->ntiles);
  (void)hipFree(p->segs);
  (void)hipFree(p->seg_begin);
  (void)hipFree(p->bin_bodies);
  (void)hipFree(p->offsets);
  (void)hipFree(p->items);
  (void)hipFree(p->sym_blocks);
  (void)hipFree(p->sym_pairs);
  (void)hipFree(p->sym_diag);
  delete p;
}

template <typename T>
int upload(T **dst, const std::vector<T> &src) {
  const size_t bytes = sizeof(T) * std::max<size_t>(1, src.size());
  GRAV_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(dst), bytes));
  if (!src.empty())
    GRAV_HIP_CHECK(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return GRAV_OK;
}

}  // namespace
}  // namespace grav

using namespace grav;

extern "C" {

int grav_batched_plan_create(const int64_t *offsets, int n_universes, int waves_per_group,
                             int i_per_lane, int tiles_per_wave, int groups,
                             grav_batched_plan **plan) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_plan_create: plan is NULL");
  *plan = nullptr;
  GRAV_REQUIRE(offsets != nullptr, "grav_batched_plan_create: offsets is NULL");
  GRAV_REQUIRE(n_universes >= 1, "grav_batched_plan_create: n_universes must be >= 1");
  GRAV_REQUIRE(waves_per_group >= 1 && waves_per_group <= 8,
               "grav_batched_plan_create: waves_per_group must be in [1, 8]");
  GRAV_REQUIRE(pick_fused(i_per_lane, tiles_per_wave) != nullptr,
               "grav_batched_plan_create: unsupported (i_per_lane, tiles_per_wave) = (%d, %d); "
               "i_per_lane in {1,2,4}, tiles_per_wave in {1,2,4,8}, product <= 8",
               i_per_lane, tiles_per_wave);
  GRAV_REQUIRE(groups >= 0, "grav_batched_plan_create: groups must be >= 0 (0 = automatic)");
  GRAV_REQUIRE(offsets[0] == 0, "grav_batched_plan_create: offsets[0] must be 0");
  const int U = n_universes;
  const int tile_len = kWave * i_per_lane;
  const int slots = waves_per_group * tiles_per_wave;
  const int cap_bodies = std::min(waves_per_group * tile_len * tiles_per_wave,
                                  kLdsBytes / (2 * static_cast<int>(sizeof(float4))));
  std::vector<int64_t> sizes(U);
  int64_t total = 0;
  int max_u = 0;
  for (int u = 0; u < U; ++u) {
    const int64_t s = offsets[u + 1] - offsets[u];
    GRAV_REQUIRE(s >= 1, "grav_batched_plan_create: universe %d is empty or offsets decrease",
                 u);
    sizes[u] = s;
    total += s;
    max_u = static_cast<int>(std::max<int64_t>(max_u, s));
  }
  GRAV_REQUIRE(total <= kMaxBodies, "grav_batched_plan_create: too many bodies");
  // The fused (whole universe per workgroup) schedule exists only if every universe fits one
  // workgroup; the per-step item list always does.
  bool fused_ok = max_u <= cap_bodies && (max_u + tile_len - 1) / tile_len <= slots;
  char why[256] = "";
  if (!fused_ok)
    snprintf(why, sizeof(why),
             "a universe of %d bodies does not fit one workgroup (capacity %d bodies, %d tiles "
             "of %d)", max_u, cap_bodies, slots, tile_len);
  int dev = 0;
  GRAV_HIP_CHECK(hipGetDevice(&dev));
  int cus = 0;
  GRAV_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (groups == 0) groups = 2 * cus;
  groups = std::max(1, std::min<int>(groups, U));

  // LPT over universes (cost N^2), respecting LDS capacity and tile slots per bin.
  std::vector<int> order(U);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return sizes[a] > sizes[b]; });
  std::vector<int64_t> bcost(groups, 0);
  std::vector<int> bbodies(groups, 0), bslots(groups, 0);
  std::vector<std::vector<int>> bunis(groups);
  for (int u : order) {
    if (!fused_ok) break;
    const int s = static_cast<int>(sizes[u]);
    const int nt = (s + tile_len - 1) / tile_len;
    int best = -1;
    for (int b = 0; b < groups; ++b) {
      if (bbodies[b] + s > cap_bodies || bslots[b] + nt > slots) continue;
      if (best < 0 || bcost[b] < bcost[best]) best = b;
    }
    if (best < 0) {
      fused_ok = false;
      snprintf(why, sizeof(why), "%d groups cannot hold the batch; raise groups", groups);
      break;
    }
    bunis[best].push_back(u);
    bcost[best] += static_cast<int64_t>(s) * s;
    bbodies[best] += s;
    bslots[best] += nt;
  }
  if (!fused_ok) {
    for (auto &v : bunis) v.clear();
    std::fill(bcost.begin(), bcost.end(), 0);
  }

  // Per bin: LDS layout, load segments, tiles LPT-assigned to waves (cost N_u per tile).
  std::vector<Tile> tiles(static_cast<size_t>(groups) * waves_per_group * tiles_per_wave);
  std::vector<int> ntiles(static_cast<size_t>(groups) * waves_per_group, 0);
  std::vector<Seg> segs;
  std::vector<int> seg_begin(groups + 1, 0), bin_bodies(groups, 0);
  int max_bin = 0;
  int64_t max_cost = 0, total_cost = 0;
  int max_segs = 0;
  for (int b = 0; b < groups; ++b) {
    seg_begin[b] = static_cast<int>(segs.size());
    int l = 0;
    struct T2 { Tile t; int64_t cost; };
    std::vector<T2> bt;
    for (int u : bunis[b]) {
      const int s = static_cast<int>(sizes[u]);
      const int g0 = static_cast<int>(offsets[u]);
      const int seg = static_cast<int>(segs.size()) - seg_begin[b];
      segs.push_back(Seg{g0, l, s, 0});
      for (int t0 = 0; t0 < s; t0 += tile_len) {
        Tile t{};
        t.gb = g0 + t0;
        t.lb = l + t0;
        t.ub = l;
        t.ue = l + s;
        t.cnt = std::min(tile_len, s - t0);
        t.seg = seg;
        bt.push_back(T2{t, s});
      }
      l += s;
    }
    bin_bodies[b] = std::max(l, 1);
    max_bin = std::max(max_bin, l);
    max_segs = std::max(max_segs, static_cast<int>(segs.size()) - seg_begin[b]);
    max_cost = std::max(max_cost, bcost[b]);
    std::stable_sort(bt.begin(), bt.end(), [](const T2 &a, const T2 &c) { return a.cost > c.cost; });
    std::vector<int64_t> wload(waves_per_group, 0);
    for (const T2 &x : bt) {
      int best = -1;
      for (int w = 0; w < waves_per_group; ++w) {
        if (ntiles[b * waves_per_group + w] >= tiles_per_wave) continue;
        if (best < 0 || wload[w] < wload[best]) best = w;
      }
      // Capacity was checked per bin, so a free slot exists.
      const int wid = b * waves_per_group + best;
      tiles[static_cast<size_t>(wid) * tiles_per_wave + ntiles[wid]] = x.t;
      ntiles[wid] += 1;
      wload[best] += x.cost;
    }
  }
  seg_begin[groups] = static_cast<int>(segs.size());

  // Per-step path: every universe's i-tiles of 64*i_per_lane bodies against the universe's
  // own bodies, costliest (largest universe) first so the long tiles start early.
  std::vector<int4> items;
  for (int u : order) {
    const int g0 = static_cast<int>(offsets[u]);
    const int s = static_cast<int>(sizes[u]);
    for (int t0 = 0; t0 < s; t0 += tile_len)
      items.push_back(make_int4(g0 + t0, g0 + std::min(s, t0 + tile_len), g0, g0 + s));
  }

  grav_batched_plan *p = new (std::nothrow) grav_batched_plan();
  GRAV_REQUIRE(p != nullptr, "grav_batched_plan_create: out of host memory");
  p->device = dev;
  p->n_bodies = total;
  p->n_universes = U;
  p->groups = groups;
  p->waves = waves_per_group;
  p->kt = i_per_lane;
  p->smax = tiles_per_wave;
  p->max_bin_bodies = max_bin;
  p->max_bin_segs = max_segs;
  p->max_bin_cost = max_cost;
  for (int u = 0; u < U; ++u) total_cost += sizes[u] * sizes[u];
  p->total_cost = total_cost;
  p->max_universe = max_u;
  p->fused_ok = fused_ok;
  snprintf(p->fused_why, sizeof(p->fused_why), "%s", why);
  p->h_offsets.assign(offsets, offsets + U + 1);
  int rc = upload(&p->tiles, tiles);
  if (rc == GRAV_OK) rc = upload(&p->ntiles, ntiles);
  if (rc == GRAV_OK) rc = upload(&p->segs, segs);
  if (rc == GRAV_OK) rc = upload(&p->seg_begin, seg_begin);
  if (rc == GRAV_OK) rc = upload(&p->bin_bodies, bin_bodies);
  if (rc == GRAV_OK) rc = upload(&p->offsets, p->h_offsets);
  if (rc == GRAV_OK) rc = upload(&p->items, items);
  p->nitems = static_cast<int>(items.size());
  if (rc != GRAV_OK) {
    free_plan(p);
    return rc;
  }
  *plan = p;
  return GRAV_OK;
}

int grav_batched_plan_destroy(grav_batched_plan *plan) {
  clear_error();
  free_plan(plan);
  return GRAV_OK;
}

int grav_batched_plan_info(const grav_batched_plan *plan, int64_t *info, int len) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr && info != nullptr, "grav_batched_plan_info: NULL argument");
  const int64_t v[] = {plan->n_bodies,      plan->n_universes, plan->groups,
                       plan->waves,         plan->kt,          plan->smax,
                       plan->max_bin_bodies, plan->max_bin_cost, plan->total_cost,
                       plan->max_universe,  plan->nitems,      plan->fused_ok ? 1 : 0};
  const int m = static_cast<int>(sizeof(v) / sizeof(v[0]));
  for (int k = 0; k < len && k < m; ++k) info[k] = v[k];
  return GRAV_OK;
}

// A plan's schedule arrays live on the device that was current at plan creation.
static int plan_on_current_device(const grav_batched_plan *plan, const char *fn) {
  int dev = 0;
  GRAV_HIP_CHECK(hipGetDevice(&dev));
  if (dev != plan->device)
    return fail(GRAV_EINVAL, "%s: the plan belongs to device %d, the current device is %d", fn,
                plan->device, dev);
  return GRAV_OK;
}

int grav_batched_step_f32(const grav_batched_plan *plan, void *posm, void *r64, void *v64,
                          double G, double eps, double dt, int steps, void *traj,
                          void *stream) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_step_f32: plan is NULL");
  if (int rc = plan_on_current_device(plan, "grav_batched_step_f32")) return rc;
  GRAV_REQUIRE(posm != nullptr && r64 != nullptr && v64 != nullptr,
               "grav_batched_step_f32: NULL state buffer");
  GRAV_REQUIRE(aligned(posm, 16) && aligned(r64, 8) && aligned(v64, 8),
               "grav_batched_step_f32: misaligned state buffer");
  GRAV_REQUIRE(traj == nullptr || aligned(traj, 16),
               "grav_batched_step_f32: traj must be 16-byte aligned");
  GRAV_REQUIRE(steps >= 0, "grav_batched_step_f32: steps must be >= 0");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "grav_batched_step_f32: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_batched_step_f32: eps must be >= 0");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300, "grav_batched_step_f32: dt must be finite");
  if (steps == 0) return GRAV_OK;
  if (!plan->fused_ok)
    return fail(GRAV_ENOTSUP, "grav_batched_step_f32: no fused schedule: %s", plan->fused_why);
  FusedFn fn = pick_fused(plan->kt, plan->smax);
  const size_t lds = 2 * sizeof(float4) * std::max(1, plan->max_bin_bodies);
  hipLaunchKernelGGL(fn, dim3(plan->groups), dim3(kWave * plan->waves), lds,
                     static_cast<hipStream_t>(stream), plan->tiles, plan->ntiles, plan->segs,
                     plan->seg_begin, plan->bin_bodies, static_cast<float4 *>(posm),
                     static_cast<double *>(r64), static_cast<double *>(v64),
                     static_cast<int>(plan->n_bodies), static_cast<float>(G),
                     static_cast<float>(eps * eps), eps == 0.0 ? 1 : 0, dt, steps, 256,
                     static_cast<float4 *>(traj));
  return check_launch("k_batched_fused");
}

int grav_batched_stage1_f32(const grav_batched_plan *plan, const void *posm, void *acc,
                            double G, double eps, int j_waves, void *stream) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_stage1_f32: plan is NULL");
  if (int rc = plan_on_current_device(plan, "grav_batched_stage1_f32")) return rc;
  GRAV_REQUIRE(posm != nullptr && acc != nullptr, "grav_batched_stage1_f32: NULL buffer");
  GRAV_REQUIRE(aligned(posm, 16) && aligned(acc, 4), "grav_batched_stage1_f32: misaligned");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "grav_batched_stage1_f32: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_batched_stage1_f32: eps must be >= 0");
  GRAV_REQUIRE(j_waves == 1 || j_waves == 2 || j_waves == 4 || j_waves == 8,
               "grav_batched_stage1_f32: j_waves must be 1, 2, 4 or 8");
  return stage1_f32_launch(static_cast<const float4 *>(posm), static_cast<float *>(acc),
                           static_cast<int>(plan->n_bodies), static_cast<float>(G),
                           static_cast<float>(eps * eps), eps == 0.0 ? 1 : 0, plan->kt, j_waves,
                           256, plan->items, plan->nitems, static_cast<hipStream_t>(stream));
}

int grav_batched_sym_prepare(grav_batched_plan *plan, int i_chunks_per_wave, int waves) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_sym_prepare: plan is NULL");
  if (int rc = plan_on_current_device(plan, "grav_batched_sym_prepare")) return rc;
  GRAV_REQUIRE(sym_batched_supported(i_chunks_per_wave, waves),
               "grav_batched_sym_prepare: (i_chunks_per_wave, waves) must be in {4,8} x {1,2,4}");
  const int ib = kWave * i_chunks_per_wave * waves;
  const int U = plan->n_universes;
  const std::vector<int64_t> &off = plan->h_offsets;
  // Largest universes first, so their long item runs start early; each universe's block
  // pairs stay contiguous and row-major from its first item.
  std::vector<int> order(U);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return off[a + 1] - off[a] > off[b + 1] - off[b];
  });
  std::vector<SymBlock> blocks;
  std::vector<int4> pairs, diag;
  int64_t npairs = 0;
  for (int u : order) {
    const int s0 = static_cast<int>(off[u]), e0 = static_cast<int>(off[u + 1]);
    const int nb = (e0 - s0 + ib - 1) / ib;
    GRAV_REQUIRE(npairs + int64_t(nb) * (nb - 1) / 2 <= 0x7fffffff,
                 "grav_batched_sym_prepare: too many block pairs; use larger blocks");
    const int base = static_cast<int>(npairs);
    for (int b = 0; b < nb; ++b) {
      const int bs = s0 + b * ib, be = std::min(e0, bs + ib);
      blocks.push_back(SymBlock{bs, be, base, nb, b, {0, 0, 0}});
      for (int t0 = bs; t0 < be; t0 += 256)
        diag.push_back(make_int4(t0, std::min(be, t0 + 256), bs, be));
    }
    for (int b = 0; b < nb; ++b)
      for (int g = b + 1; g < nb; ++g)
        pairs.push_back(make_int4(s0 + b * ib, s0 + g * ib, std::min(e0, s0 + (b + 1) * ib),
                                  std::min(e0, s0 + (g + 1) * ib)));
    npairs += int64_t(nb) * (nb - 1) / 2;
  }
  (void)hipFree(plan->sym_blocks);
  (void)hipFree(plan->sym_pairs);
  (void)hipFree(plan->sym_diag);
  plan->sym_blocks = nullptr;
  plan->sym_pairs = nullptr;
  plan->sym_diag = nullptr;
  plan->sym_k = plan->sym_w = 0;
  int rc = upload(&plan->sym_blocks, blocks);
  if (rc == GRAV_OK) rc = upload(&plan->sym_pairs, pairs);
  if (rc == GRAV_OK) rc = upload(&plan->sym_diag, diag);
  if (rc != GRAV_OK) return rc;
  plan->sym_nblocks = static_cast<int>(blocks.size());
  plan->sym_npairs = static_cast<int>(pairs.size());
  plan->sym_ndiag = static_cast<int>(diag.size());
  plan->sym_k = i_chunks_per_wave;
  plan->sym_w = waves;
  return GRAV_OK;
}

namespace {
struct SymBatchedLayout {
  int64_t diag_bytes, part_bytes, total;
};
SymBatchedLayout sym_batched_layout(const grav_batched_plan *p) {
  auto up = [](int64_t x) { return (x + 255) & ~int64_t(255); };
  SymBatchedLayout L{};
  const int64_t ib = int64_t(kWave) * p->sym_k * p->sym_w;
  L.diag_bytes = up(3 * p->n_bodies * 4);
  L.part_bytes = up(int64_t(p->sym_npairs) * 3 * ib * 4);
  L.total = L.diag_bytes + 2 * L.part_bytes;
  return L;
}
}  // namespace

int grav_batched_sym_workspace(const grav_batched_plan *plan, int64_t *bytes) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr && bytes != nullptr, "grav_batched_sym_workspace: NULL argument");
  GRAV_REQUIRE(plan->sym_k != 0, "grav_batched_sym_workspace: call grav_batched_sym_prepare first");
  *bytes = sym_batched_layout(plan).total;
  return GRAV_OK;
}

int grav_batched_stage1_sym_f32(const grav_batched_plan *plan, const void *posm, void *acc,
                                double G, double eps, void *workspace, int64_t workspace_bytes,
                                void *stream) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_stage1_sym_f32: plan is NULL");
  if (int rc = plan_on_current_device(plan, "grav_batched_stage1_sym_f32")) return rc;
  GRAV_REQUIRE(plan->sym_k != 0,
               "grav_batched_stage1_sym_f32: call grav_batched_sym_prepare first");
  GRAV_REQUIRE(posm != nullptr && acc != nullptr, "grav_batched_stage1_sym_f32: NULL buffer");
  GRAV_REQUIRE(aligned(posm, 16) && aligned(acc, 4), "grav_batched_stage1_sym_f32: misaligned");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300,
               "grav_batched_stage1_sym_f32: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_batched_stage1_sym_f32: eps must be >= 0");
  const SymBatchedLayout L = sym_batched_layout(plan);
  GRAV_REQUIRE(workspace != nullptr && aligned(workspace, 256) && workspace_bytes >= L.total,
               "grav_batched_stage1_sym_f32: workspace must be 256-byte aligned and hold %lld "
               "bytes", (long long)L.total);
  char *ws = static_cast<char *>(workspace);
  float *diag = reinterpret_cast<float *>(ws);
  float *pi = reinterpret_cast<float *>(ws + L.diag_bytes);
  float *pj = reinterpret_cast<float *>(ws + L.diag_bytes + L.part_bytes);
  return sym_batched_launch(static_cast<const float4 *>(posm), static_cast<float *>(acc),
                            static_cast<int>(plan->n_bodies), static_cast<float>(G),
                            static_cast<float>(eps * eps), eps == 0.0 ? 1 : 0, plan->sym_k,
                            plan->sym_w, plan->sym_pairs, plan->sym_npairs, plan->sym_diag,
                            plan->sym_ndiag, plan->sym_blocks, plan->sym_nblocks, diag, pi, pj,
                            static_cast<hipStream_t>(stream));
}

int grav_batched_adjoint_mu(const grav_batched_plan *plan, const void *rho64, void *nu64,
                            void *mu4, double dt, void *stream) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_adjoint_mu: plan is NULL");
  if (int rc = plan_on_current_device(plan, "grav_batched_adjoint_mu")) return rc;
  GRAV_REQUIRE(rho64 != nullptr && nu64 != nullptr && mu4 != nullptr,
               "grav_batched_adjoint_mu: NULL buffer");
  GRAV_REQUIRE(aligned(rho64, 8) && aligned(nu64, 8) && aligned(mu4, 16),
               "grav_batched_adjoint_mu: misaligned buffer");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300,
               "grav_batched_adjoint_mu: dt must be finite");
  return adjoint_mu_launch(static_cast<const double *>(rho64), static_cast<double *>(nu64),
                           static_cast<float4 *>(mu4), static_cast<int>(plan->n_bodies), dt,
                           plan->offsets, plan->n_universes, static_cast<hipStream_t>(stream));
}

int grav_batched_stage1_vjp_f32(const grav_batched_plan *plan, const void *posm,
                                const void *mu4, void *rho64, void *gamma64, double G,
                                double eps, double dt, int j_waves, void *stream) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_stage1_vjp_f32: plan is NULL");
  if (int rc = plan_on_current_device(plan, "grav_batched_stage1_vjp_f32")) return rc;
  GRAV_REQUIRE(posm != nullptr && mu4 != nullptr && rho64 != nullptr && gamma64 != nullptr,
               "grav_batched_stage1_vjp_f32: NULL buffer");
  GRAV_REQUIRE(aligned(posm, 16) && aligned(mu4, 16) && aligned(rho64, 8) &&
                   aligned(gamma64, 8),
               "grav_batched_stage1_vjp_f32: misaligned buffer");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300,
               "grav_batched_stage1_vjp_f32: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_batched_stage1_vjp_f32: eps must be >= 0");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300,
               "grav_batched_stage1_vjp_f32: dt must be finite");
  GRAV_REQUIRE(j_waves == 1 || j_waves == 2 || j_waves == 4 || j_waves == 8,
               "grav_batched_stage1_vjp_f32: j_waves must be 1, 2, 4 or 8");
  GRAV_REQUIRE(plan->kt <= 4, "grav_batched_stage1_vjp_f32: plan i_per_lane must be <= 4");
  return stage1_vjp_launch(static_cast<const float4 *>(posm), static_cast<const float4 *>(mu4),
                           static_cast<double *>(rho64), static_cast<double *>(gamma64),
                           static_cast<int>(plan->n_bodies), dt * G,
                           static_cast<float>(eps * eps), plan->kt, j_waves, plan->items,
                           plan->nitems, static_cast<hipStream_t>(stream));
}

int grav_batched_adjoint_f32(const grav_batched_plan *plan, const void *traj, void *rho64,
                             void *nu64, void *gamma64, double G, double eps, double dt,
                             int steps, void *stream) {
  clear_error();
  GRAV_REQUIRE(plan != nullptr, "grav_batched_adjoint_f32: plan is NULL");
  if (int rc = plan_on_current_device(plan, "grav_batched_adjoint_f32")) return rc;
  GRAV_REQUIRE(traj != nullptr && rho64 != nullptr && nu64 != nullptr && gamma64 != nullptr,
               "grav_batched_adjoint_f32: NULL buffer");
  GRAV_REQUIRE(aligned(traj, 16) && aligned(rho64, 8) && aligned(nu64, 8) &&
                   aligned(gamma64, 8),
               "grav_batched_adjoint_f32: misaligned buffer");
  GRAV_REQUIRE(steps >= 0, "grav_batched_adjoint_f32: steps must be >= 0");
  GRAV_REQUIRE(G == G && G > -1e300 && G < 1e300, "grav_batched_adjoint_f32: G must be finite");
  GRAV_REQUIRE(eps >= 0.0 && eps < 1e150, "grav_batched_adjoint_f32: eps must be >= 0");
  GRAV_REQUIRE(dt == dt && dt > -1e300 && dt < 1e300,
               "grav_batched_adjoint_f32: dt must be finite");
  if (steps == 0) return GRAV_OK;
  if (!plan->fused_ok)
    return fail(GRAV_ENOTSUP, "grav_batched_adjoint_f32: no fused schedule: %s",
                plan->fused_why);
  AdjointFn fn = pick_adjoint(plan->kt, plan->smax);
  const size_t lds = 2 * sizeof(float4) * std::max(1, plan->max_bin_bodies) +
                     2 * sizeof(int) * std::max(1, plan->max_bin_segs);
  hipLaunchKernelGGL(fn, dim3(plan->groups), dim3(kWave * plan->waves), lds,
                     static_cast<hipStream_t>(stream), plan->tiles, plan->ntiles, plan->segs,
                     plan->seg_begin, plan->bin_bodies, static_cast<const float4 *>(traj),
                     static_cast<double *>(rho64), static_cast<double *>(nu64),
                     static_cast<double *>(gamma64), static_cast<int>(plan->n_bodies), dt * G,
                     static_cast<float>(eps * eps), /*mask_self=*/1, dt, steps, 256);
  return check_launch("k_batched_adjoint");
}

}  // extern "C"
