/*
 * gravitation_hip.h -- C ABI of libgravhip.so, the MI355X (gfx950) direct-summation
 * N-body kernel library.
 *
 * Every entry point:
 *   - takes plain device pointers, sizes and scalars (no C++ or torch types);
 *   - validates its arguments before any HIP call and returns GRAV_OK (0) or a
 *     GRAV_E* code; the message of the last failure on the calling thread is
 *     returned by grav_last_error();
 *   - enqueues its work on the caller's HIP stream (`stream`, a hipStream_t; NULL is
 *     the null stream) and returns without synchronising.
 *
 * The reference crosses into native code with `void step_stage1(struct univ *self)`
 * (src/gravitation/kernel/_lib1_/lib.c:37-52, bound by ctypes in
 * src/gravitation/kernel/c1a.py:65-76): a struct of host SoA pointers, G and N, with no
 * return code and no error channel.  This library keeps that spirit (one flat call per
 * stage, memory owned by the caller, bound by ctypes) but passes device pointers as
 * arguments and reports errors.
 *
 * Layouts (all device memory, 16-byte aligned where a float4/double4 is named):
 *   posm   float4[n]  (x, y, z, m) per body           -- FP32 stage-1 input
 *   posm64 double4[n] (x, y, z, m) per body           -- FP64 stage-1 input / FP64 master
 *   acc    float [3*n]  SoA: ax[0..n), ay[0..n), az[0..n)
 *   acc64  double[3*n]  SoA
 *   vel    float [3*n]  SoA;  r64, v64: double[3*n] SoA
 *
 * Physics (src/gravitation/kernel/py1.py:57-64, src/gravitation/kernel/pc2.py:60-90):
 *   a_i = G * sum_{j != i} m_j (r_j - r_i) / (|r_j - r_i|^2 + eps^2)^(3/2)
 * eps = 0 is the reference formula (self term excluded); eps > 0 is Plummer softening.
 */
#ifndef GRAVITATION_HIP_H
#define GRAVITATION_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRAV_OK 0
#define GRAV_EINVAL 1     /* invalid argument (checked before any HIP call) */
#define GRAV_EHIP 2       /* a HIP runtime call failed */
#define GRAV_ENOTSUP 3    /* configuration not supported */

#define GRAV_ACCUM_PLAIN 0   /* one FP32 accumulator per component */
#define GRAV_ACCUM_KAHAN 1   /* FP32 block partials folded into an FP32 (hi, lo) TwoSum pair */

/* Library version (major * 10000 + minor * 100 + patch). */
int grav_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *grav_last_error(void);

/* Number of visible HIP devices (0 if none); returns GRAV_OK even when 0. */
int grav_device_count(int *count);

/*
 * Stage 1, FP32 pair terms: acc[3*n] <- G * sum_j m_j d_ij (d^2 + eps^2)^-3/2.
 * Replaces `step_stage1(struct univ*)` (src/gravitation/kernel/_lib1_/lib.c:52-126) and the
 * pc2 `update_all` launch (src/gravitation/kernel/pc2.py:147-162).
 * accum: GRAV_ACCUM_PLAIN or GRAV_ACCUM_KAHAN.  n in [1, 2^29] (every entry point).
 */
int grav_stage1_f32(const void *posm, void *acc, int64_t n, double G, double eps,
                    int accum, void *stream);

/*
 * As grav_stage1_f32 with an explicit launch configuration (used by the autotuner):
 *   i_per_lane in {1, 2, 4, 8}: i-bodies held in registers per lane;
 *   j_waves    in {1, 2, 4, 8}, or 16 with i_per_lane 1 or 2: wavefronts of one workgroup
 *              splitting the j-range;
 *   j_block    >= 8 (multiple of 8): j-bodies per compensation block.
 */
int grav_stage1_f32_cfg(const void *posm, void *acc, int64_t n, double G, double eps,
                        int accum, int i_per_lane, int j_waves, int j_block, void *stream);

/*
 * Stage 1, FP32, symmetric form: each unordered pair evaluated once and applied to both
 * bodies (Newton's third law, as src/gravitation/kernel/_lib1_/lib.c:72-123 and
 * src/gravitation/kernel/np1.py:81-108 do on the CPU), same result contract as
 * grav_stage1_f32 (acc[3*n], FP32 pair terms, compensated accumulation, deterministic).
 * Bodies are grouped in blocks of 64*i_chunks_per_wave*waves; (i_chunks_per_wave, waves) in
 * {4,8} x {2,4}, or (0, 0) for the library's choice by n ((4, 2) below 57344 bodies, (8, 2)
 * from there; the form pays off over grav_stage1_f32 from about 24576 bodies and over
 * grav_stage1_f32_split from about 36000 on an MI355X, below which its few block pairs cannot
 * fill the GPU).  A row of block pairs is cut into segments of up to 8 pairs, one workgroup
 * each, whose partial sums go to a device workspace (256-byte aligned).
 * grav_stage1_sym_workspace gives the size that holds every segment at once (12 n bytes +
 * 12 block (1 + seglen) bytes per segment: about 12 n + 6 n^2 / block (1 + seglen) / seglen,
 * 510 MB at 2^18 bodies); any size down to grav_stage1_sym_workspace_min's (one row of block
 * pairs, plus 24 n bytes of FP64 sums: about 50 n bytes) also works, the rows then being done
 * in slabs that fit, each folded into the FP64 sums in a fixed order.  The result depends on
 * the workspace size only through that rounding and is deterministic for a given size.
 */
int grav_stage1_sym_workspace(int64_t n, int i_chunks_per_wave, int waves, int64_t *bytes);
int grav_stage1_sym_workspace_min(int64_t n, int i_chunks_per_wave, int waves, int64_t *bytes);
int grav_stage1_sym_f32(const void *posm, void *acc, int64_t n, double G, double eps,
                        int i_chunks_per_wave, int waves, void *workspace,
                        int64_t workspace_bytes, void *stream);
/*
 * The symmetric form divided among GPUs by rows of block pairs, as the reference's np4
 * divides the rows of its pair loop among worker processes, each holding every position, and
 * adds the workers' arrays on the host (src/gravitation/kernel/np4.py:101-118, :135-163).
 * Share `share` of `shares` (0 <= share < shares <= 65536) holds about 1/shares of the block
 * pairs: rows beta and nb - 2 - beta (nb blocks) together hold nb pairs, and the share takes a
 * contiguous run of these couples, cut as np4 cuts its row batches, so that shares are equal
 * when the couples divide evenly; and the block-diagonal pairs of blocks
 * [share nb / shares, (share + 1) nb / shares).  grav_stage1_sym_f32_share writes that
 * share's contribution to every body as unscaled FP64 sums, sums64 double[3][n] (zero for
 * bodies it does not touch), with a workspace of at least bytes[1] bytes (bytes[0] holds all
 * its segments at once; below that it works in slabs of rows, deterministic for a given size).
 * grav_stage1_sym_share_plan gives rows = {first row, end row, first and end row of the
 * partner rows, first and end diagonal block} (int[6]) and bytes = {full, minimum} for a
 * share.  grav_stage1_sym_shares_combine
 * adds the shares of `count` bodies in share order in FP64, acc[c][i] = G * sum_k
 * sums64[k][c][i] (sums64 double[shares][3][count]): after an exchange that gives each GPU
 * its own bodies' sums from every share, the same call yields its rows of the accelerations,
 * deterministically.  The combined result equals grav_stage1_sym_f32's to FP64 rounding
 * (its fold adds the same terms in another order).
 */
int grav_stage1_sym_share_plan(int64_t n, int share, int shares, int i_chunks_per_wave,
                               int waves, int *rows, int64_t *bytes);
int grav_stage1_sym_f32_share(const void *posm, void *sums64, int64_t n, double eps, int share,
                              int shares, int i_chunks_per_wave, int waves, void *workspace,
                              int64_t workspace_bytes, void *stream);
int grav_stage1_sym_shares_combine(const void *sums64, int shares, int64_t count, double G,
                                   void *acc, void *stream);

/* Stage 1, FP64 pair terms and accumulation: acc64[3*n] from posm64 double4[n]. */
int grav_stage1_f64(const void *posm64, void *acc64, int64_t n, double G, double eps,
                    void *stream);

/*
 * Stage 1, FP64, symmetric form: grav_stage1_f64's result contract with every unordered pair
 * evaluated once (the scheme of grav_stage1_sym_f32, sums in FP64).  (i_chunks_per_wave,
 * waves) in {2,4} x {2,4}, or (0, 0) for the library's choice by n.  Workspace (256-byte
 * aligned) as for grav_stage1_sym_f32 with 8-byte partials: one slab's size from
 * grav_stage1_sym_f64_workspace, the smallest from grav_stage1_sym_f64_workspace_min.
 */
int grav_stage1_sym_f64_workspace(int64_t n, int i_chunks_per_wave, int waves, int64_t *bytes);
int grav_stage1_sym_f64_workspace_min(int64_t n, int i_chunks_per_wave, int waves,
                                      int64_t *bytes);
int grav_stage1_sym_f64(const void *posm64, void *acc64, int64_t n, double G, double eps,
                        int i_chunks_per_wave, int waves, void *workspace,
                        int64_t workspace_bytes, void *stream);

/*
 * Stage 2, semi-implicit Euler (src/gravitation/kernel/_base_.py:58-62,
 * src/gravitation/kernel/np2.py:110-115):  v <- v + dt*a;  r <- r + dt*v (new v).
 * FP32 state: positions in posm (m untouched), velocities vel[3*n].
 */
int grav_stage2_f32(void *posm, void *vel, const void *acc, int64_t n, double dt,
                    void *stream);

/*
 * Stage 2 with FP64 masters: r64, v64 (SoA double) advanced with FP32 acc, then the
 * FP32 stage-1 copy posm.xyz <- (float) r64 (m untouched).
 */
int grav_stage2_f64m(void *posm, void *r64, void *v64, const void *acc, int64_t n,
                     double dt, void *stream);

/* Stage 2, all-FP64: posm64.xyz and v64 advanced with acc64. */
int grav_stage2_f64(void *posm64, void *v64, const void *acc64, int64_t n, double dt,
                    void *stream);
/* The three stage-2 calls with the new state also stored straight into a host mirror, so
 * that the host views of the front end (`_point_mass._r`, `._v`, src/gravitation/kernel/
 * _base_.py:39-62, refreshed by the reference's kernels after every step) are current once the
 * stream has drained, with no copy command.  `mirror` is the device address of a buffer from
 * grav_host_mirror_alloc holding, in the call's precision: f32 and f64 posm (n, 4) then the
 * velocities (3, n), 28 or 56 bytes per body; f64m r64 (3, n) then v64 (3, n), 48 bytes per
 * body.  Alignment 16 (f32), 8 (f64m) or 32 (f64) bytes. */
int grav_stage2_f32_mirror(void *posm, void *vel, const void *acc, int64_t n, double dt,
                           void *mirror, void *stream);
int grav_stage2_f64m_mirror(void *posm, void *r64, void *v64, const void *acc, int64_t n,
                            double dt, void *mirror, void *stream);
int grav_stage2_f64_mirror(void *posm64, void *v64, const void *acc64, int64_t n, double dt,
                           void *mirror, void *stream);
/* Pinned, mapped and coherent host memory of `bytes` bytes for a mirror (*host: its host
 * address, *device: the address kernels write it through), and its release. */
int grav_host_mirror_alloc(int64_t bytes, void **host, void **device);
int grav_host_mirror_free(void *host);

/*
 * Batched universes (no counterpart in the reference, whose kernels hold one universe:
 * src/gravitation/kernel/_base_.py:64-93).  Universe u owns bodies
 * [offsets[u], offsets[u+1]) of the concatenated arrays posm float4[n], r64/v64 double[3*n].
 *
 * A plan bin-packs whole universes into `groups` workgroups (0 = twice the CU count) of
 * `waves_per_group` (1..8) wavefronts; each universe is cut into tiles of 64*i_per_lane
 * bodies, at most `tiles_per_wave` per wavefront (i_per_lane in {1,2,4}, tiles_per_wave in
 * {1,2,4,8}, product <= 8).  Every universe must fit one workgroup (LDS: <= 5120 bodies per
 * workgroup).  offsets is a HOST array of n_universes+1 entries starting at 0.  The plan
 * owns device memory on the device current at creation.
 */
typedef struct grav_batched_plan grav_batched_plan;

int grav_batched_plan_create(const int64_t *offsets, int n_universes, int waves_per_group,
                             int i_per_lane, int tiles_per_wave, int groups,
                             grav_batched_plan **plan);
int grav_batched_plan_destroy(grav_batched_plan *plan);

/* info[0..len): n_bodies, n_universes, groups, waves, i_per_lane, tiles_per_wave,
 * max bodies per workgroup, max workgroup cost (sum N^2), total cost, largest universe. */
int grav_batched_plan_info(const grav_batched_plan *plan, int64_t *info, int len);

/*
 * Advances every universe `steps` steps in ONE launch: per step, FP32 compensated stage 1
 * (the universe's own bodies only) and stage 2 on FP64 masters r64, v64; posm.xyz is the
 * FP32 copy.  traj (NULL or float4[steps][n]) receives posm at the start of every step.
 * Equivalent to `steps` x (grav_stage1_f32 + grav_stage2_f64m) on each universe.
 */
int grav_batched_step_f32(const grav_batched_plan *plan, void *posm, void *r64, void *v64,
                          double G, double eps, double dt, int steps, void *traj,
                          void *stream);

/*
 * Adjoint (reverse mode) of the step v' = v + dt a(r, m), r' = r + dt v' (no counterpart in
 * the reference).  State: rho64 = dL/dr, nu64 = dL/dv (double[3*n] SoA), gamma64 = dL/dm
 * (double[n]).  One backward step, with mu = nu + dt*rho:
 *   rho <- rho + dt J_r^T mu,  gamma <- gamma + dt J_m^T mu,  nu <- mu
 * where J = da/d(r, m) at the positions the forward step used.
 *
 * grav_adjoint_mu: the element-wise part for ONE universe of n bodies: mu = nu + dt*rho,
 *   nu <- mu, and mu4 float4[n] <- (mu * 2^-E, E), E (an int, stored bit for bit in the w
 *   float) = frexp exponent of the largest finite |mu| component (0 if none), so that
 *   |mu4.xyz| < 1.  The VJP is linear in mu; the scale keeps its FP32 pair terms, which carry
 *   factors up to eps^-3 m N, from overflowing when the adjoint of a chaotic trajectory grows
 *   by 1e28 and more, and is undone exactly in the FP64 update.
 * grav_stage1_vjp_f32: the all-pairs part at positions posm (float4 x,y,z,m) with mu4 from
 *   grav_adjoint_mu: rho += 2^E dt*G*sum_j T_kj (m_k mu_j - m_j mu_k),
 *   gamma -= 2^E dt*G*sum_j s_kj (r_j - r_k).mu_j, T = s (I - 3 d d^T/q), s = q^-3/2,
 *   q = |r_j - r_k|^2 + eps^2 (FP32 pair terms, FP64 state); E is read from mu4[k].w (a zero
 *   w means unscaled).
 * grav_batched_adjoint_f32: `steps` backward steps of every universe of a plan in ONE launch,
 *   reading the forward positions from traj float4[steps][n] (as written by
 *   grav_batched_step_f32), from step steps-1 down to 0, with the per-universe scale of
 *   grav_adjoint_mu applied in LDS.
 */
int grav_adjoint_mu(const void *rho64, void *nu64, void *mu4, int64_t n, double dt,
                    void *stream);
/* grav_adjoint_mu for one universe over ceil(n/1024) workgroups instead of one (its scale
 * is a reduction over the universe, done through the workspace: 4 bytes per 1024 bodies,
 * 4-byte aligned); the same nu and mu4 bit for bit.  (No reference counterpart.) */
int grav_adjoint_mu_ws(const void *rho64, void *nu64, void *mu4, int64_t n, double dt,
                       void *workspace, int64_t workspace_bytes, void *stream);

/*
 * Per-step batched path (one launch per stage per step, the whole GPU shared by all
 * universes' i-tiles of 64*i_per_lane bodies, largest universes first; no LDS capacity
 * limit on a universe's size beyond the plan's):
 *   grav_batched_stage1_f32     acc[3*n] of every universe (FP32 compensated), j_waves in
 *                               {1,2,4,8} wavefronts splitting each tile's j-range;
 *   grav_batched_adjoint_mu     grav_adjoint_mu for every universe, each with its own E;
 *   grav_batched_stage1_vjp_f32 the all-pairs adjoint part of grav_stage1_vjp_f32 for every
 *                               universe (plan i_per_lane <= 4).
 * Stage 2 acts per body and takes the concatenated arrays directly.  (Adjoint: no reference
 * counterpart.)
 */
int grav_batched_stage1_f32(const grav_batched_plan *plan, const void *posm, void *acc,
                            double G, double eps, int j_waves, void *stream);
int grav_batched_adjoint_mu(const grav_batched_plan *plan, const void *rho64, void *nu64,
                            void *mu4, double dt, void *stream);
int grav_batched_stage1_vjp_f32(const grav_batched_plan *plan, const void *posm,
                                const void *mu4, void *rho64, void *gamma64, double G,
                                double eps, double dt, int j_waves, void *stream);
int grav_stage1_vjp_f32(const void *posm, const void *mu4, void *rho64, void *gamma64,
                        int64_t n, double G, double eps, double dt, void *stream);
int grav_batched_adjoint_f32(const grav_batched_plan *plan, const void *traj, void *rho64,
                             void *nu64, void *gamma64, double G, double eps, double dt,
                             int steps, void *stream);

/*
 * Per-step batched stage 1 in the symmetric form (each unordered pair of a universe once, as
 * grav_stage1_sym_f32 does for one universe): grav_batched_sym_prepare cuts every universe
 * into blocks of 64*i_chunks_per_wave*waves bodies ({4,8} x {1,2,4}) and lists their block
 * pairs (host work and device lists, at set-up time); grav_batched_sym_workspace gives the
 * bytes of the caller's workspace (256-byte aligned: 12 n + 24 * block * (block pairs));
 * grav_batched_stage1_sym_f32 writes acc[3*n] with grav_batched_stage1_f32's contract.
 * (No reference counterpart: the reference holds one universe per kernel object.)
 */
int grav_batched_sym_prepare(grav_batched_plan *plan, int i_chunks_per_wave, int waves);
int grav_batched_sym_workspace(const grav_batched_plan *plan, int64_t *bytes);
int grav_batched_stage1_sym_f32(const grav_batched_plan *plan, const void *posm, void *acc,
                                double G, double eps, void *workspace, int64_t workspace_bytes,
                                void *stream);

/*
 * j-split form of stage 1 (FP32, owner-computes) and of its VJP, for universes too small to
 * fill the GPU with i-tiles alone (below about 40000 bodies): each i-tile of 64*i_per_lane
 * bodies is also split over `splits` equal j-ranges, each workgroup writes its FP64 partial
 * to the caller's workspace and one small kernel sums the ranges in a fixed order, so the
 * results stay deterministic.  Result contracts as grav_stage1_f32 (compensated) and
 * grav_stage1_vjp_f32.  (i_per_lane, j_waves, splits) in {1, 2} x {1, 2, 4, 8} x [1, 64], or
 * (0, 0, 0) for the library's choice by n.  Workspace: 256-byte aligned, the bytes
 * grav_stage1_split_workspace returns for `splits` (0: the default's), 32 * splits * n.
 * Same reference interfaces as grav_stage1_f32 and (none) grav_stage1_vjp_f32.
 */
int grav_stage1_split_workspace(int64_t n, int splits, int64_t *bytes);
int grav_stage1_f32_split(const void *posm, void *acc, int64_t n, double G, double eps,
                          int i_per_lane, int j_waves, int splits, void *workspace,
                          int64_t workspace_bytes, void *stream);
/*
 * Stage 1 of one share of the bodies, for a universe divided among GPUs by i-range: the FP32
 * compensated accelerations acc float[3][m] of the bodies [i0, i0 + m) of posm float4[n]
 * against all n bodies (every GPU holds all n positions, e.g. after an all-gather).
 * splits == 1: owner-computes, (i_per_lane, j_waves) as grav_stage1_f32_cfg, no workspace;
 * splits > 1: the j-split form, shapes as grav_stage1_f32_split, with a workspace of
 * grav_stage1_split_workspace(m, splits) bytes.  (0, 0, 0): grav_stage1_shard_cfg's shape.
 * The rows equal those of the whole universe's grav_stage1_f32_cfg (j_block 256) or
 * grav_stage1_f32_split of the same shape bit for bit, so the shares of P GPUs put together
 * are the one-GPU result.  Replaces one worker's row batch in the reference's division of
 * stage 1 among processes (src/gravitation/kernel/np4.py:101-118, :135-163), where every
 * worker holds all positions, sums its rows' pairs once and the host adds the workers'
 * arrays; here each share owns its rows and needs no reduction.
 */
int grav_stage1_shard_cfg(int64_t n, int64_t m, int *i_per_lane, int *j_waves, int *splits);
int grav_stage1_f32_shard(const void *posm, void *acc, int64_t n, int64_t i0, int64_t m,
                          double G, double eps, int i_per_lane, int j_waves, int splits,
                          void *workspace, int64_t workspace_bytes, void *stream);
int grav_stage1_vjp_f32_split(const void *posm, const void *mu4, void *rho64, void *gamma64,
                              int64_t n, double G, double eps, double dt, int i_per_lane,
                              int j_waves, int splits, void *workspace, int64_t workspace_bytes,
                              void *stream);
/* The VJP of one GPU's share of the rows (no reference counterpart: reverse mode is new): the
 * rho64 += / gamma64 -= update of grav_stage1_vjp_f32_split for the bodies [i0, i0 + m) of
 * posm float4[n] only, rho64 double[3][m], gamma64 double[m]; mu4 float4[n] holds every body's
 * scaled cotangent with the universe's one exponent (ranks agree on it, then all-gather mu4).
 * splits == 1: owner-computes, (i_per_lane, j_waves) in {1, 2} x {1, 2, 4, 8} or {4} x
 * {1, 2, 4}; splits > 1: the j-split form, workspace of grav_stage1_split_workspace(m, splits)
 * bytes; (0, 0, 0): grav_stage1_vjp_shard_cfg's shape.  The rows equal the whole-universe VJP's
 * of the same shape bit for bit when i0 is a multiple of 64 * i_per_lane (256 covers every
 * shape; the self-masked pairs then fall in the same compensation blocks), and to FP32
 * rounding otherwise. */
int grav_stage1_vjp_shard_cfg(int64_t n, int64_t m, int *i_per_lane, int *j_waves, int *splits);
int grav_stage1_vjp_f32_shard(const void *posm, const void *mu4, void *rho64, void *gamma64,
                              int64_t n, int64_t i0, int64_t m, double G, double eps, double dt,
                              int i_per_lane, int j_waves, int splits, void *workspace,
                              int64_t workspace_bytes, void *stream);
/* One time step in the j-split form with FP64 masters: grav_stage1_f32_split followed by
 * grav_stage2_f64m (posm float4[n], r64 and v64 double[3][n], updated in place), with the
 * reduce and stage 2 in one kernel and no acc written; bit for bit the same r64, v64 and
 * posm as the two calls.  Same shapes and workspace as grav_stage1_f32_split.  Replaces
 * stage 1 + stage 2 of universe_base.step (src/gravitation/kernel/_base_.py:136-161). */
int grav_step_split_f64m(void *posm, void *r64, void *v64, int64_t n, double G, double eps,
                         double dt, int i_per_lane, int j_waves, int splits, void *workspace,
                         int64_t workspace_bytes, void *stream);
/* grav_step_split_f64m with the positions read from posm_in and the new ones (x, y, z, m)
 * written to posm_out, which is posm_in (the same as grav_step_split_f64m) or a buffer that
 * does not overlap it: a step from one record slot of a trajectory to the next without a
 * copy.  r64, v64 are updated in place.  Same stage 1 + stage 2 of universe_base.step
 * (src/gravitation/kernel/_base_.py:136-161). */
int grav_step_split_f64m_io(const void *posm_in, void *posm_out, void *r64, void *v64,
                            int64_t n, double G, double eps, double dt, int i_per_lane,
                            int j_waves, int splits, void *workspace, int64_t workspace_bytes,
                            void *stream);
/* One backward step of a chain in the j-split form, two launches instead of four (no
 * reference counterpart: reverse mode is new): grav_stage1_vjp_f32_split at positions posm
 * with the scaled cotangent mu4, then, if next_mu is 1, the next backward step's
 * grav_adjoint_mu in the same reduce kernel: nu <- nu + dt rho (the updated rho) and mu4 <- its
 * FP32 copy scaled by 2^-E.  E is the exponent of the PREVIOUS mu (lagged by one step;
 * grav_adjoint_mu uses the current one), so mu4.xyz may exceed 1 in magnitude; rho and gamma
 * are those of the grav_adjoint_mu + grav_stage1_vjp_f32_split sequence bit for bit unless a
 * scaled component is subnormal.  With next_mu 0 (the trajectory's first step) nu and mu4 are
 * left alone.  eslot: a device uint32[2] carrying the exponents from step to step, zero
 * before the chain's first step (whose mu4 comes from grav_adjoint_mu); steps of one chain
 * run in order on one stream, and chains that share an eslot must not overlap. */
int grav_stage1_vjp_f32_split_chain(const void *posm, void *mu4, void *rho64, void *nu64,
                                    void *gamma64, int64_t n, double G, double eps, double dt,
                                    int i_per_lane, int j_waves, int splits, int next_mu,
                                    void *workspace, int64_t workspace_bytes, void *eslot,
                                    void *stream);
/* One chained backward step (grav_stage1_vjp_f32_split_chain with next_mu = 1, at positions
 * posm_adj) and, independent of it, one recomputed forward step (grav_step_split_f64m_io from
 * posm_in to posm_out with the masters r64, v64) in the same two launches: the first carries
 * both steps' pair partials side by side, the second both reduces.  For the backward pass that
 * recomputes one checkpoint segment while it runs the adjoint steps of the next, without a
 * second stream; results are those of the two calls bit for bit.  Shapes: the VJP's
 * (vjp_i_per_lane, j_waves, vjp_splits) and the step's (step_i_per_lane, j_waves, step_splits)
 * in {1, 2} x {4, 8} x [1, 64], or all 0 for the library's choice by n.  workspace: 256-byte
 * aligned, at least grav_backward_pair_workspace bytes (the two forms' partials).  No reference
 * counterpart (reverse mode is new; the step is universe_base.step's stage 1 + stage 2,
 * src/gravitation/kernel/_base_.py:136-161). */
int grav_backward_pair_workspace(int64_t n, int vjp_splits, int step_splits, int64_t *bytes);
int grav_backward_pair_split(const void *posm_adj, void *mu4, void *rho64, void *nu64,
                             void *gamma64, void *eslot, const void *posm_in, void *posm_out,
                             void *r64, void *v64, int64_t n, double G, double eps, double dt,
                             int vjp_i_per_lane, int vjp_splits, int step_i_per_lane,
                             int step_splits, int j_waves, void *workspace,
                             int64_t workspace_bytes, void *stream);

/* grav_stage1_f64's contract in the j-split form; (i_per_lane, j_waves, splits) in
 * {1, 2} x {4, 8} x [1, 64], or (0, 0, 0); the same workspace size rule. */
int grav_stage1_f64_split(const void *posm64, void *acc64, int64_t n, double G, double eps,
                          int i_per_lane, int j_waves, int splits, void *workspace,
                          int64_t workspace_bytes, void *stream);

/*
 * grav_stage1_f32_split with no workspace argument and no allocation: the partial sums of the
 * j-ranges live in a file-scope device array of the library, GRAV_NOALLOC_SCRATCH_BYTES
 * (32 MiB) on every GPU the library runs on, held until the process ends.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same shapes, result
 * contract and bits as grav_stage1_f32_split; requests whose partials need more than the array
 * (32 * splits * n bytes, splits = the default's for n when the shape is (0, 0, 0)) return
 * GRAV_EINVAL.  The first call on a device looks the array's address up (no allocation).
 * Every call on one GPU uses that GPU's one array: calls that overlap on one GPU (other
 * streams or threads) race; calls in stream order do not.
 */
#define GRAV_NOALLOC_SCRATCH_BYTES (32LL * 1024 * 1024)
int grav_stage1_f32_split_noalloc(const void *posm, void *acc, int64_t n, double G, double eps,
                                  int i_per_lane, int j_waves, int splits, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GRAVITATION_HIP_H */
