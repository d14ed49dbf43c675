"""Benchmark of the gravitation hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--skip-batched]
                    [--leg NAME]

Workload (the headline `value`): one universe of N = 2^18 bodies in a Plummer sphere (scale
radius 1, truncated at 10, equal masses 1/N, zero velocities), G = 1, eps = 1e-3, dt = 1e-4.
One step is FP32-compensated stage 1 plus the fused stage-2 kernel with FP64 masters, on
device-resident state (no host copies inside the timed region).  Stage 1 runs in both of its
forms -- owner-computes (every ordered pair evaluated) and symmetric (every unordered pair
once, both bodies updated) -- from the same initial state; the faster one is the headline and
both are reported under `stage1_forms`.

Further legs go into the same JSON line: `batched_forward` (4096 uneven universes, 100
steps in one hipGraph replay, against the plain per-universe loop), `batched_fwd_adjoint`
(512 uneven universes, T = 500, forward + adjoint gradients, against the plain per-universe
autograd loop), `single_fwd_adjoint` (one 4096-body Plummer universe, T = 200, checkpoints
every 20 steps) and `orbit_fit` (50 gradient-descent iterations on the initial velocities of
the 512 universes); `plugin_step` times hip1's step() through the plugin contract with the
host views refreshed every step (PCIe included), as the reference's worker times a kernel;
`c2_sizes` times stage 1 at 2^14, 2^16, 2^18 and 2^20 bodies in the form hip1 picks.

Metric: pair interactions per second, counted as N^2 per step (the self term included, as it
costs a pair evaluation), in units of 1e9; and the stage-1 kernel's fraction of the MI355X FP32
vector peak at 20 flop per interaction.  Stage-1 kernel time is measured with HIP events on
the stream the kernels run on.  The CPU baseline (the reference's own SSE + OpenMP C stage 1,
built into oracle/_ref/, or the NumPy restatement of np1 when that is absent) runs on this
host in the same process, on a bounded sample, and is reported, not targeted.
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector peak (datasheet; one wave64 FMA per 2 cycles per SIMD)
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector peak (datasheet)
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E
FLOP_PER_INTERACTION = 20
VJP_FLOP_PER_PAIR = 50         # informational count of one ordered pair of the adjoint
STAGE2_BYTES_F64M = 124      # per body: read r 24 + v 24 + a 12, write r 24 + v 24 + pos 16
STAGE2_BYTES_F32 = 68        # per body: read pos 16 + v 12 + a 12, write pos 16 + v 12


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=20)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--n', type=int, default=1 << 18)
    p.add_argument('--eps', type=float, default=1e-3)
    p.add_argument('--dt', type=float, default=1e-4)
    p.add_argument('--cfg', type=str, default=None,
                   help="stage-1 launch shape: 'i_per_lane,j_waves,j_block' (owner), "
                        "'i_per_lane,j_waves,splits' (split) or 'i_chunks_per_wave,waves' "
                        "(symmetric); needs --stage1; with --leg shard the one share shape "
                        "'i_per_lane,j_waves,splits' to time")
    p.add_argument('--stage1', choices=('auto', 'owner', 'split', 'symmetric'), default='auto',
                   help='stage-1 form; auto times both and reports the faster')
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--leg', choices=('plugin_step', 'c2_sizes', 'shard', 'galaxy_worker',
                                     'batched_forward', 'batched_shard',
                                     'batched_fwd_adjoint',
                                     'single_fwd_adjoint', 'orbit_fit'), default=None,
                   help='run only this leg and print its JSON (for profiling it alone)')
    p.add_argument('--skip-batched', action='store_true',
                   help='only the single-universe legs')
    p.add_argument('--only-headline', action='store_true',
                   help='only the headline stage-1 forms at --n (no other leg, no CPU '
                        'baseline): the run the counter profiles are taken on')
    p.add_argument('--skip-orbit-fit', action='store_true',
                   help='skip the 50-iteration batched orbit-fitting leg')
    p.add_argument('--cpu-quick', action='store_true',
                   help='CPU baseline: only the rows at the headline size')
    p.add_argument('--batched-schedule', type=str, default=None,
                   metavar='MODE,WAVES,I_PER_LANE,TILES_PER_WAVE,GROUPS,J_WAVES',
                   help="batched_forward leg only: run this schedule alone instead of the "
                        "tuner's pick (no tuning, no plain loop; for profiling its kernels)")
    p.add_argument('--calibrate-copy', type=int, default=0, metavar='BYTES',
                   help='first run one device-to-device copy of BYTES (the known-size kernel '
                        'that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters)')
    return p.parse_args()


def _traffic_from_profiles(kernel, n):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by
    a copy-kernel calibration) and written into profiles/; None when not measured for n."""
    path = os.path.join(ROOT, 'profiles', '%s_n%d_traffic.md' % (kernel, n))
    if not os.path.exists(path):
        return None
    m = re.search(r'^hbm_bytes_per_launch:\s*([0-9.eE+]+)', open(path).read(), re.M)
    return float(m.group(1)) if m else None


def _host_cpus():
    try:
        import psutil
        phys = psutil.cpu_count(logical=False)
        logical = psutil.cpu_count(logical=True)
    except Exception:
        phys = logical = os.cpu_count()
    return {'physical': phys, 'logical': logical, 'affinity': len(os.sched_getaffinity(0))}


def _cpu_time(fn, reps):
    """min and median wall time [s] of `reps` calls, timed as the reference's worker times a
    step (src/gravitation/cli/worker.py:119-136, :220): gc disabled, gc.collect() before each
    call, perf_counter_ns around it."""
    import gc
    was = gc.isenabled()
    gc.disable()
    times = []
    try:
        for _ in range(reps):
            gc.collect()
            t0 = time.perf_counter_ns()
            fn()
            times.append((time.perf_counter_ns() - t0) * 1e-9)
    finally:
        if was:
            gc.enable()
    return float(np.min(times)), float(np.median(times))


def _cpu_baseline(pos, m, quick=False, reps=3):
    """The CPU baselines on this host, reported (not targeted): the oracle's restatements of
    the reference's NumPy kernels np1 (one process, src/gravitation/kernel/np1.py:57-108) and
    np4 (joblib processes, one per physical core, np4.py:59-166), and the reference's own SSE +
    OpenMP C stage 1 (_lib4_/lib.c, built into oracle/_ref/).  Each row gives the minimum and
    median of `reps` steps.  np1 runs in full up to 2^14 bodies and np4 up to 2^16; at the
    headline's 2^18 both run a strided subset of the rows and the time is scaled by the exact
    pair count sum(N-1-i) of all rows over that of the subset (labelled extrapolated).
    Interactions are counted as N^2 per step, as on the GPU.  The headline value is np4's
    extrapolated rate at 2^18 (kind 'port'); `quick` keeps only the 2^18 rows."""
    from oracle import oracle as O
    from oracle import ref as R
    cpus = _host_cpus()
    workers = max(1, min(cpus['physical'] or 1, cpus['affinity']))
    n_full = pos.shape[0]
    table = []

    def row(kernel, n, mode, cores, t, sample=None):
        t_min, t_med = t
        d = {'kernel': kernel, 'n_bodies': n, 'mode': mode, 'cores': cores, 'steps': reps,
             's_min': t_min, 's_median': t_med,
             'Ginteractions_per_s': float(n) * n / t_min / 1e9}
        if sample:
            d['sample'] = sample
        table.append(d)
        return d

    def extrapolated(kernel, stride, fn, cores):
        # fn(rows) over strided row subsets at two strides; t = t0 + c * pairs separates the
        # per-step fixed cost t0 (process dispatch, result transfer) from the pair work c,
        # and the full step is t0 + c * N(N-1)/2
        fits = []
        for st in (2 * stride, stride):
            rows = np.arange(0, n_full - 1, st)
            fits.append((float(np.sum(n_full - 1 - rows)), _cpu_time(lambda: fn(rows), reps)))
        (p2, t2), (p1, t1) = fits
        full = n_full * (n_full - 1) / 2.0
        est = []
        for k in range(2):
            c = (t1[k] - t2[k]) / (p1 - p2)
            t0 = max(0.0, t1[k] - c * p1)
            est.append(t0 + c * full if c > 0 else t1[k] * full / p1)
        return row(kernel, n_full, 'extrapolated', cores, tuple(est),
                   'rows with stride %d and %d (%.4f and %.4f of all pairs: %.2f and %.2f s), '
                   'fitted as fixed cost + pair count' % (stride, 2 * stride, p1 / full,
                                                          p2 / full, t1[0], t2[0]))

    # warm the joblib process pool (process start-up is not a step)
    O.np4_rows(pos[:256], m[:256], 1.0, range(255), workers)
    if not quick:
        for n in (1024, 4096, 16384):
            p, w = pos[:n], m[:n]
            row('np1', n, 'full', 1, _cpu_time(lambda: O.np1_accel(p, w, 1.0), reps))
        for n in (16384, 65536):
            p, w = pos[:n], m[:n]
            row('np4', n, 'full', workers,
                _cpu_time(lambda: O.np4_rows(p, w, 1.0, range(n - 1), workers), reps))
        extrapolated('np1', 640, lambda rows: O.np1_rows(pos, m, 1.0, rows), 1)
    head = extrapolated('np4', 48, lambda rows: O.np4_rows(pos, m, 1.0, rows, workers),
                        workers)
    threads = max(2, min(16, cpus['affinity']))
    if R.available('lib4') and n_full % 4 == 0:
        lib = R.RefLib4(pos, m, 1.0, threads)
        lib.accel()
        row('lib4 (reference C, SSE + OpenMP, _mm_rsqrt_ps)', n_full, 'full', threads,
            _cpu_time(lib.accel, reps))
    return {'value': head['Ginteractions_per_s'], 'unit': 'Ginteractions/s',
            'cores': workers, 'kind': 'port',
            'sample': 'np4 restatement (src/gravitation/kernel/np4.py:59-166; %d joblib '
                      'processes, one per physical core) on the N=%d headline workload: %s; '
                      'min of %d steps, N^2 interactions/step' % (workers, n_full,
                                                                   head['sample'], reps),
            'host_cpus': cpus, 'table': table}


def _sync_time(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _best_per_mode(table):
    """Tuner table -> {mode: fastest ms per step} (the evidence behind the pick)."""
    out = {}
    for cfg, t in table:
        out[cfg[0]] = min(out.get(cfg[0], float('inf')), t * 1e3)
    return out


def _best_schedule_per_mode(table):
    """Tuner table -> {mode: its fastest schedule}."""
    best = {}
    for cfg, t in table:
        if cfg[0] not in best or t < best[cfg[0]][1]:
            best[cfg[0]] = (cfg, t)
    return {mode: _schedule(cfg) for mode, (cfg, _) in best.items()}


def _parse_schedule(text):
    """'mode,waves,i_per_lane,tiles_per_wave,groups,j_waves' -> a batched schedule tuple."""
    parts = text.split(',')
    if len(parts) != 6 or parts[0] not in ('fused', 'perstep'):
        raise SystemExit('--batched-schedule: mode (fused or perstep) and five integers')
    return (parts[0],) + tuple(int(x) for x in parts[1:])


def _schedule(cfg):
    d = dict(zip(('mode', 'waves', 'i_per_lane', 'tiles_per_wave', 'groups', 'j_waves'), cfg))
    if len(cfg) > 6:
        d['sym_block'] = list(cfg[6])
    return d


def _leg_batched_forward(dev, steps=100, seed=0, schedule=None):
    """4096 universes, N log-uniform in [32, 2048], uniform-cube positions, Gaussian
    velocities (sigma 0.1), log-uniform masses [1e-4, 1e-2], G=1, eps=1e-3, dt=1e-3: all
    universes and `steps` steps in one hipGraph-captured fused launch, against the plain loop
    of single-universe launches (2 per universe-step)."""
    from gravitation_amd import batched, tune
    from gravitation_amd.lib import scenarios as S
    r, v, m, off = S.batched(4096, 32, 2048, seed)
    cost = float(np.sum(np.diff(off).astype(np.float64) ** 2))
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=dev)
    if schedule is None:
        # 20 steps per timed call: the fused schedule's per-launch cost (loading every
        # universe into LDS) weighs on a 4-step call 5x more than on the 100-step launch
        best, table = tune.tune_batched(b, steps=20)
    else:
        best, table = schedule, []
    b.set_config(best)
    g = b.capture(steps)
    g.replay()
    t = _sync_time(g.replay)
    if schedule is not None:
        # one schedule alone, for a profile of its kernels: no plain loop
        return {'schedule': _schedule(best), 'ms_per_step': t / steps * 1e3,
                'interactions_per_s': cost * steps / t, 'plain': 'not run'}
    plain = b.plain()
    plain.step(1)
    tp = _sync_time(lambda: plain.step(2)) / 2
    return {'workload': '4096 universes, N log-uniform [32, 2048], uniform cube, v~N(0,0.1), '
                        'm log-uniform [1e-4,1e-2], G=1, eps=1e-3, dt=1e-3, FP32 pair terms, '
                        'FP64 masters, %d steps per hipGraph launch' % steps,
            'bodies': int(b.n), 'interactions_per_step': cost,
            'ms_per_step': t / steps * 1e3, 'steps_per_s': steps / t,
            'interactions_per_s': cost * steps / t,
            'pct_fp32_peak': 100.0 * FLOP_PER_INTERACTION * cost * steps / t / 1e12
                             / PEAK_FP32_TFLOPS,
            'plain_ms_per_step': tp * 1e3, 'plain_steps_per_s': 1.0 / tp,
            'speedup_vs_plain': tp / (t / steps),
            'schedule': _schedule(best),
            'tuned_ms_per_step_by_mode': _best_per_mode(table),
            'tuned_schedule_by_mode': _best_schedule_per_mode(table),
            'plan': b.plan.info()}


def _leg_batched_shard(dev, shares=8, steps=100, seed=0):
    """The batched forward workload (4096 universes, N log-uniform in [32, 2048]) divided
    among `shares` GPUs by whole universes (``batched.shard_universes``: longest processing
    time first on N_u^2): every share tuned and stepped alone, `steps` steps per hipGraph
    launch, against the whole batch / shares.  Universes are independent, so a share needs no
    communication."""
    from gravitation_amd import batched, tune
    from gravitation_amd.lib import scenarios as S
    t0 = time.perf_counter()
    r, v, m, off = S.batched(4096, 32, 2048, seed)
    cost = np.diff(off).astype(np.float64) ** 2
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=dev)
    best, _ = tune.tune_batched(b, steps=20)
    b.set_config(best)
    g = b.capture(steps)
    g.replay()
    whole = _sync_time(g.replay) / steps
    del g, b
    rows = []
    for share in batched.shard_universes(off, shares):
        rs, vs, ms, os_ = batched.take_universes(r, v, m, off, share)
        bs = batched.Batch(rs, vs, ms, os_, G=1.0, eps=1e-3, dt=1e-3, device=dev)
        own, _ = tune.tune_batched(bs, steps=20)
        bs.set_config(own)
        gs = bs.capture(steps)
        gs.replay()
        t_own = _sync_time(gs.replay) / steps
        del gs, bs
        print('batched_shard: share %d of %d: %.3f ms per step' % (
            len(rows) + 1, shares, t_own * 1e3), file=sys.stderr, flush=True)
        rows.append({'universes': int(share.shape[0]), 'bodies': int(os_[-1]),
                     'interactions_per_step': float(cost[share].sum()),
                     'ms_per_step': t_own * 1e3, 'schedule': _schedule(own)})
    loads = np.array([x['interactions_per_step'] for x in rows])
    slowest = max(x['ms_per_step'] for x in rows)
    return {'workload': '4096 universes, N log-uniform [32, 2048] (the batched forward '
                        'workload) divided among %d GPUs by whole universes, %d steps per '
                        'hipGraph launch' % (shares, steps),
            'shares': shares, 'whole_ms_per_step': whole * 1e3,
            'whole_schedule': _schedule(best),
            'whole_ms_per_step_over_shares': whole * 1e3 / shares,
            'slowest_share_ms_per_step': slowest,
            'load_max_over_mean': float(loads.max() / loads.mean()),
            'per_share': rows, 'leg_seconds': time.perf_counter() - t0}


def _leg_batched_adjoint(dev, steps=500, plain_steps=10, seed=0):
    """512 universes, N log-uniform in [64, 4096], otherwise as the batched forward; loss
    sum |r_T - target|^2 over T=500 steps; forward + adjoint (gradients to r0, v0, m) batched
    against the plain per-universe loop of the same autograd function for one universe
    (4 launches per universe-step), inputs resident on the device for both.  The plain
    loop's cost per step is its marginal cost: (time at 2 T0 - time at T0) / T0 with
    T0 = `plain_steps`, which leaves out the fixed per-universe setup (leaf tensors, loss,
    autograd graph) the batched call pays once."""
    import torch

    from gravitation_amd import adjoint, batched, tune
    from gravitation_amd.lib import scenarios as S
    r, v, m, off = S.batched(512, 64, 4096, seed)
    cost = float(np.sum(np.diff(off).astype(np.float64) ** 2))
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=dev)
    best, table = tune.tune_batched_adjoint(b, steps=4)
    del b
    engine = batched.Engine(off, best)
    rng = np.random.default_rng(seed + 7)
    target = torch.from_numpy(r + rng.normal(0, 1e-2, r.shape)).to(dev)

    rd = torch.from_numpy(r).to(dev)
    vd = torch.from_numpy(v).to(dev)
    md = torch.from_numpy(m).to(dev)

    def run(rr, vv, mm, tg, fn):
        r0 = rr.detach().clone().requires_grad_(True)
        v0 = vv.detach().clone().requires_grad_(True)
        m0 = mm.detach().clone().requires_grad_(True)
        rT, _ = fn(r0, v0, m0)
        loss = ((rT - tg) ** 2).sum()
        loss.backward()
        return loss

    bat = lambda T: (lambda a, b_, c: adjoint.batched_trajectory(engine, a, b_, c, T, 1.0,
                                                                 1e-3, 1e-3))
    run(rd, vd, md, target, bat(2))
    tb = _sync_time(lambda: run(rd, vd, md, target, bat(steps)))

    parts = [(rd[a:e].contiguous(), vd[a:e].contiguous(), md[a:e].contiguous(),
              target[a:e].contiguous()) for a, e in zip(off[:-1], off[1:])]

    def plain(T, form):
        for ru, vu, mu, tu in parts:
            run(ru, vu, mu, tu, lambda x, y, z: adjoint.trajectory(x, y, z, T, 1.0, 1e-3, 1e-3,
                                                                   form=form))
    # the baseline is the faster of the two per-universe choices: forms by size (more
    # launches for the larger universes) or owner-computes only (one launch per stage)
    plain_by_form = {}
    for form in ('auto', 'owner'):
        plain(1, form)
        t1 = _sync_time(lambda: plain(plain_steps, form))
        t2 = _sync_time(lambda: plain(2 * plain_steps, form))
        plain_by_form[form] = (t1, t2, (t2 - t1) / plain_steps)
    plain_form = min(plain_by_form, key=lambda f: plain_by_form[f][2])
    t1, t2, tp = plain_by_form[plain_form]
    return {'workload': '512 universes, N log-uniform [64, 4096], uniform cube, v~N(0,0.1), '
                        'm log-uniform [1e-4,1e-2], G=1, eps=1e-3, dt=1e-3, T=%d; forward + '
                        'adjoint of sum |r_T - target|^2 w.r.t. r0, v0, m' % steps,
            'bodies': int(r.shape[0]), 'interactions_per_step': cost, 'steps': steps,
            'fwd_adjoint_s': tb, 'steps_per_s': steps / tb,
            'interactions_per_s_fwd_plus_adjoint': 2.0 * cost * steps / tb,
            'pct_fp32_peak_informational': 100.0 * (FLOP_PER_INTERACTION + VJP_FLOP_PER_PAIR)
                                           * cost * steps / tb / 1e12 / PEAK_FP32_TFLOPS,
            'plain_s_per_step': tp, 'plain_steps_per_s': 1.0 / tp,
            'plain_timed_steps': [plain_steps, 2 * plain_steps],
            'plain_s_at_timed_steps': [t1, t2],
            'plain_form': plain_form,
            'plain_s_per_step_by_form': {f: v[2] for f, v in plain_by_form.items()},
            'speedup_vs_plain': tp / (tb / steps),
            'schedule': _schedule(best),
            'tuned_ms_per_step_by_mode': _best_per_mode(table)}


def _leg_c2_sizes(dev, eps=1e-3):
    """FP32 stage 1 on the Plummer spheres of every size of the single-universe benchmark
    (2^14, 2^16, 2^18, 2^20; G = 1, eps = 1e-3) in the form hip1 takes by size: best of a few
    calls after a warm-up, HIP events on the stream."""
    import torch

    from gravitation_amd import stages
    from gravitation_amd.lib import scenarios as S
    out = {}
    for lg in (14, 16, 18, 20):
        n = 1 << lg
        r, _, m = S.plummer(n, seed=0)
        posm = torch.empty((n, 4), dtype=torch.float32)
        posm[:, :3] = torch.from_numpy(r)
        posm[:, 3] = torch.from_numpy(m)
        posm = posm.to(dev)
        acc = torch.empty((3, n), dtype=torch.float32, device=dev)
        form = stages.auto_form(n, 'kahan', torch.cuda.mem_get_info(dev)[0])
        fn = {'owner': lambda: stages.stage1_f32(posm, acc, 1.0, eps, 'kahan'),
              'split': lambda: stages.stage1_split_f32(posm, acc, 1.0, eps),
              'symmetric': lambda: stages.stage1_sym_f32(posm, acc, 1.0, eps)}[form]
        fn()
        ts = []
        for _ in range(2 if lg == 20 else 5):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        t = min(ts)
        stages.release_workspaces()
        out['2^%d' % lg] = {'n_bodies': n, 'stage1_form': form, 'stage1_ms': t * 1e3,
                            'interactions_per_s': float(n) * n / t,
                            'pct_fp32_peak': 100.0 * FLOP_PER_INTERACTION * float(n) * n / t
                                             / 1e12 / PEAK_FP32_TFLOPS}
    return out


SHARD_CANDIDATES = [(2, 16, 1), (1, 16, 1), (2, 8, 1), (1, 8, 1), (4, 8, 1),
                    (1, 8, 8), (1, 8, 16), (1, 8, 64), (2, 8, 8), (2, 8, 16), (2, 8, 32),
                    (2, 8, 64), (2, 4, 32), (2, 4, 64)]


def _leg_shard(dev, n=1 << 18, shares=8, eps=1e-3, reps=5, candidates=None):
    """Stage 1 of one GPU's share of the 2^18-body Plummer sphere divided among `shares` GPUs
    by i-range: bodies [p n/P, (p+1) n/P) against all n (the positions every GPU holds after
    the all-gather).  Every launch shape of both forms (or ``candidates``) on share 0, then
    every share with the fastest, beside the whole universe's owner-computes time / P.  HIP
    events on the stream, best of `reps` calls after a warm-up."""
    import torch

    from gravitation_amd import stages
    from gravitation_amd.lib import scenarios as S
    r, _, m = S.plummer(n, seed=0)
    posm = torch.empty((n, 4), dtype=torch.float32)
    posm[:, :3] = torch.from_numpy(r)
    posm[:, 3] = torch.from_numpy(m)
    posm = posm.to(dev)
    size = n // shares

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return min(ts)

    acc = torch.empty((3, size), dtype=torch.float32, device=dev)

    def share(i0, cfg):
        ws = stages.split_workspace(posm, size, cfg) if cfg[2] > 1 else None
        return lambda: stages.stage1_shard_f32(posm, acc, i0, 1.0, eps, cfg, ws)

    table = [(cfg, timed(share(0, cfg))) for cfg in (candidates or SHARD_CANDIDATES)]
    default = tuple(stages.shard_cfg(n, size))
    default_ms = dict(table).get(default)
    if default_ms is None:
        default_ms = timed(share(0, default))
    best, best_ms = min(table, key=lambda x: x[1])
    per_share = [timed(share(p * size, best)) for p in range(shares)]
    whole = torch.empty((3, n), dtype=torch.float32, device=dev)
    whole_ms = timed(lambda: stages.stage1_f32(posm, whole, 1.0, eps, 'kahan'))
    stages.release_workspaces()
    # the symmetric form divided by rows of block pairs: each share's FP64 sums of all n
    # bodies, then the combination of the shares' sums in share order
    sums = torch.empty((shares, 3, n), dtype=torch.float64, device=dev)
    sym_ms = []
    for p in range(shares):
        sym_ms.append(timed(lambda: stages.stage1_sym_f32_share(posm, sums[p], p, shares, eps)))
        stages.release_workspaces()
    combine_ms = timed(lambda: stages.sym_shares_combine(sums, 1.0, whole))
    sym_whole_ms = timed(lambda: stages.stage1_sym_f32(posm, whole, 1.0, eps))
    stages.release_workspaces()
    # the adjoint's VJP on one share of the rows (mu4 of all n bodies, as after the all-gather)
    from gravitation_amd import adjoint
    mu4 = torch.empty((n, 4), dtype=torch.float32, device=dev)
    nu = torch.ones((3, n), dtype=torch.float64, device=dev)
    adjoint.adjoint_mu(torch.zeros_like(nu), nu, mu4, 1.0)
    rho = torch.zeros((3, n), dtype=torch.float64, device=dev)
    gam = torch.zeros(n, dtype=torch.float64, device=dev)
    vjp_whole_ms = timed(lambda: adjoint.stage1_vjp_split_f32(posm, mu4, rho, gam, 1.0, eps,
                                                              1e-4))
    stages.release_workspaces()
    vjp_cfg = adjoint.vjp_shard_cfg(n, size)
    rs = torch.zeros((3, size), dtype=torch.float64, device=dev)
    gs = torch.zeros(size, dtype=torch.float64, device=dev)
    vjp_share_ms = timed(lambda: adjoint.stage1_vjp_shard_f32(posm, mu4, rs, gs, 0, 1.0, eps,
                                                              1e-4))
    stages.release_workspaces()
    del mu4, nu, rho, gam, rs, gs
    rows = [stages.sym_share_plan(n, p, shares)[0] for p in range(shares)]
    nb = rows[-1][5]
    pairs = [sum(nb - 1 - b for b in list(range(r[0], r[1])) + list(range(r[2], r[3])))
             for r in rows]
    inter = float(size) * n
    return {'workload': 'share of %d of a Plummer sphere of %d bodies (eps=%g): %d x %d '
                        'interactions' % (size, n, eps, size, n),
            'n_bodies': n, 'shares': shares, 'share_bodies': size,
            'default_cfg': list(default), 'default_ms': default_ms,
            'best_cfg': list(best), 'best_ms': best_ms,
            'interactions_per_s': inter / (best_ms * 1e-3),
            'pct_fp32_peak': 100.0 * FLOP_PER_INTERACTION * inter / (best_ms * 1e-3) / 1e12
                             / PEAK_FP32_TFLOPS,
            'every_share_ms': per_share,
            'whole_owner_ms': whole_ms, 'whole_owner_ms_over_shares': whole_ms / shares,
            'table_ms': [[list(c), t] for c, t in table],
            'vjp': {'default_cfg': list(vjp_cfg), 'share_ms': vjp_share_ms,
                    'whole_j_split_ms': vjp_whole_ms,
                    'whole_ms_over_shares': vjp_whole_ms / shares},
            'symmetric': {'rows_partner_rows_diagonal_blocks': [list(r) for r in rows],
                          'block_pairs': pairs, 'every_share_ms': sym_ms,
                          'slowest_share_ms': max(sym_ms),
                          'sums_bytes_per_share': 3 * 8 * n,
                          'combine_ms': combine_ms,
                          'whole_ms': sym_whole_ms,
                          'whole_ms_over_shares': sym_whole_ms / shares}}


def _stage2_large(dev, n=1956492, reps=20):
    """Stage 2 alone on n bodies (the C3 body count: 243 MB of algorithmic traffic in the
    FP64-master mode, about the Infinity Cache's size, where the 2^18 case is cache-resident),
    in the FP64-master mode (k_stage2_f64m, 124 B per body) and the FP32 mode (k_stage2_f32,
    68 B per body): average of `reps` launches each, HIP events on the stream."""
    import torch

    from gravitation_amd import stages
    g = torch.Generator(device=dev).manual_seed(0)
    posm = torch.rand((n, 4), dtype=torch.float32, device=dev, generator=g)
    r64 = torch.rand((3, n), dtype=torch.float64, device=dev, generator=g)
    v64 = torch.rand((3, n), dtype=torch.float64, device=dev, generator=g)
    acc = torch.rand((3, n), dtype=torch.float32, device=dev, generator=g)
    v32 = torch.rand((3, n), dtype=torch.float32, device=dev, generator=g)

    def avg(fn):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    out = {'n_bodies': n}
    for mode, nbytes, fn in (
            ('f64m', STAGE2_BYTES_F64M, lambda: stages.stage2_f64m(posm, r64, v64, acc, 1e-6)),
            ('f32', STAGE2_BYTES_F32, lambda: stages.stage2_f32(posm, v32, acc, 1e-6))):
        t = avg(fn)
        out[mode] = {'ms_avg': t * 1e3, 'bytes_per_body': nbytes,
                     'GB_per_s': nbytes * n / t / 1e9,
                     'pct_hbm_peak': 100.0 * nbytes * n / t / 1e9 / PEAK_HBM_GBS}
    return out


def _leg_stage1_f64(dev, r, m, eps, reps=3):
    """Informational: FP64 stage 1 (hip1's dtype='float64' mode) on the headline workload in
    both forms, HIP events around each call; 20 flop per interaction against the FP64 vector
    peak."""
    import torch

    from gravitation_amd import stages
    n = r.shape[0]
    posm = torch.empty((n, 4), dtype=torch.float64)
    posm[:, :3] = torch.from_numpy(r)
    posm[:, 3] = torch.from_numpy(m)
    posm = posm.to(dev)
    acc = torch.empty((3, n), dtype=torch.float64, device=dev)
    inter = float(n) * n
    out = {}
    for form, fn in (('owner', lambda: stages.stage1_f64(posm, acc, 1.0, eps)),
                     ('symmetric', lambda: stages.stage1_sym_f64(posm, acc, 1.0, eps))):
        fn()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        t = float(np.mean(ts))
        tf = FLOP_PER_INTERACTION * inter / t / 1e12
        out[form] = {'ms_avg': t * 1e3, 'interactions_per_s': inter / t, 'TFLOP_per_s': tf,
                     'pct_fp64_peak': 100.0 * tf / PEAK_FP64_TFLOPS}
    stages.release_workspaces()
    return out


def _leg_single_adjoint(dev, n=4096, steps=200, interval=20, seed=4):
    """One Plummer universe of n bodies, FP32 stage 1 + FP64-master stage 2, T = `steps`,
    dt = 1e-4, eps = 1e-3, an adjoint checkpoint every `interval` steps; loss
    sum |r_T - target|^2 with the target integrated from the initial state perturbed by
    N(0, 1e-3^2); forward + adjoint (gradients to r0, v0, m).  The adjoint recomputes each
    segment's forward (all but its last step, whose result is the next checkpoint), so a
    differentiated step costs 3 - 1/interval all-pairs passes: forward, recompute and VJP."""
    import torch

    from gravitation_amd import adjoint
    from gravitation_amd.lib import scenarios as S
    G, eps, dt = 1.0, 1e-3, 1e-4
    r0, v0, m = S.plummer(n, seed=seed)
    rng = np.random.default_rng(seed + 40)
    rd, vd, md = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (r0, v0, m))
    rp = rd + torch.from_numpy(rng.normal(0, 1e-3, r0.shape)).to(dev)
    vp = vd + torch.from_numpy(rng.normal(0, 1e-3, v0.shape)).to(dev)
    with torch.no_grad():
        target, _ = adjoint.trajectory(rp, vp, md, steps, G, eps, dt, interval=interval)

    def forward():
        with torch.no_grad():
            adjoint.trajectory(rd, vd, md, steps, G, eps, dt, interval=interval)

    def run():
        a = rd.clone().requires_grad_(True)
        b = vd.clone().requires_grad_(True)
        c = md.clone().requires_grad_(True)
        rT, _ = adjoint.trajectory(a, b, c, steps, G, eps, dt, interval=interval)
        ((rT - target) ** 2).sum().backward()

    run()
    tf = min(_sync_time(forward) for _ in range(3))
    tb = min(_sync_time(run) for _ in range(3))
    pairs = float(n) * n * steps
    return {'workload': 'one Plummer universe, N=%d, FP32 stage 1 + FP64 stage 2, G=1, '
                        'eps=1e-3, dt=1e-4, T=%d, checkpoint every %d steps; forward + '
                        'adjoint of sum |r_T - target|^2 w.r.t. r0, v0, m'
                        % (n, steps, interval),
            'forward_s': tf, 'fwd_adjoint_s': tb, 'adjoint_s': tb - tf,
            'steps_per_s_fwd_adjoint': steps / tb,
            'interactions_per_s_forward': pairs / tf,
            'pair_passes_per_step': 3 - 1.0 / interval,
            'pct_fp32_peak_informational': 100.0 * ((2 - 1.0 / interval) * FLOP_PER_INTERACTION
                                                    + VJP_FLOP_PER_PAIR) * pairs / tb / 1e12
                                           / PEAK_FP32_TFLOPS}


def _leg_plugin_step(dev, steps=20):
    """hip1 through the reference's plugin contract: universe.step() wall clock with
    sync_host on (r and v copied back into the host views after every step, so the time
    includes the PCIe transfers), timed as the reference's worker does
    (src/gravitation/cli/worker.py:119-136, :220): gc disabled, gc.collect() between steps,
    minimum and median over `steps` steps after one warm-up step.  Workloads: the
    CPU-runnable 1024-body uniform cube (G = 6.6743e-11, T = 1, FP32) and the 2^18-body
    Plummer sphere of the headline (G = 1, eps = 1e-3, T = 1e-4, FP64 stage-2 masters)."""
    import gc

    from gravitation_amd.kernel import hip1
    from gravitation_amd.lib import scenarios as S
    out = {}
    cases = (('cube_n1024', S.uniform_cube(1024, seed=0), dict(T=1.0, G=6.6743e-11)),
             ('plummer_n262144', S.plummer(1 << 18, seed=0),
              dict(T=1e-4, G=1.0, eps=1e-3, stage2_dtype='float64')))
    for name, (r, v, m), kw in cases:
        u = hip1.universe(device=dev, **kw)
        u.add_objects(r.copy(), v.copy(), m.copy())
        u.start()
        u.step()
        times = []
        was = gc.isenabled()
        gc.disable()
        try:
            for _ in range(steps):
                t0 = time.perf_counter()
                u.step()
                times.append(time.perf_counter() - t0)
                gc.collect()
        finally:
            if was:
                gc.enable()
        # the same steps back to back (no gc.collect() gap between them): if the median drops
        # to the minimum, the spread above comes from the host gaps between steps
        b2b = []
        for _ in range(steps):
            t0 = time.perf_counter()
            u.step()
            b2b.append(time.perf_counter() - t0)
        n = len(u)
        form = u._form
        u.stop()
        t_min, t_med = float(np.min(times)), float(np.median(times))
        out[name] = {'n_bodies': n, 'stage1_form': form, 'step_ms_min': t_min * 1e3,
                     'step_ms_median': t_med * 1e3,
                     'back_to_back_ms_min': float(np.min(b2b)) * 1e3,
                     'back_to_back_ms_median': float(np.median(b2b)) * 1e3,
                     'interactions_per_s_at_min': float(n) * n / t_min}
    return out


# The reference's published best step() wall times for the galaxy scenario [ms]
# (demo/benchmark_long.log; Intel i5-4570 + GTX 1050 Ti, README.md:22): INPUT line / best
# STEP line per entry.
REFERENCE_GALAXY_MS = {
    'pc2 (PyCUDA all-pairs, reference GPU stage 1)': {
        1024: (0.5855, 'demo/benchmark_long.log:36224/:36249'),
        4096: (1.795, 'demo/benchmark_long.log:36960/:37077'),
        16384: (16.17, 'demo/benchmark_long.log:37636/:37695')},
    'np4 (NumPy + joblib, 4 processes)': {
        1024: (35.43, 'demo/benchmark_long.log:27008/:27053'),
        4096: (215.7, 'demo/benchmark_long.log:27192/:27209'),
        16384: (2169.0, 'demo/benchmark_long.log:27274/:27280')},
    'np1 (NumPy, 1 process)': {
        1024: (66.19, 'demo/benchmark_long.log:20741/:20808'),
        4096: (569.3, 'demo/benchmark_long.log:20973/:20988'),
        16384: (6990.0, 'demo/benchmark_long.log:21031/:21039')},
    'c4a (C SSE + OpenMP, 4 threads)': {
        16384: (139.9, 'demo/benchmark_long.log:6739/:6760')},
}


def _leg_galaxy_worker(dev, sizes=(1024, 4096, 16384), iterations=30):
    """The reference's own benchmark, like for like: the default `galaxy` scenario
    (src/gravitation/lib/simulation.py:57-83, :99-184; eps = 0, scale_r 1e-10, scale_m 1e-30)
    through the JSON-lines worker with the hip1 kernel (gravitation_amd.lib.worker, the
    semantics of src/gravitation/cli/worker.py:119-136: gc disabled, gc.collect() around each
    step, wall clock of universe.step() with the host views refreshed, BEST_TIME = the minimum),
    at the sizes the reference published, beside the reference's published best times."""
    import io
    import tempfile

    from gravitation_amd.lib import worker
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for n in sizes:
            log = io.StringIO()
            rc = worker.run_worker('hip1', 'galaxy', {'stars_len': n},
                                   os.path.join(tmp, 'g.npz'), min_iterations=iterations,
                                   min_total_runtime=0, seed=n,
                                   kernel_param={'device': str(dev)}, out=log)
            lines = [json.loads(x) for x in log.getvalue().split('\n') if x.strip()]
            if rc != 0:
                raise RuntimeError('worker failed: %s' % lines[-2:])
            steps = [d['runtime'] for d in lines if d['log'] == 'STEP']
            best = [d['value'] for d in lines if d['log'] == 'BEST_TIME'][-1]
            ref = {k: {'best_step_ms': v[n][0], 'source': v[n][1],
                       'hip1_speedup': v[n][0] / (best * 1e-6)}
                   for k, v in REFERENCE_GALAXY_MS.items() if n in v}
            out['stars_%d' % n] = {'best_step_ms': best * 1e-6,
                                   'median_step_ms': float(np.median(steps)) * 1e-6,
                                   'steps': len(steps),
                                   'directed_interactions_per_s_at_best':
                                       float(n) * (n - 1) / (best * 1e-9),
                                   'reference': ref}
    return out


def _orbit_fit_errors(dev, engine, r, v, m, off, target, steps, dt=1e-3, sample=6):
    """Loss and gradient error of the batched forward + adjoint (the first orbit-fitting
    iteration's) against the FP64 oracle.  Gradients: on universes sampled across the size
    range, the oracle's backward recursion (oracle.adjoint_backward, FP64 VJP of the C oracle)
    along the trajectory the GPU took, from the GPU's own r_T (chaos makes whole T = 500
    trajectories incomparable beyond a few hundred bodies, see the error table's C5 rows).
    Loss: on the 8 smallest universes, against the oracle's FP64 trajectory, beside the
    oracle's own FP32 trajectory's error against it (FP32 arithmetic alone moves r_T this much
    over T = 500 steps of these chaotic universes, so that is the floor to compare with)."""
    import torch

    from gravitation_amd import adjoint
    from oracle import oracle as O
    from oracle import ref as R
    acc, vjp = (R.oracle_accel_c, R.oracle_vjp_c) if R.available('oracle') else (None, None)
    n = r.shape[0]
    tr = torch.tensor(r, device=dev)
    tv = torch.tensor(v, device=dev, requires_grad=True)
    tm = torch.tensor(m, device=dev)
    rT, _ = adjoint.batched_trajectory(engine, tr, tv, tm, steps, 1.0, 1e-3, dt)
    d2 = ((rT - torch.tensor(target, device=dev)) ** 2).sum(1)
    d2.sum().backward()
    g = tv.grad.cpu().numpy()
    r_t = rT.detach().cpu().numpy()
    loss_u = np.add.reduceat(d2.detach().cpu().numpy(), off[:-1])
    posm = torch.empty((n, 4), dtype=torch.float32, device=dev)
    posm[:, :3] = tr
    posm[:, 3] = tm
    r64 = tr.t().contiguous()
    v64 = torch.tensor(np.ascontiguousarray(v.T), device=dev)
    sizes = np.diff(off)
    order = np.argsort(sizes, kind='stable')
    picks = [int(order[int(q * (len(order) - 1))]) for q in np.linspace(0, 0.7, sample)]
    rows = np.concatenate([np.arange(off[u], off[u + 1]) for u in picks])
    rec = torch.empty((steps, rows.shape[0], 4), dtype=torch.float32)
    traj = torch.empty((1, n, 4), dtype=torch.float32, device=dev)
    idx = torch.tensor(rows, device=dev)
    t0 = time.perf_counter()
    for t in range(steps):        # the same forward again, one recorded step at a time
        engine.forward(posm, r64, v64, 1.0, 1e-3, dt, 1, traj)
        rec[t] = traj[0].index_select(0, idx).cpu()
    replayed = bool(np.array_equal(r64.t().cpu().numpy(), r_t))
    rec = rec.numpy().astype(np.float64)
    worst = 0.0
    o = 0
    for u in picks:
        a, e = off[u], off[u + 1]
        k = e - a
        _, gv, _ = O.adjoint_backward([rec[t, o:o + k, :3] for t in range(steps)],
                                      rec[0, o:o + k, 3], 1.0, 1e-3, dt,
                                      2.0 * (r_t[a:e] - target[a:e]), np.zeros((k, 3)), vjp)
        worst = max(worst, O.rel_err(g[a:e], gv))
        o += k
    gpu_err, o32_err = [], []
    for u in order[:8]:
        a, e = off[u], off[u + 1]
        r32, _, _ = O.simulate(r[a:e], v[a:e], m[a:e], 1.0, 1e-3, dt, steps, np.float32, acc)
        r64o, _, _ = O.simulate(r[a:e], v[a:e], m[a:e], 1.0, 1e-3, dt, steps, np.float64, acc)
        lo = float(((r64o - target[a:e]) ** 2).sum())
        gpu_err.append(abs(loss_u[u] - lo) / lo)
        o32_err.append(abs(float(((r32 - target[a:e]) ** 2).sum()) - lo) / lo)
    return {'grad_v0_rel_err_max': worst,
            'grad_universes': [int(sizes[u]) for u in picks],
            'grad_method': 'oracle adjoint (FP64 VJP) along the GPU trajectory, T=%d' % steps,
            'loss_rel_err_max': max(gpu_err), 'loss_rel_err_median': float(np.median(gpu_err)),
            'loss_rel_err_oracle_fp32_max': max(o32_err),
            'loss_rel_err_oracle_fp32_median': float(np.median(o32_err)),
            'loss_universes': [int(sizes[u]) for u in order[:8]],
            'loss_method': "per-universe loss against the oracle's FP64 trajectory; "
                           "'oracle_fp32' is the oracle's own FP32 run against the same",
            'trajectory_replayed_bitwise': replayed,
            'check_s': time.perf_counter() - t0}


def _leg_orbit_fit(dev, steps=500, iters=50, seed=0):
    """Batched differentiable orbit fitting: the 512 universes of the forward + adjoint leg,
    targets = positions after T = 500 steps from initial velocities perturbed by N(0, 0.01^2);
    50 iterations of gradient descent on v0 from the unperturbed velocities."""
    import torch

    from gravitation_amd import adjoint, batched, tune
    from gravitation_amd.lib import scenarios as S
    r, v, m, off = S.batched(512, 64, 4096, seed)
    rng = np.random.default_rng(seed + 11)
    v_true = v + rng.normal(0, 1e-2, v.shape)
    b = batched.Batch(r, v_true, m, off, G=1.0, eps=1e-3, dt=1e-3, device=dev)
    best, _ = tune.tune_batched_adjoint(b, steps=4)
    b.set_config(best)
    b.step(steps)
    target = b.r64.t().contiguous()
    engine = b.engine
    errors = _orbit_fit_errors(dev, engine, r, v, m, off, b.r64.t().cpu().numpy(), steps)
    r0 = torch.from_numpy(r).to(dev)
    m0 = torch.from_numpy(m).to(dev)
    v0 = torch.from_numpy(v).to(dev)
    t0 = time.perf_counter()
    stats = {}
    v_fit, losses = adjoint.fit_initial_velocities(engine, r0, v0, m0, target, steps, iters,
                                                   G=1.0, eps=1e-3, dt=1e-3, stats=stats)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    vt = torch.from_numpy(v_true).to(dev)
    err_v = float((v_fit - vt).norm() / (v0 - vt).norm())
    e_fit = ((v_fit - vt) ** 2).sum(1).cpu().numpy()
    e_0 = ((v0 - vt) ** 2).sum(1).cpu().numpy()
    per_u = np.sqrt(np.add.reduceat(e_fit, off[:-1]) / np.add.reduceat(e_0, off[:-1]))
    return {'workload': '512 universes (as batched_fwd_adjoint), T=%d, %d gradient-descent '
                        'iterations on v0 towards targets from perturbed velocities'
                        % (steps, iters),
            'loss_first': losses[0], 'loss_last': losses[-1],
            'universes_improved': int((stats['best'] < stats['first']).sum()),
            'universes_10x': int((stats['best'] < 0.1 * stats['first']).sum()),
            'median_universe_loss_ratio': float(np.median(stats['best'] / stats['first'])),
            'rejected_updates': stats['rejected'],
            'v0_error_ratio': err_v,
            'v0_error_ratio_median_universe': float(np.median(per_u)),
            's_per_iteration': t / iters, 'total_s': t,
            'universe_steps_per_s': 512 * steps * iters / t,
            'first_iteration_vs_oracle': errors}


def main():
    args = _parse()
    if args.gpus != 1:
        print('bench.py: measures one GPU only (the stage kernels run on a single device; '
              'multi-GPU sharding is not implemented); --gpus %d is not supported' % args.gpus,
              file=sys.stderr)
        return 2
    import torch

    from gravitation_amd import _native, stages
    from gravitation_amd.lib import scenarios as S

    _native.load()
    if not torch.cuda.is_available():
        print('bench.py: no HIP device visible', file=sys.stderr)
        return 1
    dev = torch.device('cuda:0')
    if args.leg is not None:
        legs = {'plugin_step': _leg_plugin_step, 'c2_sizes': _leg_c2_sizes,
                'shard': _leg_shard,
                'galaxy_worker': _leg_galaxy_worker,
                'batched_forward': _leg_batched_forward, 'batched_shard': _leg_batched_shard,
                'batched_fwd_adjoint': _leg_batched_adjoint,
                'single_fwd_adjoint': _leg_single_adjoint, 'orbit_fit': _leg_orbit_fit}
        if args.batched_schedule is not None:
            if args.leg != 'batched_forward':
                raise SystemExit('--batched-schedule applies to --leg batched_forward only')
            out = _leg_batched_forward(dev, schedule=_parse_schedule(args.batched_schedule))
        elif args.leg == 'shard' and args.cfg:
            out = _leg_shard(dev, eps=args.eps,
                             candidates=[tuple(int(x) for x in args.cfg.split(','))])
        else:
            out = legs[args.leg](dev)
        print(json.dumps({args.leg: out}), flush=True)
        return 0
    if args.calibrate_copy > 0:
        words = args.calibrate_copy // 4
        src = torch.ones(words, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    n = args.n
    cfg = tuple(int(x) for x in args.cfg.split(',')) if args.cfg else None
    r, v, m = S.plummer(n, seed=0)

    def run(form, cfg):
        """K timed steps (after W warm-up steps) of stage 1 in `form` + FP64-master stage 2,
        on freshly initialised device state; HIP events around each stage on its stream."""
        posm = torch.empty((n, 4), dtype=torch.float32)
        posm[:, :3] = torch.from_numpy(r)
        posm[:, 3] = torch.from_numpy(m)
        posm = posm.to(dev)
        r64 = torch.from_numpy(np.ascontiguousarray(r.T)).to(dev)
        v64 = torch.from_numpy(np.ascontiguousarray(v.T)).to(dev)
        acc = torch.zeros((3, n), dtype=torch.float32, device=dev)

        def step(ev=None):
            if ev is not None:
                ev[0].record()
            if form == 'symmetric':
                stages.stage1_sym_f32(posm, acc, 1.0, args.eps, cfg)
            elif form == 'split':
                stages.stage1_split_f32(posm, acc, 1.0, args.eps, cfg)
            else:
                stages.stage1_f32(posm, acc, 1.0, args.eps, 'kahan', cfg)
            if ev is not None:
                ev[1].record()
            stages.stage2_f64m(posm, r64, v64, acc, args.dt)
            if ev is not None:
                ev[2].record()

        for _ in range(args.warmup):
            step()
        events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)]
                  for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step(events[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        s1 = np.array([e[0].elapsed_time(e[1]) for e in events]) * 1e-3   # s
        s2 = np.array([e[1].elapsed_time(e[2]) for e in events]) * 1e-3
        stages.release_workspaces()
        return {'form': form, 'cfg': cfg, 'wall': wall, 's1': s1, 's2': s2, 'posm': posm}

    forms = ['owner', 'split', 'symmetric'] if args.stage1 == 'auto' else [args.stage1]
    runs = [run(f, cfg if args.stage1 != 'auto' else None) for f in forms]
    best = min(runs, key=lambda x: x['wall'])
    posm = best['posm']
    s1, s2, wall = best['s1'], best['s2'], best['wall']
    s1_avg = float(s1.mean())
    inter = float(n) * float(n)
    value = inter * args.steps / wall / 1e9
    achieved = FLOP_PER_INTERACTION * inter / s1_avg / 1e12
    traffic = _traffic_from_profiles({'symmetric': 'stage1_sym_plummer',
                                      'owner': 'stage1_f32_plummer'}.get(best['form'], 'none'),
                                     n)
    kernel = {'symmetric': 'stage 1, symmetric form: k_stage1_f32 (block diagonal) + '
                           'k_sym_segments + k_sym_fold, timed together',
              'split': 'stage 1, j-split form: k_stage1_f32 (j-split) + k_split_reduce_f32, '
                       'timed together',
              'owner': 'k_stage1_f32'}[best['form']]

    line = {
        'metric': 'G pair-interactions/s and % FP32 peak, N=2^18; batched fwd+adjoint '
                  'steps/s vs plain',
        'value': value,
        'unit': 'Ginteractions/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded Plummer sphere)',
        'config': {'workload': 'single universe, Plummer N=%d, FP32-compensated stage 1 + '
                               'FP64-master stage 2, G=1, eps=%g, dt=%g'
                               % (n, args.eps, args.dt),
                   'n_bodies': n, 'eps': args.eps, 'dt': args.dt,
                   'stage1_form': best['form'],
                   'stage1_cfg': list(best['cfg']) if best['cfg'] else 'default'},
        'roofline': {'bound': 'valu', 'achieved': achieved, 'peak': PEAK_FP32_TFLOPS,
                     'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP32_TFLOPS,
                     'traffic': traffic,
                     'kernel': kernel, 'kernel_ms_avg': s1_avg * 1e3,
                     'flop_per_interaction': FLOP_PER_INTERACTION},
        'stage1_ms_min': float(s1.min()) * 1e3,
        'stage1_interactions_per_s': inter / s1_avg,
        'pct_fp32_peak': 100.0 * achieved / PEAK_FP32_TFLOPS,
        'stage1_forms': {x['form']: {'ms_per_step': x['wall'] / args.steps * 1e3,
                                     'stage1_ms_avg': float(x['s1'].mean()) * 1e3,
                                     'interactions_per_s': inter / float(x['s1'].mean())}
                         for x in runs},
    }
    if args.only_headline:
        print(json.dumps(line), flush=True)
        return 0
    line.update({
        'stage2': {'kernel': 'k_stage2_f64m', 'ms_avg': float(s2.mean()) * 1e3,
                   'GB_per_s': STAGE2_BYTES_F64M * n / float(s2.mean()) / 1e9,
                   'bytes_per_body': STAGE2_BYTES_F64M,
                   'c3_body_count': _stage2_large(dev)},
        'stage1_f64': _leg_stage1_f64(dev, r, m, args.eps),
        'plugin_step': _leg_plugin_step(dev),
        'galaxy_worker': _leg_galaxy_worker(dev),
        'c2_sizes': _leg_c2_sizes(dev, args.eps),
        'shard': _leg_shard(dev, eps=args.eps),
    })
    if not args.skip_batched:
        line['batched_forward'] = _leg_batched_forward(dev)
        line['batched_shard'] = _leg_batched_shard(dev)
        line['batched_fwd_adjoint'] = _leg_batched_adjoint(dev)
        line['single_fwd_adjoint'] = _leg_single_adjoint(dev)
        if not args.skip_orbit_fit:
            line['orbit_fit'] = _leg_orbit_fit(dev)
    if not args.no_cpu:
        pos32 = posm[:, :3].double().cpu().numpy()
        m32 = posm[:, 3].double().cpu().numpy()
        line['cpu_baseline'] = _cpu_baseline(pos32, m32, args.cpu_quick)
    else:
        line['cpu_baseline'] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
