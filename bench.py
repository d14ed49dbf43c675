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
                        "(symmetric); needs --stage1")
    p.add_argument('--stage1', choices=('auto', 'owner', 'split', 'symmetric'), default='auto',
                   help='stage-1 form; auto times both and reports the faster')
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--leg', choices=('plugin_step', 'c2_sizes', 'batched_forward',
                                     'batched_fwd_adjoint',
                                     'single_fwd_adjoint', 'orbit_fit'), default=None,
                   help='run only this leg and print its JSON (for profiling it alone)')
    p.add_argument('--skip-batched', action='store_true',
                   help='only the single-universe leg')
    p.add_argument('--skip-orbit-fit', action='store_true',
                   help='skip the 50-iteration batched orbit-fitting leg')
    p.add_argument('--cpu-budget', type=float, default=15.0,
                   help='seconds of CPU work for the CPU baseline sample')
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


def _cpu_baseline(pos, m, budget):
    """The reference's C stage 1 (oracle/_ref/lib4.so, SSE + OpenMP) timed on this host; the
    oracle's np1 restatement on a row sample when the reference build is absent."""
    import gc
    from oracle import oracle as O
    from oracle import ref as R
    n_full = pos.shape[0]
    cores = len(os.sched_getaffinity(0))
    threads = max(2, min(16, cores))
    try:
        import psutil
        phys = psutil.cpu_count(logical=False)
        logical = psutil.cpu_count(logical=True)
    except Exception:
        phys = logical = os.cpu_count()
    if R.available('lib4'):
        # Calibrate at N = 2^14, then run one full step at the largest N <= n_full (a power of
        # two) whose predicted time fits the budget.
        probe = 1 << 14
        lib = R.RefLib4(pos[:probe], m[:probe], 1.0, threads)
        lib.accel()
        t0 = time.perf_counter()
        lib.accel()
        rate = probe * probe / (time.perf_counter() - t0)
        n = probe
        while n * 2 <= n_full and (2 * n) ** 2 / rate <= budget:
            n *= 2
        lib = R.RefLib4(pos[:n], m[:n], 1.0, threads)
        gc.collect()
        gc.disable()
        t0 = time.perf_counter()
        lib.accel()
        dt = time.perf_counter() - t0
        gc.enable()
        return {'value': n * n / dt / 1e9, 'unit': 'Ginteractions/s', 'cores': threads,
                'kind': 'reference',
                'sample': 'one full stage-1 step of src/gravitation/kernel/_lib4_/lib.c '
                          '(SSE + OpenMP, built into oracle/_ref) on the first %d bodies of '
                          'the workload, %d OpenMP threads, %.2f s; N^2 interactions/step'
                          % (n, threads, dt),
                'host_cpus': {'physical': phys, 'logical': logical, 'affinity': cores}}
    # Fallback: np1 restatement on strided rows, scaled by the exact pair count.
    rows = np.arange(0, n_full - 1, max(1, (n_full - 1) // 64))
    t0 = time.perf_counter()
    O.np1_rows(pos, m, 1.0, rows)
    dt = time.perf_counter() - t0
    pairs = float(np.sum(n_full - 1 - rows))
    rate_pairs = pairs / dt
    return {'value': 2 * rate_pairs / 1e9, 'unit': 'Ginteractions/s', 'cores': 1,
            'kind': 'port',
            'sample': 'oracle np1 restatement (src/gravitation/kernel/np1.py:81-108), %d '
                      'strided rows of N=%d, %.2f s; extrapolated by pair count'
                      % (rows.shape[0], n_full, dt),
            'host_cpus': {'physical': phys, 'logical': logical, 'affinity': cores}}


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


def _schedule(cfg):
    d = dict(zip(('mode', 'waves', 'i_per_lane', 'tiles_per_wave', 'groups', 'j_waves'), cfg))
    if len(cfg) > 6:
        d['sym_block'] = list(cfg[6])
    return d


def _leg_batched_forward(dev, steps=100, seed=0):
    """4096 universes, N log-uniform in [32, 2048], uniform-cube positions, Gaussian
    velocities (sigma 0.1), log-uniform masses [1e-4, 1e-2], G=1, eps=1e-3, dt=1e-3: all
    universes and `steps` steps in one hipGraph-captured fused launch, against the plain loop
    of single-universe launches (2 per universe-step)."""
    from gravitation_amd import batched, tune
    from gravitation_amd.lib import scenarios as S
    r, v, m, off = S.batched(4096, 32, 2048, seed)
    cost = float(np.sum(np.diff(off).astype(np.float64) ** 2))
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=dev)
    best, table = tune.tune_batched(b, steps=4)
    b.set_config(best)
    g = b.capture(steps)
    g.replay()
    t = _sync_time(g.replay)
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
            'plan': b.plan.info()}


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
    segment's forward, so a differentiated step costs three all-pairs passes: forward,
    recompute and VJP."""
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
            'pair_passes_per_step': 3,
            'pct_fp32_peak_informational': 100.0 * (2 * FLOP_PER_INTERACTION
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
        n = len(u)
        form = u._form
        u.stop()
        t_min, t_med = float(np.min(times)), float(np.median(times))
        out[name] = {'n_bodies': n, 'stage1_form': form, 'step_ms_min': t_min * 1e3,
                     'step_ms_median': t_med * 1e3,
                     'interactions_per_s_at_min': float(n) * n / t_min}
    return out


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
            'universe_steps_per_s': 512 * steps * iters / t}


def main():
    args = _parse()
    if args.gpus != 1:
        print('bench.py: this build measures one GPU only (the gravitation hot path is a '
              'single-device kernel here); --gpus %d is not supported' % args.gpus,
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
                'batched_forward': _leg_batched_forward,
                'batched_fwd_adjoint': _leg_batched_adjoint,
                'single_fwd_adjoint': _leg_single_adjoint, 'orbit_fit': _leg_orbit_fit}
        print(json.dumps({args.leg: legs[args.leg](dev)}), flush=True)
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
                           'k_sym_pairs + k_sym_reduce, timed together',
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
        'stage2': {'kernel': 'k_stage2_f64m', 'ms_avg': float(s2.mean()) * 1e3,
                   'GB_per_s': STAGE2_BYTES_F64M * n / float(s2.mean()) / 1e9,
                   'bytes_per_body': STAGE2_BYTES_F64M,
                   'c3_body_count': _stage2_large(dev)},
        'stage1_f64': _leg_stage1_f64(dev, r, m, args.eps),
        'plugin_step': _leg_plugin_step(dev),
        'c2_sizes': _leg_c2_sizes(dev, args.eps),
    }
    if not args.skip_batched:
        line['batched_forward'] = _leg_batched_forward(dev)
        line['batched_fwd_adjoint'] = _leg_batched_adjoint(dev)
        line['single_fwd_adjoint'] = _leg_single_adjoint(dev)
        if not args.skip_orbit_fit:
            line['orbit_fit'] = _leg_orbit_fit(dev)
    if not args.no_cpu:
        pos32 = posm[:, :3].double().cpu().numpy()
        m32 = posm[:, 3].double().cpu().numpy()
        line['cpu_baseline'] = _cpu_baseline(pos32, m32, args.cpu_budget)
    else:
        line['cpu_baseline'] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
