"""Autotuner for the launch shapes of the stage-1 and batched kernels.

Each tuner times a list of candidate shapes on the caller's device with HIP events, on
scratch copies of the state (the caller's state is not advanced), and returns the fastest
shape plus the full timing table.  Results are cached per process, keyed by the workload
signature, and can be persisted to / loaded from a JSON file.
"""

import json
import os

import numpy as np
import torch

from . import _native, batched, stages

_cache = {}

STAGE1_CANDIDATES = [(k, s, jb) for k in (1, 2, 4, 8) for s in (1, 2, 4, 8)
                     for jb in (256,)]

# (mode, waves, i_per_lane, tiles_per_wave, groups, j_waves[, (K, W)]); the per-step
# schedule ignores the workgroup-bin shape, so it only varies i_per_lane and j_waves, and
# the symmetric per-step forward uses neither, only its block shape (K, W).
BATCHED_CANDIDATES = (
    [('fused', w, kt, sm, g, 4) for w in (4, 8)
     for kt, sm in ((1, 8), (1, 4), (2, 4), (2, 2), (4, 2))
     for g in (256, 512, 1024, 2048)]
    + [('perstep', 8, kt, 8 // kt, 0, jw) for kt in (1, 2, 4) for jw in (1, 2, 4, 8)]
    + [('perstep-sym', 8, 1, 8, 0, 4, sym) for sym in batched.SYM_SHAPES])


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e-3)
    return float(min(ts))


def _size_bucket(n):
    return int(2 ** round(np.log2(max(1, n))))


def tune_stage1(posm, G=1.0, eps=1e-3, candidates=None, reps=3):
    """Fastest (i_per_lane, j_waves, j_block) for FP32 stage 1 on ``posm`` (n, 4)."""
    n = posm.shape[0]
    key = ('stage1', _size_bucket(n), eps == 0.0)
    if key in _cache and candidates is None:
        return _cache[key]
    acc = torch.empty((3, n), dtype=torch.float32, device=posm.device)
    table = []
    for cfg in candidates or STAGE1_CANDIDATES:
        t = _time(lambda: stages.stage1_f32(posm, acc, G, eps, 'kahan', cfg), reps)
        table.append((cfg, t))
    best = min(table, key=lambda x: x[1])[0]
    _cache[key] = (best, table)
    return best, table


SYM_CANDIDATES = [(4, 2), (4, 4), (8, 2), (8, 4)]
SYM_F64_CANDIDATES = [(2, 2), (2, 4), (4, 2), (4, 4)]
OWNER_FORM_CANDIDATES = [(1, 8, 256), (2, 8, 256), (2, 16, 256), (4, 4, 256), (8, 2, 256)]
SPLIT_CANDIDATES = [(1, 8, 4), (1, 8, 8), (1, 4, 8), (2, 8, 8), (2, 8, 16), (1, 8, 16)]
SPLIT_F64_CANDIDATES = [(1, 8, 16), (1, 8, 32), (2, 8, 16), (1, 4, 32)]


def tune_stage1_form(posm, G=1.0, eps=1e-3, reps=3, free_bytes=None):
    """Fastest stage-1 (form, cfg) for ``posm`` (n, 4), float32 or float64: owner-computes
    shapes, j-split shapes and the symmetric shapes (each with the workspace
    ``stages.sym_workspace_size`` gives), timed on the device.  Returns ((form, cfg), table);
    cfg None = the library's shape.  The result is cached per (precision, n, eps == 0) for
    the process."""
    n = posm.shape[0]
    f64 = posm.dtype == torch.float64
    key = ('stage1_form', 'f64' if f64 else 'f32', n, eps == 0.0)
    if key in _cache:
        return _cache[key]
    if free_bytes is None:
        free_bytes, _ = torch.cuda.mem_get_info(posm.device)
    acc = torch.empty((3, n), dtype=posm.dtype, device=posm.device)
    table = []
    if f64:
        table.append((('owner', None), _time(lambda: stages.stage1_f64(posm, acc, G, eps), reps)))
        for cfg in SPLIT_F64_CANDIDATES:
            ws = stages.split_workspace(posm, n, cfg)
            table.append((('split', cfg),
                          _time(lambda: stages.stage1_split_f64(posm, acc, G, eps, cfg, ws),
                                reps)))
        sym, dtype = SYM_F64_CANDIDATES, 'float64'
        run = stages.stage1_sym_f64
    else:
        for cfg in OWNER_FORM_CANDIDATES:
            table.append((('owner', cfg),
                          _time(lambda: stages.stage1_f32(posm, acc, G, eps, 'kahan', cfg), reps)))
        for cfg in SPLIT_CANDIDATES:
            ws = stages.split_workspace(posm, n, cfg)
            table.append((('split', cfg),
                          _time(lambda: stages.stage1_split_f32(posm, acc, G, eps, cfg, ws),
                                reps)))
        sym, dtype = SYM_CANDIDATES, 'float32'
        run = stages.stage1_sym_f32
    for cfg in sym:
        size = stages.sym_workspace_size(n, cfg, free_bytes, dtype)
        if size > free_bytes // 2:
            continue
        ws = torch.empty((max(size, 256),), dtype=torch.uint8, device=posm.device)
        table.append((('symmetric', cfg), _time(lambda: run(posm, acc, G, eps, cfg, ws), reps)))
        del ws
    best = min(table, key=lambda x: x[1])[0]
    _cache[key] = (best, table)
    return best, table


def _batch_signature(offsets):
    sizes = np.diff(np.asarray(offsets))
    hist = np.bincount(np.minimum(np.log2(sizes).astype(int), 20), minlength=21)
    return ('batched', int(sizes.shape[0]), tuple(int(x) for x in hist))


def tune_batched(batch, steps=4, candidates=None, reps=2):
    """Fastest schedule (mode, waves, i_per_lane, tiles_per_wave, groups, j_waves) for one
    forward call of ``steps`` steps over a Batch; the batch's own state is restored after
    timing."""
    key = _batch_signature(batch.offsets)
    if key in _cache and candidates is None:
        return _cache[key]
    saved = (batch.posm.clone(), batch.r64.clone(), batch.v64.clone(), batch.config)
    table = []
    for cfg in candidates or BATCHED_CANDIDATES:
        try:
            batch.set_config(cfg)
        except Exception:
            continue  # shape cannot hold this batch (capacity), skip it
        try:
            t = _time(lambda: batch.step(steps), reps)
        except _native.GravError:
            continue  # e.g. no fused schedule for this batch
        table.append((batch.config, t / steps))
    batch.posm.copy_(saved[0])
    batch.r64.copy_(saved[1])
    batch.v64.copy_(saved[2])
    batch.set_config(saved[3])
    if not table:
        raise RuntimeError('no candidate launch shape can hold this batch')
    best = min(table, key=lambda x: x[1])[0]
    _cache[key] = (best, table)
    return best, table


ADJOINT_CANDIDATES = (
    [('fused', w, kt, sm, g, 4) for w in (4, 8)
     for kt, sm in ((1, 8), (1, 4), (2, 4), (2, 2))
     for g in (256, 512, 1024, 2048)]
    + [('perstep', 8, kt, 8 // kt, 0, jw) for kt in (1, 2, 4) for jw in (2, 4, 8)]
    + [('perstep-sym', 8, kt, 8 // kt, 0, jw, sym) for kt in (1, 2) for jw in (4, 8)
       for sym in ((4, 1), (4, 2), (8, 1))])


def tune_batched_adjoint(batch, steps=4, candidates=None, reps=2):
    """Fastest schedule for one forward (with trajectory record) plus one adjoint pass of
    ``steps`` steps over a Batch; the batch's state is restored after timing."""
    key = ('adjoint',) + _batch_signature(batch.offsets)[1:]
    if key in _cache and candidates is None:
        return _cache[key]
    n = batch.n
    saved = (batch.posm.clone(), batch.r64.clone(), batch.v64.clone(), batch.config)
    traj = torch.empty((steps, n, 4), dtype=torch.float32, device=batch.device)
    rho = torch.zeros((3, n), dtype=torch.float64, device=batch.device)
    nu = torch.zeros_like(rho)
    gam = torch.zeros((n,), dtype=torch.float64, device=batch.device)
    table = []
    for cfg in candidates or ADJOINT_CANDIDATES:
        try:
            batch.set_config(cfg)
        except Exception:
            continue

        def run():
            batch.step(steps, traj=traj)
            batch.engine.adjoint(traj, rho, nu, gam, batch.G, batch.eps, batch.dt, steps)
        try:
            t = _time(run, reps)
        except _native.GravError:
            continue
        table.append((batch.config, t / steps))
    batch.posm.copy_(saved[0])
    batch.r64.copy_(saved[1])
    batch.v64.copy_(saved[2])
    batch.set_config(saved[3])
    if not table:
        raise RuntimeError('no candidate launch shape can hold this batch')
    best = min(table, key=lambda x: x[1])[0]
    _cache[key] = (best, table)
    return best, table


def save(path):
    data = [{'key': repr(k), 'best': list(v[0]), 'table': [[list(c), t] for c, t in v[1]]}
            for k, v in _cache.items()]
    with open(path, 'w') as f:
        json.dump(data, f, indent=1)


def load(path):
    if not os.path.exists(path):
        return 0
    import ast
    with open(path) as f:
        data = json.load(f)
    for e in data:
        _cache[ast.literal_eval(e['key'])] = (tuple(e['best']),
                                              [(tuple(c), t) for c, t in e['table']])
    return len(data)


__all__ = ['tune_stage1', 'tune_stage1_form', 'tune_batched', 'tune_batched_adjoint', 'save', 'load', 'batched']
