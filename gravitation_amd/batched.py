"""Batched universes on one device: many independent universes of uneven size advanced by
one fused kernel launch per call (``grav_batched_step_f32``), optionally captured in a
hipGraph, next to the plain per-universe loop of single-universe launches it replaces.

The reference holds one universe per kernel object (src/gravitation/kernel/_base_.py:64-93);
its physics per universe is stage 1 (py1.py:57-64, pc2.py:60-90) and stage 2
(_base_.py:58-62).  Here every universe gets that physics with FP32 compensated pair terms
and FP64 position/velocity masters.

State layout (concatenated over universes, universe u = bodies offsets[u]:offsets[u+1]):
``posm`` float32 (n, 4) = x, y, z, m (the FP32 copy stage 1 reads), ``r64`` and ``v64``
float64 (3, n).
"""

import ctypes

import numpy as np
import torch

from . import _native, stages

_i64p = ctypes.POINTER(ctypes.c_int64)

INFO_KEYS = ('n_bodies', 'n_universes', 'groups', 'waves', 'i_per_lane', 'tiles_per_wave',
             'max_bin_bodies', 'max_bin_cost', 'total_cost', 'max_universe')

DEFAULT_CONFIG = (8, 1, 8, 0)   # waves per group, i_per_lane, tiles_per_wave, groups (0=auto)


class BatchedPlan:
    """Owner of a ``grav_batched_plan`` (schedule of universes onto workgroups/wavefronts)."""

    def __init__(self, offsets, waves=8, i_per_lane=1, tiles_per_wave=8, groups=0):
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        if offs.ndim != 1 or offs.shape[0] < 2:
            raise ValueError('offsets must be a 1-D array of n_universes + 1 entries')
        self._offs = offs
        h = ctypes.c_void_p()
        _native.call('grav_batched_plan_create', offs.ctypes.data_as(_i64p),
                     offs.shape[0] - 1, int(waves), int(i_per_lane), int(tiles_per_wave),
                     int(groups), ctypes.byref(h))
        self._h = h
        self.config = (int(waves), int(i_per_lane), int(tiles_per_wave), int(groups))

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError('plan was closed')
        return self._h

    def info(self):
        buf = (ctypes.c_int64 * len(INFO_KEYS))()
        _native.call('grav_batched_plan_info', self.handle, buf, len(INFO_KEYS))
        return dict(zip(INFO_KEYS, [int(x) for x in buf]))

    def close(self):
        if getattr(self, '_h', None) is not None:
            _native.load().grav_batched_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A batch of universes resident on one device."""

    def __init__(self, r, v, m, offsets, G=1.0, eps=0.0, dt=1e-3, device='cuda',
                 config=DEFAULT_CONFIG):
        self.device = torch.device(device)
        r = np.asarray(r, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        m = np.asarray(m, dtype=np.float64)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n = r.shape[0]
        if r.shape != (n, 3) or v.shape != (n, 3) or m.shape != (n,):
            raise ValueError('r, v must be (n, 3) and m (n,)')
        if self.offsets[0] != 0 or self.offsets[-1] != n:
            raise ValueError('offsets must start at 0 and end at n')
        self.n = n
        self.G, self.eps, self.dt = float(G), float(eps), float(dt)
        posm = torch.empty((n, 4), dtype=torch.float32)
        posm[:, :3] = torch.from_numpy(r)
        posm[:, 3] = torch.from_numpy(m)
        self.posm = posm.to(self.device)
        self.r64 = torch.from_numpy(np.ascontiguousarray(r.T)).to(self.device)
        self.v64 = torch.from_numpy(np.ascontiguousarray(v.T)).to(self.device)
        self.set_config(config)

    @property
    def n_universes(self):
        return self.offsets.shape[0] - 1

    def set_config(self, config):
        self.plan = BatchedPlan(self.offsets, *config)
        self.config = self.plan.config

    def step(self, steps=1, traj=None):
        """Advances every universe ``steps`` steps in one launch on the current stream.
        ``traj`` (optional float32 tensor (steps, n, 4)) receives posm at each step start."""
        tp = None
        if traj is not None:
            if (traj.dtype != torch.float32 or tuple(traj.shape) != (steps, self.n, 4)
                    or not traj.is_contiguous() or traj.device != self.device):
                raise ValueError('traj must be a contiguous float32 (%d, %d, 4) tensor on %s'
                                 % (steps, self.n, self.device))
            tp = traj.data_ptr()
        with torch.cuda.device(self.device):
            _native.call('grav_batched_step_f32', self.plan.handle, self.posm.data_ptr(),
                         self.r64.data_ptr(), self.v64.data_ptr(), self.G, self.eps, self.dt,
                         int(steps), tp, torch.cuda.current_stream(self.device).cuda_stream)

    def capture(self, steps):
        """A hipGraph of one ``step(steps)`` launch (replay with ``.replay()``)."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.step(steps)
        torch.cuda.current_stream(self.device).wait_stream(s)
        return g

    def state(self):
        """Host copies (r (n,3), v (n,3)) of the FP64 masters."""
        torch.cuda.synchronize(self.device)
        return self.r64.cpu().numpy().T.copy(), self.v64.cpu().numpy().T.copy()

    def plain(self):
        """The plain per-universe loop on a copy of the current state."""
        return PlainLoop(self)


class PlainLoop:
    """One universe at a time with the single-universe kernels: for each step and each
    universe, grav_stage1_f32 then grav_stage2_f64m (2 launches per universe-step).  The
    baseline the batched launch is compared with."""

    def __init__(self, batch):
        self.G, self.eps, self.dt = batch.G, batch.eps, batch.dt
        self.device = batch.device
        self.universes = []
        for u in range(batch.n_universes):
            a, b = int(batch.offsets[u]), int(batch.offsets[u + 1])
            nu = b - a
            self.universes.append({
                'posm': batch.posm[a:b].clone(),
                'r64': batch.r64[:, a:b].contiguous(),
                'v64': batch.v64[:, a:b].contiguous(),
                'acc': torch.zeros((3, nu), dtype=torch.float32, device=self.device),
            })

    def step(self, steps=1):
        for _ in range(steps):
            for st in self.universes:
                stages.stage1_f32(st['posm'], st['acc'], self.G, self.eps, 'kahan')
                stages.stage2_f64m(st['posm'], st['r64'], st['v64'], st['acc'], self.dt)

    def state(self):
        torch.cuda.synchronize(self.device)
        r = np.concatenate([st['r64'].cpu().numpy().T for st in self.universes])
        v = np.concatenate([st['v64'].cpu().numpy().T for st in self.universes])
        return r, v
