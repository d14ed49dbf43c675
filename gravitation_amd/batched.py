"""Batched universes on one device: many independent universes of uneven size advanced
together, next to the plain per-universe loop of single-universe launches they replace.

The reference holds one universe per kernel object (src/gravitation/kernel/_base_.py:64-93);
its physics per universe is stage 1 (py1.py:57-64, pc2.py:60-90) and stage 2
(_base_.py:58-62).  Here every universe gets that physics with FP32 compensated pair terms
and FP64 position/velocity masters, by one of two schedules of the same plan:

* ``fused``: whole universes bin-packed into workgroups, positions in LDS, all steps of a
  call in ONE launch (``grav_batched_step_f32`` / ``grav_batched_adjoint_f32``); best when
  the universes are small and many;
* ``perstep``: one launch per stage per step over every universe's i-tiles, largest first,
  so a big universe is spread over the whole GPU (``grav_batched_stage1_f32`` +
  ``grav_stage2_f64m``; ``grav_batched_adjoint_mu`` + ``grav_batched_stage1_vjp_f32``);
  best when a few large universes dominate the cost;
* ``perstep-sym``: ``perstep`` with stage 1 in the symmetric form (each pair of bodies of a
  universe once, in blocks of 64*K*W bodies: ``grav_batched_stage1_sym_f32``), which halves
  the pair work of the universes much larger than one block; the adjoint is ``perstep``'s.

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
             'max_bin_bodies', 'max_bin_cost', 'total_cost', 'max_universe', 'n_items',
             'fused_ok')

# (mode, waves per group, i_per_lane, tiles_per_wave, groups (0 = auto), j_waves), and for
# 'perstep-sym' a 7th entry (K, W): symmetric blocks of 64*K*W bodies
DEFAULT_CONFIG = ('fused', 8, 1, 8, 0, 4)
MODES = ('fused', 'perstep', 'perstep-sym')
SYM_SHAPES = ((4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (8, 4))
DEFAULT_SYM = (4, 2)


def shard_universes(offsets, shares):
    """The universes of each of ``shares`` GPUs when a batch is divided among GPUs by whole
    universes: longest processing time first on the pair count N_u^2 (larger universes
    first, ties to the lower index; each to the GPU with the least work so far, ties to the
    lower rank), every share's list ascending.  Universes are independent, so the shares
    need no communication, and the heaviest share holds at most the mean plus one universe's
    pair count.  Returns a list of ``shares`` int64 arrays (some empty when there are fewer
    universes than shares)."""
    import heapq
    offs = np.ascontiguousarray(offsets, dtype=np.int64)
    if offs.ndim != 1 or offs.shape[0] < 1 or offs[0] != 0 or np.any(np.diff(offs) < 0):
        raise ValueError('offsets must be a non-decreasing 1-D array starting at 0')
    shares = int(shares)
    if shares < 1:
        raise ValueError('shares must be >= 1 (got %d)' % shares)
    sizes = np.diff(offs)
    cost = sizes * sizes
    order = np.lexsort((np.arange(sizes.shape[0]), -cost))
    heap = [(0, p) for p in range(shares)]
    bins = [[] for _ in range(shares)]
    for u in order:
        load, p = heapq.heappop(heap)
        bins[p].append(int(u))
        heapq.heappush(heap, (load + int(cost[u]), p))
    return [np.array(sorted(b), dtype=np.int64) for b in bins]


def take_universes(r, v, m, offsets, universes):
    """(r, v, m, offsets) of the listed universes, in the listed order: host arrays of one
    share of a batch, the same layout as the whole (universe k = bodies offsets[k]:offsets[k+1]
    of the result)."""
    offs = np.ascontiguousarray(offsets, dtype=np.int64)
    us = np.ascontiguousarray(universes, dtype=np.int64)
    if us.size and (us.min() < 0 or us.max() >= offs.shape[0] - 1):
        raise ValueError('universe index out of range')
    sizes = offs[us + 1] - offs[us]
    new = np.zeros(us.shape[0] + 1, dtype=np.int64)
    np.cumsum(sizes, out=new[1:])
    idx = (np.concatenate([np.arange(offs[u], offs[u + 1]) for u in us]) if us.size
           else np.zeros(0, dtype=np.int64))
    return (np.asarray(r)[idx], np.asarray(v)[idx], np.asarray(m)[idx], new)


class BatchedPlan:
    """Owner of a ``grav_batched_plan`` (schedule of universes onto workgroups/wavefronts and
    the per-step tile list)."""

    def __init__(self, offsets, waves=8, i_per_lane=1, tiles_per_wave=8, groups=0,
                 device=None):
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        if offs.ndim != 1 or offs.shape[0] < 2:
            raise ValueError('offsets must be a 1-D array of n_universes + 1 entries')
        self._offs = offs
        # the plan's schedule lives on this device; the library rejects it elsewhere
        self.device = (torch.device('cuda', torch.cuda.current_device()) if device is None
                       else torch.device(device))
        if self.device.index is None:
            self.device = torch.device('cuda', torch.cuda.current_device())
        self._h = None
        h = ctypes.c_void_p()
        with stages.on_device(self.device):
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

    @property
    def n_bodies(self):
        return int(self._offs[-1])

    @property
    def offsets(self):
        """Host int64 array of n_universes + 1 body offsets (a copy)."""
        return self._offs.copy()

    def info(self):
        buf = (ctypes.c_int64 * len(INFO_KEYS))()
        _native.call('grav_batched_plan_info', self.handle, buf, len(INFO_KEYS))
        return dict(zip(INFO_KEYS, [int(x) for x in buf]))

    def close(self):
        if getattr(self, '_h', None) is not None:
            with stages.on_device(self.device):
                _native.load().grav_batched_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream(dev):
    return torch.cuda.current_stream(dev).cuda_stream


def _check_mode(mode, sym):
    if mode not in MODES:
        raise ValueError('mode must be one of %s (got %r)' % (MODES, mode))
    if mode == 'perstep-sym' and tuple(sym) not in SYM_SHAPES:
        raise ValueError('symmetric shape (K, W) must be one of %s (got %r)'
                         % (SYM_SHAPES, sym))


class Engine:
    """A plan plus a schedule ('fused', 'perstep' or 'perstep-sym'); advances and
    differentiates the concatenated state of all universes on the current stream."""

    def __init__(self, offsets, config=DEFAULT_CONFIG, device=None):
        if len(config) not in (6, 7):
            raise ValueError('config is (mode, waves, i_per_lane, tiles_per_wave, groups, '
                             'j_waves[, (K, W)])')
        mode, waves, kt, smax, groups, jw = config[:6]
        sym = tuple(int(x) for x in config[6]) if len(config) == 7 else DEFAULT_SYM
        _check_mode(mode, sym)
        self.plan = BatchedPlan(offsets, waves, kt, smax, groups, device)
        self._init(mode, jw, sym)

    @classmethod
    def wrap(cls, plan, mode='fused', j_waves=4, sym=DEFAULT_SYM):
        """An engine over an existing BatchedPlan."""
        sym = tuple(int(x) for x in sym)
        _check_mode(mode, sym)
        e = cls.__new__(cls)
        e.plan = plan
        e._init(mode, j_waves, sym)
        return e

    def _init(self, mode, jw, sym):
        self.mode = mode
        self.j_waves = int(jw)
        self.config = (mode,) + self.plan.config + (int(jw),)
        self._acc = None
        self._mu4 = None
        self._ws = None
        if mode == 'perstep-sym':
            self.config += (sym,)
            nbytes = ctypes.c_int64()
            with stages.on_device(self.plan.device):
                # block lists (host work and uploads) and the partials' workspace, once
                _native.call('grav_batched_sym_prepare', self.plan.handle, sym[0], sym[1])
                _native.call('grav_batched_sym_workspace', self.plan.handle,
                             ctypes.byref(nbytes))
                self._ws = torch.empty((max(nbytes.value, 1),), dtype=torch.uint8,
                                       device=self.plan.device)

    @property
    def n(self):
        return self.plan.n_bodies

    def _device(self, *tensors):
        """Checks that the tensors of a call live on the plan's device; returns the context
        that makes it current."""
        for t in tensors:
            if t is not None and t.device != self.plan.device:
                raise ValueError('tensor on %s, the plan was made on %s'
                                 % (t.device, self.plan.device))
        return stages.on_device(self.plan.device)

    def _buf(self, name, shape, dtype, dev):
        b = getattr(self, name)
        if b is None or b.device != dev or tuple(b.shape) != shape:
            b = torch.empty(shape, dtype=dtype, device=dev)
            setattr(self, name, b)
        return b

    def forward(self, posm, r64, v64, G, eps, dt, steps, traj=None):
        """``steps`` steps of every universe; traj (steps, n, 4) float32 or None receives posm
        at the start of every step."""
        dev = posm.device
        with self._device(posm, r64, v64, traj):
            if self.mode == 'fused':
                _native.call('grav_batched_step_f32', self.plan.handle, posm.data_ptr(),
                             r64.data_ptr(), v64.data_ptr(), G, eps, dt, int(steps),
                             None if traj is None else traj.data_ptr(), _stream(dev))
                return
            acc = self._buf('_acc', (3, self.n), torch.float32, dev)
            for t in range(steps):
                if traj is not None:
                    traj[t].copy_(posm)
                if self._ws is not None:
                    _native.call('grav_batched_stage1_sym_f32', self.plan.handle,
                                 posm.data_ptr(), acc.data_ptr(), G, eps, self._ws.data_ptr(),
                                 self._ws.numel(), _stream(dev))
                else:
                    _native.call('grav_batched_stage1_f32', self.plan.handle, posm.data_ptr(),
                                 acc.data_ptr(), G, eps, self.j_waves, _stream(dev))
                stages.stage2_f64m(posm, r64, v64, acc, dt)

    def adjoint(self, traj, rho, nu, gamma, G, eps, dt, steps):
        """``steps`` backward steps (t = steps-1 .. 0) at the positions traj[t]."""
        dev = rho.device
        with self._device(traj, rho, nu, gamma):
            if self.mode == 'fused':
                _native.call('grav_batched_adjoint_f32', self.plan.handle, traj.data_ptr(),
                             rho.data_ptr(), nu.data_ptr(), gamma.data_ptr(), G, eps, dt,
                             int(steps), _stream(dev))
                return
            mu4 = self._buf('_mu4', (self.n, 4), torch.float32, dev)
            for t in range(steps - 1, -1, -1):
                _native.call('grav_batched_adjoint_mu', self.plan.handle, rho.data_ptr(),
                             nu.data_ptr(), mu4.data_ptr(), dt, _stream(dev))
                _native.call('grav_batched_stage1_vjp_f32', self.plan.handle,
                             traj[t].data_ptr(), mu4.data_ptr(), rho.data_ptr(),
                             gamma.data_ptr(), G, eps, dt, self.j_waves, _stream(dev))


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
        self.device = self.posm.device   # with its index
        self.r64 = torch.from_numpy(np.ascontiguousarray(r.T)).to(self.device)
        self.v64 = torch.from_numpy(np.ascontiguousarray(v.T)).to(self.device)
        self.set_config(config)

    @property
    def n_universes(self):
        return self.offsets.shape[0] - 1

    @property
    def plan(self):
        return self.engine.plan

    def set_config(self, config):
        if len(config) == 4:   # plan shape only: the fused schedule
            config = ('fused',) + tuple(config) + (4,)
        self.engine = Engine(self.offsets, config, self.device)
        self.config = self.engine.config

    def step(self, steps=1, traj=None):
        """Advances every universe ``steps`` steps on the current stream.  ``traj``
        (optional float32 tensor (steps, n, 4)) receives posm at each step start."""
        if traj is not None:
            if (traj.dtype != torch.float32 or tuple(traj.shape) != (steps, self.n, 4)
                    or not traj.is_contiguous() or traj.device != self.device):
                raise ValueError('traj must be a contiguous float32 (%d, %d, 4) tensor on %s'
                                 % (steps, self.n, self.device))
        with torch.cuda.device(self.device):
            self.engine.forward(self.posm, self.r64, self.v64, self.G, self.eps, self.dt,
                                steps, traj)

    def capture(self, steps):
        """A hipGraph of ``step(steps)`` (replay with ``.replay()``)."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            if self.engine.mode != 'fused':
                self.engine._buf('_acc', (3, self.n), torch.float32, self.device)
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
    baseline the batched schedules are compared with, kept as lean as a host loop gets: the
    C ABI is called directly with pointers prepared once (every status still checked)."""

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
        self._args = [(st['posm'].data_ptr(), st['acc'].data_ptr(), st['r64'].data_ptr(),
                       st['v64'].data_ptr(), st['posm'].shape[0]) for st in self.universes]

    def step(self, steps=1):
        lib = _native.load()
        s1, s2 = lib.grav_stage1_f32, lib.grav_stage2_f64m
        G, eps, dt, kahan = self.G, self.eps, self.dt, _native.ACCUM_KAHAN
        with stages.on_device(self.device):
            stream = _stream(self.device)
            for _ in range(steps):
                for p, acc, r, v, n in self._args:
                    rc = s1(p, acc, n, G, eps, kahan, stream)
                    if rc:
                        _native.check('grav_stage1_f32', rc)
                    rc = s2(p, r, v, acc, n, dt, stream)
                    if rc:
                        _native.check('grav_stage2_f64m', rc)

    def state(self):
        torch.cuda.synchronize(self.device)
        r = np.concatenate([st['r64'].cpu().numpy().T for st in self.universes])
        v = np.concatenate([st['v64'].cpu().numpy().T for st in self.universes])
        return r, v
