"""Differentiable trajectories: hand-written adjoint kernels behind torch.autograd.Function.

The forward is T time steps of the reference's scheme (stage 1, src/gravitation/kernel/
py1.py:57-64 with optional softening; stage 2, src/gravitation/kernel/_base_.py:58-62) with
FP32 pair terms and FP64 position/velocity masters.  The backward is the discrete adjoint
(per step, with mu = nu + dt*rho: rho += dt J_r^T mu, gamma += dt J_m^T mu, nu = mu), computed
for one universe by ``grav_adjoint_mu`` + ``grav_stage1_vjp_f32`` (its j-split form
``grav_stage1_vjp_f32_split`` from 1024 bodies), and for many universes by
``grav_batched_adjoint_f32`` (fused schedule: one launch per segment) or
``grav_batched_adjoint_mu`` + ``grav_batched_stage1_vjp_f32`` (per-step schedules).

Memory: the backward needs the positions every forward step used.  If T float4 position
sets fit the HBM budget they are recorded during the forward pass; otherwise the forward keeps
FP64 checkpoints (r, v) every C steps and the backward recomputes each segment's positions from
its checkpoint (one extra forward pass in total).  ``checkpoint_interval`` picks C.
"""

import collections
import ctypes
import math

import numpy as np
import torch

from . import _native, batched, stages


MU_TILES_FROM = 2048

# Backward steps in the j-split form run chained (``stage1_vjp_split_chain``: the reduce also
# forms the next step's cotangent, two launches a step instead of four); False runs the
# separate ``adjoint_mu`` + ``stage1_vjp_split_f32`` calls (the same rho, nu and gamma).
CHAIN = True
# The backward pass recomputes segment c - 1 while segment c's adjoint steps run: each
# backward step shares its two launches with one recomputed step (``backward_pair_split``,
# two record buffers).  False runs them one after the other (the same bits: the same
# arithmetic on the same data).
OVERLAP = True


def adjoint_mu(rho, nu, mu4, dt, workspace=None):
    """nu (3, n) <- mu = nu + dt rho (FP64) and mu4 (n, 4) FP32 <- mu * 2^-E with E in w,
    for ONE universe of n bodies.  With a device ``workspace`` (4 bytes per 1024 bodies) and
    n >= MU_TILES_FROM, over many workgroups (``grav_adjoint_mu_ws``, the same bits)."""
    n = rho.shape[1]
    for t, name, dtype, shape in ((rho, 'rho', torch.float64, (3, n)),
                                  (nu, 'nu', torch.float64, (3, n)),
                                  (mu4, 'mu4', torch.float32, (n, 4))):
        stages._check(t, name, dtype, shape)
    stages.same_device(rho, nu, mu4)
    stream = torch.cuda.current_stream(rho.device).cuda_stream
    with stages.on_device(rho):
        if workspace is not None and n >= MU_TILES_FROM:
            stages.same_device(rho, workspace)
            _native.call('grav_adjoint_mu_ws', rho.data_ptr(), nu.data_ptr(), mu4.data_ptr(), n,
                         float(dt), workspace.data_ptr(), workspace.numel(), stream)
        else:
            _native.call('grav_adjoint_mu', rho.data_ptr(), nu.data_ptr(), mu4.data_ptr(), n,
                         float(dt), stream)


def stage1_vjp_f32(posm, mu4, rho, gamma, G, eps, dt):
    """rho += 2^E dt G J_r^T mu and gamma += 2^E dt G J_m^T mu at positions posm (n, 4),
    with the scaled cotangent mu4 of ``adjoint_mu``."""
    n = posm.shape[0]
    for t, name, dtype, shape in ((posm, 'posm', torch.float32, (n, 4)),
                                  (mu4, 'mu4', torch.float32, (n, 4)),
                                  (rho, 'rho', torch.float64, (3, n)),
                                  (gamma, 'gamma', torch.float64, (n,))):
        stages._check(t, name, dtype, shape)
    stages.same_device(posm, mu4, rho, gamma)
    with stages.on_device(posm):
        _native.call('grav_stage1_vjp_f32', posm.data_ptr(), mu4.data_ptr(), rho.data_ptr(),
                     gamma.data_ptr(), n, float(G), float(eps), float(dt),
                     torch.cuda.current_stream(posm.device).cuda_stream)


def stage1_vjp_split_f32(posm, mu4, rho, gamma, G, eps, dt, cfg=None, workspace=None):
    """``stage1_vjp_f32`` in the j-split form (small n); cfg = (i_per_lane, j_waves, splits)
    or None for the library's shape, workspace as for ``stages.stage1_split_f32``."""
    n = posm.shape[0]
    for t, name, dtype, shape in ((posm, 'posm', torch.float32, (n, 4)),
                                  (mu4, 'mu4', torch.float32, (n, 4)),
                                  (rho, 'rho', torch.float64, (3, n)),
                                  (gamma, 'gamma', torch.float64, (n,))):
        stages._check(t, name, dtype, shape)
    stages.same_device(posm, mu4, rho, gamma)
    if workspace is None:
        workspace = stages.split_workspace(posm, n, cfg)
    stages.same_device(posm, workspace)
    k, s, p = cfg if cfg is not None else (0, 0, 0)
    with stages.on_device(posm):
        _native.call('grav_stage1_vjp_f32_split', posm.data_ptr(), mu4.data_ptr(),
                     rho.data_ptr(), gamma.data_ptr(), n, float(G), float(eps), float(dt),
                     int(k), int(s), int(p), workspace.data_ptr(), workspace.numel(),
                     torch.cuda.current_stream(posm.device).cuda_stream)


def vjp_shard_cfg(n, m):
    """The library's (i_per_lane, j_waves, splits) for the VJP of a share of m of n bodies."""
    k, s, p = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _native.call('grav_stage1_vjp_shard_cfg', int(n), int(m), ctypes.byref(k), ctypes.byref(s),
                 ctypes.byref(p))
    return k.value, s.value, p.value


def stage1_vjp_shard_f32(posm, mu4, rho, gamma, i0, G, eps, dt, cfg=None, workspace=None):
    """``stage1_vjp_f32`` for the bodies [i0, i0 + m) of posm (n, 4) only, one GPU's share of
    the rows: rho (3, m) and gamma (m,) are the share's; mu4 (n, 4) holds every body's
    cotangent with the universe's one scale exponent.  With i0 a multiple of 64 * k (256 for
    every shape), bit for bit rows i0 .. i0 + m - 1 of the whole universe's owner-computes VJP
    of shape (k, s) for cfg (k, s, 1), or of ``stage1_vjp_split_f32`` with cfg (k, s, splits >
    1); elsewhere equal to FP32 rounding.  cfg None: ``vjp_shard_cfg``'s; workspace
    (splits > 1): at least ``stages.split_workspace_bytes(m, cfg)`` bytes."""
    n = posm.shape[0]
    m = gamma.shape[0] if gamma.dim() == 1 else -1
    for t, name, dtype, shape in ((posm, 'posm', torch.float32, (n, 4)),
                                  (mu4, 'mu4', torch.float32, (n, 4)),
                                  (rho, 'rho', torch.float64, (3, m)),
                                  (gamma, 'gamma', torch.float64, (m,))):
        stages._check(t, name, dtype, shape)
    stages.same_device(posm, mu4, rho, gamma)
    cfg = tuple(cfg) if cfg is not None else vjp_shard_cfg(n, m)
    ptr, size = 0, 0
    if cfg[2] > 1:
        if workspace is None:
            workspace = stages.split_workspace(posm, m, cfg)
        stages.same_device(posm, workspace)
        ptr, size = workspace.data_ptr(), workspace.numel()
    with stages.on_device(posm):
        _native.call('grav_stage1_vjp_f32_shard', posm.data_ptr(), mu4.data_ptr(),
                     rho.data_ptr(), gamma.data_ptr(), n, int(i0), m, float(G), float(eps),
                     float(dt), int(cfg[0]), int(cfg[1]), int(cfg[2]), ptr, size,
                     torch.cuda.current_stream(posm.device).cuda_stream)


def stage1_vjp_split_chain(posm, mu4, rho, nu, gamma, G, eps, dt, next_mu, eslot, cfg=None,
                           workspace=None):
    """One chained backward step in the j-split form: ``stage1_vjp_split_f32`` at positions
    posm with the scaled cotangent mu4, and, with ``next_mu``, the next step's
    mu = nu + dt rho into nu and its scaled copy into mu4 in the same reduce kernel.  The
    scale exponent lags by one step (the previous mu's), which changes no bit of rho, nu or
    gamma unless a scaled component is subnormal.  ``eslot``: a zeroed int32 (2,) device
    tensor before the chain's first step, whose mu4 comes from ``adjoint_mu``."""
    n = posm.shape[0]
    for t, name, dtype, shape in ((posm, 'posm', torch.float32, (n, 4)),
                                  (mu4, 'mu4', torch.float32, (n, 4)),
                                  (rho, 'rho', torch.float64, (3, n)),
                                  (nu, 'nu', torch.float64, (3, n)),
                                  (gamma, 'gamma', torch.float64, (n,)),
                                  (eslot, 'eslot', torch.int32, (2,))):
        stages._check(t, name, dtype, shape)
    stages.same_device(posm, mu4, rho, nu, gamma, eslot)
    if workspace is None:
        workspace = stages.split_workspace(posm, n, cfg)
    k, s, p = cfg if cfg is not None else (0, 0, 0)
    with stages.on_device(posm):
        _native.call('grav_stage1_vjp_f32_split_chain', posm.data_ptr(), mu4.data_ptr(),
                     rho.data_ptr(), nu.data_ptr(), gamma.data_ptr(), n, float(G), float(eps),
                     float(dt), int(k), int(s), int(p), 1 if next_mu else 0,
                     workspace.data_ptr(), workspace.numel(), eslot.data_ptr(),
                     torch.cuda.current_stream(posm.device).cuda_stream)


def backward_pair_workspace_bytes(n, cfg=None):
    """Workspace bytes of ``backward_pair_split`` for n bodies (cfg as there)."""
    b = ctypes.c_int64(0)
    pa, pr = (cfg[1], cfg[3]) if cfg is not None else (0, 0)
    _native.call('grav_backward_pair_workspace', int(n), int(pa), int(pr), ctypes.byref(b))
    return b.value


def backward_pair_split(posm_adj, mu4, rho, nu, gamma, eslot, pos_in, pos_out, r64, v64, G,
                        eps, dt, workspace, cfg=None):
    """One chained backward step at positions posm_adj (``stage1_vjp_split_chain`` with
    next_mu) and one recomputed forward step from pos_in to pos_out with the masters r64, v64
    (``stages.step_split_f64m`` with out=pos_out) in the same two launches; the same bits as
    the two calls.  cfg: (vjp_i_per_lane, vjp_splits, step_i_per_lane, step_splits, j_waves)
    or None for the library's shapes; workspace: ``backward_pair_workspace_bytes``."""
    n = posm_adj.shape[0]
    for t, name, dtype, shape in ((posm_adj, 'posm_adj', torch.float32, (n, 4)),
                                  (mu4, 'mu4', torch.float32, (n, 4)),
                                  (rho, 'rho', torch.float64, (3, n)),
                                  (nu, 'nu', torch.float64, (3, n)),
                                  (gamma, 'gamma', torch.float64, (n,)),
                                  (eslot, 'eslot', torch.int32, (2,)),
                                  (pos_in, 'pos_in', torch.float32, (n, 4)),
                                  (pos_out, 'pos_out', torch.float32, (n, 4)),
                                  (r64, 'r64', torch.float64, (3, n)),
                                  (v64, 'v64', torch.float64, (3, n))):
        stages._check(t, name, dtype, shape)
    stages.same_device(posm_adj, mu4, rho, nu, gamma, eslot, pos_in, pos_out, r64, v64,
                       workspace)
    ka, pa, kr, pr, w = cfg if cfg is not None else (0, 0, 0, 0, 0)
    with stages.on_device(posm_adj):
        _native.call('grav_backward_pair_split', posm_adj.data_ptr(), mu4.data_ptr(),
                     rho.data_ptr(), nu.data_ptr(), gamma.data_ptr(), eslot.data_ptr(),
                     pos_in.data_ptr(), pos_out.data_ptr(), r64.data_ptr(), v64.data_ptr(), n,
                     float(G), float(eps), float(dt), int(ka), int(pa), int(kr), int(pr), int(w),
                     workspace.data_ptr(), workspace.numel(),
                     torch.cuda.current_stream(posm_adj.device).cuda_stream)


def default_budget(device):
    """A quarter of the device's free memory."""
    free, _ = torch.cuda.mem_get_info(device)
    return free // 4


def checkpoint_interval(n, steps, budget, buffers=1):
    """Returns (C, store): store=True records all T position sets (C = T, no recompute) when
    T*16n bytes fit ``budget``; otherwise C = ceil(sqrt(3T/B)) (the minimiser of
    ceil(T/C)*48n + B*C*16n bytes, with B = ``buffers`` segment records of C position sets),
    which must fit."""
    if steps <= 0:
        return max(1, steps), True
    if steps * 16 * n <= budget:
        return steps, True
    c = max(1, min(steps, int(math.ceil(math.sqrt(3.0 * steps / buffers)))))
    need = math.ceil(steps / c) * 48 * n + buffers * c * 16 * n
    if need > budget:
        raise MemoryError('adjoint needs %d bytes of checkpoints, budget is %d' % (need, budget))
    return c, False


def _check_state(r0, v0, m):
    n = r0.shape[0]
    for name, t, shape in (('r0', r0, (n, 3)), ('v0', v0, (n, 3)), ('m', m, (n,))):
        if t.dtype != torch.float64 or tuple(t.shape) != shape or t.device.type != 'cuda':
            raise ValueError('%s must be a float64 %s tensor on a HIP device' % (name, shape))
    return n


def _stage1_form(n, form):
    return stages.auto_form(n) if form == 'auto' else form


def _stage1(posm, acc, G, eps, form='auto', ws=None):
    """FP32 compensated stage 1 in the faster form for n, or owner-computes with
    form='owner' (forward and recompute use the same one, so a recomputed segment equals the
    recorded one bit for bit).  ``ws``: the form's workspace (default: the shared one)."""
    form = _stage1_form(posm.shape[0], form)
    if form == 'symmetric':
        stages.stage1_sym_f32(posm, acc, G, eps, None, ws)
    elif form == 'split':
        stages.stage1_split_f32(posm, acc, G, eps, None, ws)
    else:
        stages.stage1_f32(posm, acc, G, eps, 'kahan')


def _step(posm, acc, r64, v64, G, eps, dt, form='auto', ws=None):
    """One time step: stage 1 in the form for n, then the FP64-master stage 2; the j-split
    form as one fused step (``stages.step_split_f64m``, the same bits).  Every forward and
    recomputed step goes through here, so they agree bit for bit."""
    if _stage1_form(posm.shape[0], form) == 'split':
        stages.step_split_f64m(posm, r64, v64, G, eps, dt, None, ws)
    else:
        _stage1(posm, acc, G, eps, form, ws)
        stages.stage2_f64m(posm, r64, v64, acc, dt)


def _vjp_split(n, form):
    return form == 'auto' and n >= stages.VJP_SPLIT_FROM


def _vjp(posm, mu4, rho, gamma, G, eps, dt, form='auto', ws=None):
    """The VJP in the faster form for n (j-split from VJP_SPLIT_FROM bodies), or
    owner-computes with form='owner'.  ``ws``: the j-split workspace (default: the shared
    one)."""
    if _vjp_split(posm.shape[0], form):
        stage1_vjp_split_f32(posm, mu4, rho, gamma, G, eps, dt, None, ws)
    else:
        stage1_vjp_f32(posm, mu4, rho, gamma, G, eps, dt)


def _capture(fn):
    """A HIP graph of fn() on a side stream (fn's buffers and workspaces already exist, so
    nothing is allocated while capturing)."""
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    return g


class _SegmentGraphs:
    """Fixed buffers of one universe configuration (device, n, G, eps, dt, checkpoint
    interval, stage-1 form) and the HIP graphs of one checkpoint segment on them: the forward
    steps, the backward's recomputation into a record buffer, and the adjoint steps over a
    record buffer.  Kept across calls (the last few configurations), so a trajectory replays
    graphs that were captured once.  It owns the workspaces its kernels use, so the graphs
    never point at memory that something else frees.

    The forward and recompute state (posm, acc, r64, v64, ws) and the adjoint state (rho, nu,
    gamma, mu4, eslot, ws_adj) are separate, and there are two record buffers, so that the
    recomputation of one segment can share launches with the adjoint steps of the next.  In
    the j-split form a recomputed step goes from one record slot to the next (no copy), and a
    backward step is chained (its reduce forms the next cotangent)."""

    def __init__(self, n, dev, G, eps, dt, interval, form, records=2):
        f32, f64 = torch.float32, torch.float64
        self.cfg = (G, eps, dt, interval, form)
        self.posm = torch.empty((n, 4), dtype=f32, device=dev)
        self.acc = torch.empty((3, n), dtype=f32, device=dev)
        self.r64 = torch.empty((3, n), dtype=f64, device=dev)
        self.v64 = torch.empty((3, n), dtype=f64, device=dev)
        self.traj = [torch.empty((interval, n, 4), dtype=f32, device=dev)
                     for _ in range(records)]
        self.rho = torch.empty((3, n), dtype=f64, device=dev)
        self.nu = torch.empty((3, n), dtype=f64, device=dev)
        self.gamma = torch.empty((n,), dtype=f64, device=dev)
        self.mu4 = torch.empty((n, 4), dtype=f32, device=dev)
        self.eslot = torch.zeros((2,), dtype=torch.int32, device=dev)
        s1 = _stage1_form(n, form)
        need = (stages.sym_workspace_size(n, device=dev) if s1 == 'symmetric'
                else stages.split_workspace_bytes(n) if s1 == 'split' else 0)
        self.ws = torch.empty((max(need, 256),), dtype=torch.uint8, device=dev)
        self.io = s1 == 'split'
        self.chain = CHAIN and _vjp_split(n, form)
        self.paired = OVERLAP and self.io and self.chain and records == 2
        # the adjoint's own workspace: the j-split VJP's partials (or the paired steps' two
        # sets) and grav_adjoint_mu_ws's exponents, each finished with it before the next
        # starts on the same stream
        need = stages.split_workspace_bytes(n) if _vjp_split(n, form) else 0
        if self.paired:
            need = max(need, backward_pair_workspace_bytes(n))
        need = max(need, 4 * ((n + 1023) // 1024))
        self.ws_adj = torch.empty((max(need, 256),), dtype=torch.uint8, device=dev)
        self.graphs = {}
        # the token of the forward pass whose last segment the record buffers hold (the
        # backward of that pass need not recompute it); any other use of them clears it
        self.recorded = None

    def step(self):
        G, eps, dt, _, form = self.cfg
        _step(self.posm, self.acc, self.r64, self.v64, G, eps, dt, form, self.ws)

    def forward(self, k):
        for _ in range(k):
            self.step()

    def forward_record(self, k, buf):
        """k forward steps from the current state (as ``forward``), recording the positions
        each one uses into record buffer ``buf``: slots 0 .. k-1 are then what the backward's
        recomputation of this segment would write (the same step kernels on the same state),
        and the last step goes back into posm."""
        G, eps, dt, _, _ = self.cfg
        traj = self.traj[buf]
        traj[0].copy_(self.posm)
        for t in range(k):
            stages.step_split_f64m(traj[t], self.r64, self.v64, G, eps, dt, None, self.ws,
                                   out=traj[t + 1] if t + 1 < k else self.posm)

    def load(self, r_c, v_c, mass, buf):
        """A checkpoint (r, v) into the recompute state and slot 0 of record buffer ``buf``."""
        self.r64.copy_(r_c)
        self.v64.copy_(v_c)
        p0 = self.traj[buf][0]
        p0[:, :3] = self.r64.t()
        p0[:, 3] = mass

    def recompute(self, k, buf):
        """Slots 1 .. k-1 of record buffer ``buf`` from slot 0 and the loaded masters (the
        segment's last step is not needed: its result is the next checkpoint)."""
        G, eps, dt, _, form = self.cfg
        traj = self.traj[buf]
        if self.io:
            for t in range(k - 1):
                stages.step_split_f64m(traj[t], self.r64, self.v64, G, eps, dt, None, self.ws,
                                       out=traj[t + 1])
            return
        if k > 1:
            self.posm.copy_(traj[0])
        for t in range(k - 1):
            self.step()
            traj[t + 1].copy_(self.posm)

    def begin_backward(self):
        """The first backward step's cotangent (and a cleared exponent slot for the chain);
        with the chain every later one comes from the previous step's reduce."""
        if self.chain:
            adjoint_mu(self.rho, self.nu, self.mu4, self.cfg[2], self.ws_adj)
            self.eslot.zero_()

    def adjoint(self, k, last, traj):
        """Backward steps k-1 .. 0 over the record ``traj``; ``last``: the segment holds step
        0 of the trajectory, after which nu must keep that step's mu."""
        G, eps, dt, _, form = self.cfg
        for t in range(k - 1, -1, -1):
            if self.chain:
                stage1_vjp_split_chain(traj[t], self.mu4, self.rho, self.nu, self.gamma, G, eps,
                                       dt, not (last and t == 0), self.eslot, None, self.ws_adj)
            else:
                adjoint_mu(self.rho, self.nu, self.mu4, dt, self.ws_adj)
                _vjp(traj[t], self.mu4, self.rho, self.gamma, G, eps, dt, form, self.ws_adj)

    def pair(self, k, buf):
        """Backward steps k-1 .. 0 over record buffer ``buf`` (chained, never the
        trajectory's step 0) together with the recomputation of the other buffer's slots
        1 .. C-1 from its slot 0 (the previous segment, always a full one of C steps): step j
        of each shares the launches of ``backward_pair_split``; the longer one finishes
        alone."""
        G, eps, dt, interval, _ = self.cfg
        ta, tr = self.traj[buf], self.traj[1 - buf]
        for j in range(max(k, interval - 1)):
            t = k - 1 - j
            if j < k and j < interval - 1:
                backward_pair_split(ta[t], self.mu4, self.rho, self.nu, self.gamma, self.eslot,
                                    tr[j], tr[j + 1], self.r64, self.v64, G, eps, dt,
                                    self.ws_adj)
            elif j < k:
                stage1_vjp_split_chain(ta[t], self.mu4, self.rho, self.nu, self.gamma, G, eps,
                                       dt, True, self.eslot, None, self.ws_adj)
            else:
                stages.step_split_f64m(tr[j], self.r64, self.v64, G, eps, dt, None, self.ws,
                                       out=tr[j + 1])

    def run(self, which, k, buf=0, last=False, graph=True):
        """A segment of k steps: 'fwd' (forward), 'rec' (recompute into buffer buf), 'adj'
        (adjoint over buffer buf) or 'pair' (adjoint over buffer buf with the other buffer's
        recomputation).  Full segments with ``graph`` replay their graph if it exists, else
        run eagerly once and are captured for the next time."""
        fn = {'fwd': lambda: self.forward(k), 'fwdrec': lambda: self.forward_record(k, buf),
              'rec': lambda: self.recompute(k, buf),
              'adj': lambda: self.adjoint(k, last, self.traj[buf]),
              'pair': lambda: self.pair(k, buf)}[which]
        if not graph or k != self.cfg[3]:
            fn()
            return
        key = (which, buf, last)
        g = self.graphs.get(key)
        if g is not None:
            g.replay()
            return
        fn()
        self.graphs[key] = _capture(fn)


_SEGMENT_GRAPHS = collections.OrderedDict()


def _segment_graphs(n, dev, G, eps, dt, interval, form):
    key = (dev, n, G, eps, dt, interval, form, CHAIN, OVERLAP)
    sg = _SEGMENT_GRAPHS.pop(key, None)
    if sg is None:
        sg = _SegmentGraphs(n, dev, G, eps, dt, interval, form, 2 if OVERLAP else 1)
        while len(_SEGMENT_GRAPHS) >= 4:
            _SEGMENT_GRAPHS.popitem(last=False)
    _SEGMENT_GRAPHS[key] = sg
    return sg


def _backward_segments(sg, checkpoints, mass, steps, interval, graph, recorded=False):
    """The checkpointed backward pass: for c = C-1 .. 0, segment c's positions are
    recomputed into record buffer c % 2 from its checkpoint, then its adjoint steps run.
    Paired (``sg.paired``), the recomputation of segment c - 1 shares the launches of
    segment c's adjoint steps; otherwise each segment is recomputed, then differentiated.
    ``recorded``: the forward pass left segment C-1's positions in its record buffer."""
    C = len(checkpoints)
    seg_k = [min(steps, (c + 1) * interval) - c * interval for c in range(C)]
    sg.recorded = None
    if not sg.paired:
        for c in range(C - 1, -1, -1):
            r_c, v_c = checkpoints[c]
            sg.load(r_c, v_c, mass, 0)
            sg.run('rec', seg_k[c], 0, graph=graph)
            sg.run('adj', seg_k[c], 0, c == 0, graph=graph)
        return
    if not recorded:
        r_c, v_c = checkpoints[C - 1]
        sg.load(r_c, v_c, mass, (C - 1) % 2)
        sg.run('rec', seg_k[C - 1], (C - 1) % 2, graph=graph)
    for c in range(C - 1, -1, -1):
        if c == 0:
            sg.run('adj', seg_k[0], 0, True, graph=graph)
            break
        r_c, v_c = checkpoints[c - 1]
        sg.load(r_c, v_c, mass, (c - 1) % 2)
        sg.run('pair', seg_k[c], c % 2, graph=graph)


class _Single(torch.autograd.Function):
    """One universe; per step FP32 stage 1 (owner-computes, j-split or symmetric by n) +
    grav_stage2_f64m forward and grav_adjoint_mu + the VJP (grav_stage1_vjp_f32, j-split
    from VJP_SPLIT_FROM bodies, chained) backward.  With checkpoints and ``graph``, every full
    segment of ``interval`` steps runs as one HIP graph replay (forward, recompute and
    adjoint each), captured once per configuration on fixed buffers: the same launches in the
    same order as eager stepping, without a host launch per kernel."""

    @staticmethod
    def forward(ctx, r0, v0, m, G, eps, dt, steps, interval, store, form, graph):
        n = r0.shape[0]
        dev = r0.device
        sg = (_segment_graphs(n, dev, G, eps, dt, interval, form)
              if graph and not store else None)
        if sg is not None:
            posm, acc, r64, v64 = sg.posm, sg.acc, sg.r64, sg.v64
            r64.copy_(r0.detach().t())
            v64.copy_(v0.detach().t())
        else:
            posm = torch.empty((n, 4), dtype=torch.float32, device=dev)
            acc = torch.empty((3, n), dtype=torch.float32, device=dev)
            r64 = r0.detach().t().contiguous()
            v64 = v0.detach().t().contiguous()
        posm[:, :3] = r0
        posm[:, 3] = m
        traj = torch.empty((steps, n, 4), dtype=torch.float32, device=dev) if store else None
        checkpoints = []
        t = 0
        while t < steps:
            if store:
                traj[t].copy_(posm)
                _step(posm, acc, r64, v64, G, eps, dt, form)
                t += 1
                continue
            checkpoints.append((r64.clone(), v64.clone()))
            k = min(interval, steps - t)
            if sg is not None and sg.paired and t + k == steps:
                # the last segment records its positions for the backward pass
                sg.run('fwdrec', k, (len(checkpoints) - 1) % 2)
                ctx.rec_token = sg.recorded = object()
            elif sg is not None:
                sg.run('fwd', k)
            else:
                for _ in range(k):
                    _step(posm, acc, r64, v64, G, eps, dt, form)
            t += k
        ctx.cfg = (G, eps, dt, steps, interval, store, form, graph)
        ctx.traj = traj
        ctx.checkpoints = checkpoints
        ctx.mass = posm[:, 3].clone()
        return r64.t().contiguous(), v64.t().contiguous()

    @staticmethod
    def backward(ctx, g_r, g_v):
        G, eps, dt, steps, interval, store, form, graph = ctx.cfg
        n = ctx.mass.shape[0]
        dev = ctx.mass.device
        if store:
            sg = _SegmentGraphs(n, dev, G, eps, dt, 1, form, records=0)
        elif not graph:
            sg = _SegmentGraphs(n, dev, G, eps, dt, interval, form, 2 if OVERLAP else 1)
        else:
            sg = _segment_graphs(n, dev, G, eps, dt, interval, form)
        if g_r is not None:
            sg.rho.copy_(g_r.t())
        else:
            sg.rho.zero_()
        if g_v is not None:
            sg.nu.copy_(g_v.t())
        else:
            sg.nu.zero_()
        sg.gamma.zero_()
        sg.begin_backward()
        if store:
            sg.adjoint(steps, True, ctx.traj)
        else:
            _backward_segments(sg, ctx.checkpoints, ctx.mass, steps, interval, graph,
                               recorded=sg.recorded is getattr(ctx, 'rec_token', 0))
        ctx.traj = None
        ctx.checkpoints = None
        return (sg.rho.t().clone(), sg.nu.t().clone(), sg.gamma.clone(), None, None, None, None,
                None, None, None, None)


def trajectory(r0, v0, m, steps, G=1.0, eps=0.0, dt=1e-3, interval=None, budget=None,
               form='auto', graph=True):
    """Differentiable T-step trajectory of one universe: returns (r_T, v_T), both (n, 3)
    float64, with gradients to r0 (n, 3), v0 (n, 3) and m (n,) (all float64 on the device).
    ``interval`` forces a checkpoint interval (no trajectory record); else it follows
    ``budget`` bytes (default: a quarter of free device memory).  ``form`` 'auto' picks the
    stage-1 and VJP forms by n, 'owner' keeps the one-launch owner-computes kernels.  With
    checkpoints, ``graph`` replays every full segment after the first as a HIP graph
    (bitwise the same result as eager stepping)."""
    if form not in ('auto', 'owner'):
        raise ValueError("form must be 'auto' or 'owner'")
    n = _check_state(r0, v0, m)
    if interval is None:
        interval, store = checkpoint_interval(
            n, steps, default_budget(r0.device) if budget is None else budget,
            2 if OVERLAP else 1)
    else:
        interval, store = int(interval), False
    return _Single.apply(r0, v0, m, float(G), float(eps), float(dt), int(steps), interval, store,
                         form, bool(graph))


class _Batched(torch.autograd.Function):
    """Many universes through a batched.Engine: 'fused' (one launch per segment and pass) or
    'perstep' (launches per step over every universe's tiles)."""

    @staticmethod
    def forward(ctx, r0, v0, m, engine, G, eps, dt, steps, interval, store):
        n = r0.shape[0]
        dev = r0.device
        posm = torch.empty((n, 4), dtype=torch.float32, device=dev)
        posm[:, :3] = r0
        posm[:, 3] = m
        r64 = r0.detach().t().contiguous()
        v64 = v0.detach().t().contiguous()
        traj = None
        checkpoints = []
        if store:
            traj = torch.empty((steps, n, 4), dtype=torch.float32, device=dev)
            engine.forward(posm, r64, v64, G, eps, dt, steps, traj)
        else:
            for t0 in range(0, steps, interval):
                checkpoints.append((r64.clone(), v64.clone()))
                engine.forward(posm, r64, v64, G, eps, dt, min(interval, steps - t0))
        ctx.cfg = (engine, G, eps, dt, steps, interval, store)
        ctx.traj = traj
        ctx.checkpoints = checkpoints
        ctx.mass = posm[:, 3].clone()
        return r64.t().contiguous(), v64.t().contiguous()

    @staticmethod
    def backward(ctx, g_r, g_v):
        engine, G, eps, dt, steps, interval, store = ctx.cfg
        n = ctx.mass.shape[0]
        dev = ctx.mass.device
        rho = (g_r.t().contiguous().double() if g_r is not None
               else torch.zeros((3, n), dtype=torch.float64, device=dev))
        nu = (g_v.t().contiguous().double() if g_v is not None
              else torch.zeros((3, n), dtype=torch.float64, device=dev))
        gamma = torch.zeros((n,), dtype=torch.float64, device=dev)
        if store:
            engine.adjoint(ctx.traj, rho, nu, gamma, G, eps, dt, steps)
        else:
            posm = torch.empty((n, 4), dtype=torch.float32, device=dev)
            traj = torch.empty((interval, n, 4), dtype=torch.float32, device=dev)
            for c in range(len(ctx.checkpoints) - 1, -1, -1):
                cnt = min(interval, steps - c * interval)
                r64, v64 = ctx.checkpoints[c][0].clone(), ctx.checkpoints[c][1].clone()
                posm[:, :3] = r64.t()
                posm[:, 3] = ctx.mass
                engine.forward(posm, r64, v64, G, eps, dt, cnt, traj[:cnt])
                engine.adjoint(traj[:cnt], rho, nu, gamma, G, eps, dt, cnt)
        ctx.traj = None
        ctx.checkpoints = None
        return rho.t(), nu.t(), gamma, None, None, None, None, None, None, None


def batched_trajectory(engine, r0, v0, m, steps, G=1.0, eps=0.0, dt=1e-3, interval=None,
                       budget=None):
    """Differentiable T-step trajectories of all universes of ``engine`` (a batched.Engine,
    or a batched.BatchedPlan, which runs the fused schedule), over the concatenated bodies;
    same contract as ``trajectory``."""
    n = _check_state(r0, v0, m)
    if isinstance(engine, batched.BatchedPlan):
        engine = batched.Engine.wrap(engine)
    if not isinstance(engine, batched.Engine):
        raise TypeError('engine must be a gravitation_amd.batched.Engine or BatchedPlan')
    if engine.n != n:
        raise ValueError('engine holds %d bodies, state has %d' % (engine.n, n))
    if interval is None:
        interval, store = checkpoint_interval(
            n, steps, default_budget(r0.device) if budget is None else budget)
    else:
        interval, store = int(interval), False
    return _Batched.apply(r0, v0, m, engine, float(G), float(eps), float(dt), int(steps),
                          interval, store)


def fit_initial_velocities(engine, r0, v0, m, target, steps, iters=50, lr=None, G=1.0,
                           eps=0.0, dt=1e-3, interval=None, budget=None, stats=None,
                           trust=0.1):
    """Orbit fitting by gradient descent on the initial velocities of every universe of
    ``engine``: minimises L_u = sum over universe u of |r_T - target|^2 with the batched
    adjoint, every universe on its own.

    Under free motion r_T = r0 + T dt v0, so dL_u/dv0 = 2 (T dt)^2 (v0 - v*) and the default
    step ``lr = 0.5 / (T dt)^2`` is that case's Newton step.  Close encounters make the map
    v0 -> r_T far from linear and its gradient huge (up to 1e60 over a few hundred chaotic
    steps), so each universe also keeps a trust radius delta_u on the RMS velocity change
    per body: the step is -min(lr, delta_u / |g_u|_rms) g_u.  delta_u starts at ``trust``
    times the universe's RMS distance from its centre divided by T dt (the velocity that
    crosses the system within the horizon), doubles after an accepted update and shrinks to a
    quarter of the rejected step after an update that does not lower the universe's loss (or
    makes it non-finite), which is undone.  Each iteration is one batched forward + adjoint
    of all universes.

    Returns (best v0 (n, 3) float64, list of the total loss of the best state after each
    iteration, non-increasing).  A dict passed as ``stats`` receives per-universe first and
    best losses and the number of rejected updates."""
    if lr is None:
        lr = 0.5 / (steps * dt) ** 2
    plan = engine.plan if isinstance(engine, batched.Engine) else engine
    dev = v0.device
    counts = torch.as_tensor(np.diff(plan.offsets), device=dev)
    n_uni = counts.shape[0]
    uid = torch.repeat_interleave(torch.arange(n_uni, device=dev), counts)
    cnt = counts.double()

    def per_universe(x):
        return torch.zeros(n_uni, dtype=torch.float64, device=dev).index_add_(0, uid, x)

    r0d = r0.detach().double()
    centre = torch.zeros((n_uni, 3), dtype=torch.float64, device=dev).index_add_(0, uid, r0d)
    centre /= cnt[:, None]
    size = torch.sqrt(per_universe(((r0d - centre[uid]) ** 2).sum(1)) / cnt)
    delta = float(trust) * size / (steps * dt)
    delta = torch.where(delta > 0, delta, torch.full_like(delta, float('inf')))
    v = v0.detach().clone()
    v_ok = g_ok = l_ok = l_first = s_try = None
    losses = []
    rejected = 0
    for _ in range(iters):
        vv = v.clone().requires_grad_(True)
        r_t, _ = batched_trajectory(engine, r0, vv, m, steps, G, eps, dt, interval, budget)
        d2 = ((r_t - target) ** 2).sum(1)
        d2.sum().backward()
        g = vv.grad.detach()
        per = per_universe(d2.detach().double())
        g_bad = per_universe((~torch.isfinite(g).all(1)).double()) > 0
        good = torch.isfinite(per) & ~g_bad
        if l_ok is None:
            g = torch.where(good[uid][:, None], g, torch.zeros_like(g))
            l_first = per.clone()
            v_ok, g_ok, l_ok = v, g, per
        else:
            ok = good & (per <= l_ok)
            rejected += int((~ok).sum())
            okb = ok[uid][:, None]
            v_ok = torch.where(okb, v, v_ok)
            g_ok = torch.where(okb, g, g_ok)
            l_ok = torch.where(ok, per, l_ok)
            delta = torch.where(ok, 2.0 * delta, 0.25 * torch.minimum(delta, s_try))
        losses.append(float(l_ok.sum()))
        # step length: the Newton step of free motion, capped by the trust radius
        g_rms = torch.sqrt(per_universe((g_ok.double() ** 2).sum(1)) / cnt)
        k = torch.where(g_rms > 0, torch.clamp(delta / g_rms, max=lr), torch.zeros_like(g_rms))
        s_try = k * g_rms
        v = v_ok - k[uid][:, None] * g_ok
    if stats is not None:
        stats['first'] = l_first.cpu().numpy()
        stats['best'] = l_ok.cpu().numpy()
        stats['rejected'] = rejected
    return v_ok, losses
