"""Torch-facing wrappers of the stage kernels.

Tensors are plumbing here: device memory, shapes and streams.  Each function checks the
tensors it is handed (device, dtype, shape, contiguity) and passes raw device pointers and
torch's current HIP stream to libgravhip.  Layouts:

* ``posm``   float32 (n, 4): x, y, z, m per body (16-byte rows)
* ``posm64`` float64 (n, 4)
* ``acc``, ``vel``, ``r64``, ``v64``: (3, n) SoA, i.e. the F-order (n, 3) layout that
  src/gravitation/kernel/np2.py:64-66 uses for its state arrays.
"""

import contextlib
import ctypes

import torch

from . import _native

_ACCUM = {'plain': _native.ACCUM_PLAIN, 'kahan': _native.ACCUM_KAHAN}


_SAME = contextlib.nullcontext()


def on_device(dev):
    """Context that makes ``dev`` (a torch device or a tensor's) the current HIP device for
    the library calls inside it; free when it already is.  The library launches on the current
    device, into the stream it is handed, so both must be the tensors' device."""
    if isinstance(dev, torch.Tensor):
        dev = dev.device
    elif not isinstance(dev, torch.device):
        dev = torch.device(dev)
    idx = dev.index
    if idx is None or idx == torch.cuda.current_device():
        return _SAME
    return torch.cuda.device(idx)


def _stream(t):
    return ctypes_stream(torch.cuda.current_stream(t.device))


def ctypes_stream(s):
    return s.cuda_stream


def _check(t, name, dtype, shape):
    if not isinstance(t, torch.Tensor):
        raise TypeError('%s must be a torch.Tensor' % name)
    if t.device.type != 'cuda':
        raise ValueError('%s must live on a HIP device (got %s)' % (name, t.device))
    if t.dtype != dtype:
        raise TypeError('%s must be %s (got %s)' % (name, dtype, t.dtype))
    if tuple(t.shape) != tuple(shape):
        raise ValueError('%s must have shape %s (got %s)' % (name, tuple(shape), tuple(t.shape)))
    if not t.is_contiguous():
        raise ValueError('%s must be contiguous' % name)


def same_device(first, *others):
    for t in others:
        if t.device != first.device:
            raise ValueError('all tensors of a call must be on one device (%s and %s)'
                             % (first.device, t.device))


def accum_code(accum):
    try:
        return _ACCUM[accum]
    except KeyError:
        raise ValueError("accum must be 'plain' or 'kahan' (got %r)" % (accum,))


def stage1_f32(posm, acc, G, eps=0.0, accum='kahan', cfg=None):
    """acc (3, n) <- FP32 accelerations of posm (n, 4).  cfg = (i_per_lane, j_waves, j_block)."""
    n = posm.shape[0]
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(acc, 'acc', torch.float32, (3, n))
    same_device(posm, acc)
    code = accum_code(accum)
    if cfg is None:
        with on_device(posm):
            _native.call('grav_stage1_f32', posm.data_ptr(), acc.data_ptr(), n, float(G),
                         float(eps), code, _stream(posm))
    else:
        k, s, jb = cfg
        with on_device(posm):
            _native.call('grav_stage1_f32_cfg', posm.data_ptr(), acc.data_ptr(), n, float(G),
                         float(eps), code, int(k), int(s), int(jb), _stream(posm))
    return acc


_workspaces = {}


# Form thresholds, from stage-1 times on Plummer spheres on an MI355X.  FP32: one launch
# (owner-computes) beats the two of the j-split form below 2048 bodies; the j-split form beats
# the symmetric one, whose few block pairs cannot fill the GPU, below 65536 bodies (ms, best
# of 10, j-split against the symmetric form's default shape: 40960 0.44 against 0.48, 49152
# 0.61 against 0.60, 57344 0.80 against 0.94, 61440 0.91 against 0.94, 65536 1.04 against
# 0.98).  FP64: j-split from 1024 bodies
# (3.2 against 11.8 us of kernel time at 512), symmetric from 14336, between the measured
# points (ms, best of 20, j-split against symmetric: 8192 0.053 against 0.076, 12288 0.096
# against 0.110, 16384 0.160 against 0.140, 20480 0.250 against 0.213).
SPLIT_FROM = 2048
SYMMETRIC_FROM = 65536
SPLIT_F64_FROM = 1024
SYMMETRIC_F64_FROM = 14336
VJP_SPLIT_FROM = 1024


def auto_form(n, accum='kahan', free_bytes=None, dtype='float32'):
    """The faster stage-1 form for n bodies.  FP32 (compensated accumulation only): 'owner'
    below SPLIT_FROM, where one launch beats two, 'split' up to SYMMETRIC_FROM, 'symmetric'
    from there; FP64: 'owner' below SPLIT_F64_FROM, 'split' up to SYMMETRIC_F64_FROM,
    'symmetric' from there.  With ``free_bytes`` (free device memory), 'symmetric' also needs
    its smallest workspace (about 50 bytes per body in FP32; the form works in slabs of rows
    when it is given less than its one-slab size) to fit in half of it, else 'owner'."""
    if dtype == 'float64':
        if n < SPLIT_F64_FROM:
            return 'owner'
        if n < SYMMETRIC_F64_FROM:
            return 'split'
        need = sym_f64_workspace_bytes(n, minimum=True) if free_bytes is not None else 0
    else:
        if accum != 'kahan' or n < SPLIT_FROM:
            return 'owner'
        if n < SYMMETRIC_FROM:
            return 'split'
        need = sym_workspace_bytes(n, minimum=True) if free_bytes is not None else 0
    if free_bytes is not None and need > free_bytes // 2:
        return 'owner'
    return 'symmetric'


def sym_workspace_bytes(n, cfg=None, minimum=False):
    """Workspace bytes of the symmetric form for one slab (every block pair at once), or with
    ``minimum`` the smallest it accepts (one row of block pairs per slab); cfg None = the
    library's shape for n."""
    k, w = cfg if cfg is not None else (0, 0)
    b = ctypes.c_int64(0)
    fn = 'grav_stage1_sym_workspace_min' if minimum else 'grav_stage1_sym_workspace'
    _native.call(fn, int(n), int(k), int(w), ctypes.byref(b))
    return b.value


def sym_workspace_size(n, cfg=None, free_bytes=None, dtype='float32', device=None):
    """The workspace to allocate for the symmetric form: its one-slab size when that fits a
    quarter of the free device memory (``free_bytes``, default: measured on ``device``), else
    that quarter (at least the form's minimum).  Any size from the minimum up gives the same
    accelerations to rounding; fewer slabs save a few launches and fold passes."""
    of = sym_f64_workspace_bytes if dtype == 'float64' else sym_workspace_bytes
    full = of(n, cfg)
    if free_bytes is None:
        import torch
        free_bytes, _ = torch.cuda.mem_get_info(device)
    cap = free_bytes // 4
    if full <= cap:
        return full
    return max(of(n, cfg, minimum=True), cap & ~255)


def stage1_sym_f32(posm, acc, G, eps=0.0, cfg=None, workspace=None):
    """acc (3, n) <- FP32 accelerations of posm (n, 4), symmetric form (each pair once).
    cfg = (i_chunks_per_wave, waves), None for the library's shape for n.  ``workspace``: a
    uint8 device tensor of at least ``sym_workspace_bytes(n, cfg, minimum=True)`` bytes (the
    form runs in slabs of rows below ``sym_workspace_bytes(n, cfg)``); by default one of
    ``sym_workspace_size`` bytes is kept per (device, n, cfg)."""
    n = posm.shape[0]
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(acc, 'acc', torch.float32, (3, n))
    same_device(posm, acc)
    if workspace is None:
        key = (posm.device, n, None if cfg is None else tuple(cfg))
        workspace = _workspaces.get(key)
        if workspace is None:
            size = sym_workspace_size(n, cfg, device=posm.device)
            workspace = torch.empty((max(size, 256),), dtype=torch.uint8, device=posm.device)
            _workspaces[key] = workspace
    same_device(posm, workspace)
    need = sym_workspace_bytes(n, cfg, minimum=True)
    if workspace.numel() < need:
        raise ValueError('workspace holds %d bytes, at least %d needed'
                         % (workspace.numel(), need))
    k, w = cfg if cfg is not None else (0, 0)
    with on_device(posm):
        _native.call('grav_stage1_sym_f32', posm.data_ptr(), acc.data_ptr(), n, float(G),
                     float(eps), int(k), int(w), workspace.data_ptr(), workspace.numel(),
                     _stream(posm))
    return acc


def sym_f64_workspace_bytes(n, cfg=None, minimum=False):
    """Workspace bytes of the FP64 symmetric form, one slab or (``minimum``) the smallest;
    cfg None = the library's shape for n."""
    k, w = cfg if cfg is not None else (0, 0)
    b = ctypes.c_int64(0)
    fn = 'grav_stage1_sym_f64_workspace_min' if minimum else 'grav_stage1_sym_f64_workspace'
    _native.call(fn, int(n), int(k), int(w), ctypes.byref(b))
    return b.value


def stage1_sym_f64(posm64, acc64, G, eps=0.0, cfg=None, workspace=None):
    """acc64 (3, n) <- FP64 accelerations of posm64 (n, 4), symmetric form (each pair once).
    cfg = (i_chunks_per_wave, waves) in {2,4} x {2,4}, None for the library's shape for n;
    ``workspace`` as in stage1_sym_f32 (sizes from ``sym_f64_workspace_bytes``)."""
    n = posm64.shape[0]
    _check(posm64, 'posm64', torch.float64, (n, 4))
    _check(acc64, 'acc64', torch.float64, (3, n))
    same_device(posm64, acc64)
    if workspace is None:
        key = ('f64', posm64.device, n, None if cfg is None else tuple(cfg))
        workspace = _workspaces.get(key)
        if workspace is None:
            size = sym_workspace_size(n, cfg, dtype='float64', device=posm64.device)
            workspace = torch.empty((max(size, 256),), dtype=torch.uint8, device=posm64.device)
            _workspaces[key] = workspace
    same_device(posm64, workspace)
    need = sym_f64_workspace_bytes(n, cfg, minimum=True)
    if workspace.numel() < need:
        raise ValueError('workspace holds %d bytes, at least %d needed'
                         % (workspace.numel(), need))
    k, w = cfg if cfg is not None else (0, 0)
    with on_device(posm64):
        _native.call('grav_stage1_sym_f64', posm64.data_ptr(), acc64.data_ptr(), n, float(G),
                     float(eps), int(k), int(w), workspace.data_ptr(), workspace.numel(),
                     _stream(posm64))
    return acc64


def check_split_cfg(cfg, dtype='float32'):
    """Raises ValueError unless cfg is None or a supported (i_per_lane, j_waves, splits):
    j_waves in {1, 2, 4, 8} for FP32, {4, 8} for FP64."""
    if cfg is None:
        return
    waves = (4, 8) if dtype == 'float64' else (1, 2, 4, 8)
    ok = len(cfg) == 3 and cfg[0] in (1, 2) and cfg[1] in waves and 1 <= int(cfg[2]) <= 64
    if not ok:
        raise ValueError('split shape (i_per_lane, j_waves, splits) must be in {1, 2} x %s x '
                         '[1, 64] (got %r)' % (set(waves), cfg))


def split_workspace_bytes(n, cfg=None):
    """Workspace bytes of the j-split forms (stage 1 and its VJP share the size); cfg =
    (i_per_lane, j_waves, splits), None for the library's shape for n."""
    p = int(cfg[2]) if cfg is not None else 0
    b = ctypes.c_int64(0)
    _native.call('grav_stage1_split_workspace', int(n), p, ctypes.byref(b))
    return b.value


def split_workspace(t, n, cfg=None):
    """A cached uint8 workspace of the j-split forms on t's device for n bodies."""
    key = ('split', t.device, n, None if cfg is None else tuple(cfg))
    ws = _workspaces.get(key)
    if ws is None:
        ws = torch.empty((max(split_workspace_bytes(n, cfg), 256),), dtype=torch.uint8,
                         device=t.device)
        _workspaces[key] = ws
    return ws


def stage1_split_f32(posm, acc, G, eps=0.0, cfg=None, workspace=None):
    """acc (3, n) <- FP32-compensated accelerations of posm (n, 4) in the j-split
    owner-computes form (each i-tile also split over j-ranges, for small n).  cfg =
    (i_per_lane, j_waves, splits), None for the library's shape for n.  ``workspace``: a
    uint8 device tensor of at least ``split_workspace_bytes(n, cfg)`` bytes; by default one is
    kept per (device, n, cfg)."""
    n = posm.shape[0]
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(acc, 'acc', torch.float32, (3, n))
    same_device(posm, acc)
    if workspace is None:
        workspace = split_workspace(posm, n, cfg)
    same_device(posm, workspace)
    k, s, p = cfg if cfg is not None else (0, 0, 0)
    with on_device(posm):
        _native.call('grav_stage1_f32_split', posm.data_ptr(), acc.data_ptr(), n, float(G),
                     float(eps), int(k), int(s), int(p), workspace.data_ptr(),
                     workspace.numel(), _stream(posm))
    return acc


def sym_share_plan(n, share, shares, cfg=None):
    """((first row, end row, first partner row, end partner row, first diagonal block, end
    diagonal block), full bytes, minimum bytes) of share `share` of `shares` of the symmetric
    form for n bodies; cfg = (i_chunks_per_wave, waves), None for the library's shape for n."""
    k, w = cfg if cfg is not None else (0, 0)
    rows = (ctypes.c_int * 6)()
    b = (ctypes.c_int64 * 2)()
    _native.call('grav_stage1_sym_share_plan', int(n), int(share), int(shares), int(k), int(w),
                 rows, b)
    return tuple(rows), b[0], b[1]


def stage1_sym_f32_share(posm, sums, share, shares, eps=0.0, cfg=None, workspace=None):
    """sums (3, n) float64 <- the unscaled contribution of share `share` of `shares` of the
    symmetric form (its rows of block pairs and their diagonal blocks) to every body's
    acceleration: one GPU's part when the form is divided among GPUs.  ``workspace``: a uint8
    device tensor of at least the share's minimum bytes (``sym_share_plan``); by default one of
    its full size is kept per (device, n, share, shares, cfg)."""
    n = posm.shape[0]
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(sums, 'sums', torch.float64, (3, n))
    same_device(posm, sums)
    if workspace is None:
        key = ('sym_share', posm.device, n, share, shares, None if cfg is None else tuple(cfg))
        workspace = _workspaces.get(key)
        if workspace is None:
            _, full, _ = sym_share_plan(n, share, shares, cfg)
            workspace = torch.empty((full,), dtype=torch.uint8, device=posm.device)
            _workspaces[key] = workspace
    same_device(posm, workspace)
    k, w = cfg if cfg is not None else (0, 0)
    with on_device(posm):
        _native.call('grav_stage1_sym_f32_share', posm.data_ptr(), sums.data_ptr(), n,
                     float(eps), int(share), int(shares), int(k), int(w), workspace.data_ptr(),
                     workspace.numel(), _stream(posm))
    return sums


def sym_shares_combine(sums, G, acc):
    """acc (3, count) float32 <- G * the sum over shares, in share order, of sums (shares, 3,
    count) float64: the accelerations from every share's FP64 sums."""
    if sums.dim() != 3:
        raise ValueError('sums must be (shares, 3, count)')
    shares, count = sums.shape[0], sums.shape[2]
    _check(sums, 'sums', torch.float64, (shares, 3, count))
    _check(acc, 'acc', torch.float32, (3, count))
    same_device(sums, acc)
    with on_device(sums):
        _native.call('grav_stage1_sym_shares_combine', sums.data_ptr(), int(shares), int(count),
                     float(G), acc.data_ptr(), _stream(sums))
    return acc


def shard_cfg(n, m):
    """The library's (i_per_lane, j_waves, splits) for a share of m of n bodies
    (splits == 1: owner-computes)."""
    k, s, p = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _native.call('grav_stage1_shard_cfg', int(n), int(m), ctypes.byref(k), ctypes.byref(s),
                 ctypes.byref(p))
    return k.value, s.value, p.value


def stage1_shard_f32(posm, acc, i0, G, eps=0.0, cfg=None, workspace=None):
    """acc (3, m) <- FP32-compensated accelerations of the bodies [i0, i0 + m) of posm (n, 4)
    against all n: one GPU's share when a universe is divided among GPUs by i-range.  Bit for
    bit rows i0 .. i0 + m - 1 of ``stage1_f32(..., 'kahan', (k, s, 256))`` for cfg (k, s, 1),
    or of ``stage1_split_f32(..., cfg)`` for cfg (k, s, splits > 1).  cfg None: ``shard_cfg``'s.
    ``workspace`` (splits > 1): at least ``split_workspace_bytes(m, cfg)`` bytes; by default one
    is kept per (device, m, cfg)."""
    n = posm.shape[0]
    m = acc.shape[1] if acc.dim() == 2 else -1
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(acc, 'acc', torch.float32, (3, m))
    same_device(posm, acc)
    cfg = tuple(cfg) if cfg is not None else shard_cfg(n, m)
    ptr, size = 0, 0
    if cfg[2] > 1:
        if workspace is None:
            workspace = split_workspace(posm, m, cfg)
        same_device(posm, workspace)
        ptr, size = workspace.data_ptr(), workspace.numel()
    with on_device(posm):
        _native.call('grav_stage1_f32_shard', posm.data_ptr(), acc.data_ptr(), n, int(i0), m,
                     float(G), float(eps), int(cfg[0]), int(cfg[1]), int(cfg[2]), ptr, size,
                     _stream(posm))
    return acc


def step_split_f64m(posm, r64, v64, G, eps, dt, cfg=None, workspace=None, out=None):
    """One time step in the j-split form with FP64 masters: ``stage1_split_f32`` then
    ``stage2_f64m``, with the reduce and stage 2 in one kernel (bit for bit the same r64, v64
    and positions; no acc is written).  The new positions (x, y, z, m) go to ``out`` (n, 4),
    a buffer that does not overlap posm, or, with out=None, back into posm."""
    n = posm.shape[0]
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(r64, 'r64', torch.float64, (3, n))
    _check(v64, 'v64', torch.float64, (3, n))
    same_device(posm, r64, v64)
    if out is not None:
        _check(out, 'out', torch.float32, (n, 4))
        same_device(posm, out)
    if workspace is None:
        workspace = split_workspace(posm, n, cfg)
    same_device(posm, workspace)
    k, s, p = cfg if cfg is not None else (0, 0, 0)
    with on_device(posm):
        if out is None:
            _native.call('grav_step_split_f64m', posm.data_ptr(), r64.data_ptr(), v64.data_ptr(),
                         n, float(G), float(eps), float(dt), int(k), int(s), int(p),
                         workspace.data_ptr(), workspace.numel(), _stream(posm))
        else:
            _native.call('grav_step_split_f64m_io', posm.data_ptr(), out.data_ptr(),
                         r64.data_ptr(), v64.data_ptr(), n, float(G), float(eps), float(dt),
                         int(k), int(s), int(p), workspace.data_ptr(), workspace.numel(),
                         _stream(posm))


def stage1_split_f64(posm64, acc64, G, eps=0.0, cfg=None, workspace=None):
    """acc64 (3, n) <- FP64 accelerations of posm64 (n, 4) in the j-split form; cfg =
    (i_per_lane, j_waves, splits) or None, workspace as for ``stage1_split_f32``."""
    n = posm64.shape[0]
    _check(posm64, 'posm64', torch.float64, (n, 4))
    _check(acc64, 'acc64', torch.float64, (3, n))
    same_device(posm64, acc64)
    if workspace is None:
        workspace = split_workspace(posm64, n, cfg)
    same_device(posm64, workspace)
    k, s, p = cfg if cfg is not None else (0, 0, 0)
    with on_device(posm64):
        _native.call('grav_stage1_f64_split', posm64.data_ptr(), acc64.data_ptr(), n, float(G),
                     float(eps), int(k), int(s), int(p), workspace.data_ptr(),
                     workspace.numel(), _stream(posm64))
    return acc64


def release_workspaces():
    _workspaces.clear()


def stage1_f64(posm64, acc64, G, eps=0.0):
    """acc64 (3, n) <- FP64 accelerations of posm64 (n, 4)."""
    n = posm64.shape[0]
    _check(posm64, 'posm64', torch.float64, (n, 4))
    _check(acc64, 'acc64', torch.float64, (3, n))
    same_device(posm64, acc64)
    with on_device(posm64):
        _native.call('grav_stage1_f64', posm64.data_ptr(), acc64.data_ptr(), n, float(G),
                     float(eps), _stream(posm64))
    return acc64


def stage2_f32(posm, vel, acc, dt):
    n = posm.shape[0]
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(vel, 'vel', torch.float32, (3, n))
    _check(acc, 'acc', torch.float32, (3, n))
    same_device(posm, vel, acc)
    with on_device(posm):
        _native.call('grav_stage2_f32', posm.data_ptr(), vel.data_ptr(), acc.data_ptr(), n,
                     float(dt), _stream(posm))


def stage2_f64m(posm, r64, v64, acc, dt):
    n = posm.shape[0]
    _check(posm, 'posm', torch.float32, (n, 4))
    _check(r64, 'r64', torch.float64, (3, n))
    _check(v64, 'v64', torch.float64, (3, n))
    _check(acc, 'acc', torch.float32, (3, n))
    same_device(posm, r64, v64, acc)
    with on_device(posm):
        _native.call('grav_stage2_f64m', posm.data_ptr(), r64.data_ptr(), v64.data_ptr(),
                     acc.data_ptr(), n, float(dt), _stream(posm))


def stage2_f64(posm64, v64, acc64, dt):
    n = posm64.shape[0]
    _check(posm64, 'posm64', torch.float64, (n, 4))
    _check(v64, 'v64', torch.float64, (3, n))
    _check(acc64, 'acc64', torch.float64, (3, n))
    same_device(posm64, v64, acc64)
    with on_device(posm64):
        _native.call('grav_stage2_f64', posm64.data_ptr(), v64.data_ptr(), acc64.data_ptr(), n,
                     float(dt), _stream(posm64))


def accelerations(pos, m, G, eps=0.0, dtype='float32', accum='kahan', cfg=None):
    """Convenience: accelerations (n, 3) of positions pos (n, 3) and masses m (n,) on the
    device of ``pos``; FP32 (compensated by default) or FP64 pair arithmetic."""
    n = pos.shape[0]
    if dtype == 'float32':
        posm = torch.empty((n, 4), dtype=torch.float32, device=pos.device)
        posm[:, :3] = pos
        posm[:, 3] = m
        acc = torch.empty((3, n), dtype=torch.float32, device=pos.device)
        stage1_f32(posm, acc, G, eps, accum, cfg)
    elif dtype == 'float64':
        posm = torch.empty((n, 4), dtype=torch.float64, device=pos.device)
        posm[:, :3] = pos
        posm[:, 3] = m
        acc = torch.empty((3, n), dtype=torch.float64, device=pos.device)
        stage1_f64(posm, acc, G, eps)
    else:
        raise ValueError("dtype must be 'float32' or 'float64'")
    return acc.t()
