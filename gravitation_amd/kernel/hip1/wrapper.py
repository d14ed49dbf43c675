"""The ``hip1`` kernel: a ``universe_base`` subclass whose stages run as HIP kernels on an
MI355X.

Contract (src/gravitation/kernel/_base_.py:64-177): the constructor, ``add_object``,
``start``/``step``/``stop`` and the per-body iterable are the base class's.  This class only
overrides the four hooks the reference lets a kernel override: ``start_kernel``,
``step_stage1``, ``step_stage2`` and ``stop_kernel``.

Options (keyword arguments of the constructor; they travel through ``**kwargs`` into
``self._meta`` as the base class provides, src/gravitation/kernel/_base_.py:93):

* ``eps`` (float, default 0.0): Plummer softening length; 0 is the reference formula.
* ``accum`` ('kahan' | 'plain', default 'kahan'): FP32 stage-1 accumulation.
* ``stage2_dtype`` ('float32' | 'float64', default = ``dtype``): precision of the position and
  velocity masters updated by stage 2.  With dtype='float32' and stage2_dtype='float64' the
  pair terms are FP32 and r, v accumulate in FP64.
* ``graph`` (bool, default False): replay the two stages as HIP graphs captured at start();
  the results are those of the eager path, bit for bit.  Off by
  default: on ROCm 7 one graph replay costs about 20 us against about 10 us for the direct
  launch of the same kernel, so small universes step slower with graphs (57 against 46 us at
  256 bodies) and large ones gain 1-8 %.
* ``sync_host`` (bool, default True): refresh the host views after every step, so that
  ``pm._r`` / ``pm._v`` hold the new state when ``step()`` returns (README.md:50 of the
  reference requires it): the stage-2 kernel stores the new state into the mapped host mirror
  as well, and step() waits for the stream.  With False the state stays on the device until
  ``sync_to_host()`` is called.
* ``device`` (default 'cuda'): torch device of the state.
* ``stage1`` ('auto' | 'tune' | 'owner' | 'split' | 'symmetric', default 'auto'): stage-1
  form.
  'owner' sums every body's acceleration in one lane (N^2 pair evaluations); 'split' is
  'owner' with each i-tile also split over j-ranges whose FP64 partials one more launch sums
  (it fills the GPU with small universes; a device workspace of 32 B per body and range,
  held from start() to stop()); 'symmetric' evaluates each unordered pair once and
  updates both bodies, as the reference's CPU kernels do (N^2/2 pair evaluations, a device
  workspace held from start() to stop(): the size that does every block pair at once (about
  0.5 GB at 2^18 bodies in FP32, twice that in FP64) when it fits a quarter of the free
  device memory, else that quarter, the form then working through its rows in slabs).
  'auto' takes the fastest for N: with dtype='float32' and accum='kahan' 'owner' below
  SPLIT_FROM bodies, 'split' up to SYMMETRIC_FROM, 'symmetric' from there (see
  ``gravitation_amd.stages``); with dtype='float64' likewise from SPLIT_F64_FROM and
  SYMMETRIC_F64_FROM.  'tune' times every form and launch shape on this universe's state at
  start() (``gravitation_amd.tune.tune_stage1_form``, a few calls each) and keeps the
  fastest; a universe of the same size, precision and eps = 0-ness tuned earlier in the
  process reuses that result.  It needs accum='kahan' in FP32.
* ``stage1_cfg`` (tuple or None): launch shape, (i_per_lane, j_waves, j_block) for FP32
  'owner', (i_per_lane, j_waves, splits) for 'split', (i_chunks_per_wave, waves) for
  'symmetric'; None uses the library's default.

The device state is the truth.  ``pm._r`` and ``pm._v`` are row views of one pinned, mapped
host mirror of it (the rebinding pattern of src/gravitation/kernel/np2.py:70-75 and
pc2.py:124-129), refreshed after each step; writing into them does not change the simulation.  ``pm._a`` views an array
of zeros, the state the reference's stage 2 leaves behind (src/gravitation/kernel/
_base_.py:62); the last stage-1 result is ``self.acc`` on the device.
"""

import numpy as np

from .._base_ import universe_base


class _MirrorMemory:
    """Owns one grav_host_mirror_alloc buffer; freed when the last array over it goes."""

    def __init__(self, nbytes):
        import ctypes
        from gravitation_amd import _native
        h, d = ctypes.c_void_p(), ctypes.c_void_p()
        _native.call('grav_host_mirror_alloc', int(nbytes), ctypes.byref(h), ctypes.byref(d))
        self.host, self.device = h.value, d.value

    def __del__(self):
        try:
            from gravitation_amd import _native
            _native.call('grav_host_mirror_free', self.host)
        except Exception:   # interpreter shutdown: the process frees it
            pass


def _host_mirror(count, dtype):
    """(array, device address): a numpy array of ``count`` elements of the torch dtype
    ``dtype`` in pinned, mapped, coherent host memory, and the address kernels write it
    through.  The memory lives as long as any array or view over it."""
    import ctypes
    import torch
    ct = ctypes.c_double if dtype == torch.float64 else ctypes.c_float
    mem = _MirrorMemory(count * ctypes.sizeof(ct))
    buf = (ct * count).from_address(mem.host)
    buf._owner = mem
    return np.frombuffer(buf, dtype=np.float64 if ct is ctypes.c_double else np.float32), \
        mem.device


class universe(universe_base):

    def start_kernel(self):
        import torch

        # absolute import: the package also runs as src/gravitation/kernel/hip1/ of the
        # reference tree, where ``.._base_`` is the reference's own base class
        from gravitation_amd import _native, stages

        _native.load()  # fails loudly without the HIP library
        if not torch.cuda.is_available():
            raise RuntimeError('hip1 needs a HIP device; none is visible')
        meta = self._meta
        self._stages = stages
        self._eps = float(meta.get('eps', 0.0))
        self._accum = meta.get('accum', 'kahan')
        stages.accum_code(self._accum)
        if self._dtype not in ('float32', 'float64'):
            raise ValueError("dtype must be 'float32' or 'float64'")
        self._stage2_dtype = meta.get('stage2_dtype', self._dtype)
        if self._stage2_dtype not in ('float32', 'float64'):
            raise ValueError("stage2_dtype must be 'float32' or 'float64'")
        if self._dtype == 'float64' and self._stage2_dtype != 'float64':
            raise ValueError("dtype='float64' needs stage2_dtype='float64'")
        self._sync_host = bool(meta.get('sync_host', True))
        self._cfg = meta.get('stage1_cfg', None)
        self._form = meta.get('stage1', 'auto')
        if self._form not in ('auto', 'tune', 'owner', 'split', 'symmetric'):
            raise ValueError("stage1 must be 'auto', 'tune', 'owner', 'split' or 'symmetric'")
        if (self._form in ('symmetric', 'split', 'tune') and self._accum != 'kahan'
                and self._dtype != 'float64'):
            raise ValueError("stage1=%r always uses compensated accumulation" % self._form)
        if self._form in ('auto', 'tune') and self._cfg is not None:
            raise ValueError("stage1_cfg needs stage1='owner', 'split' or 'symmetric'")
        dev = torch.device(meta.get('device', 'cuda'))
        self._device = dev
        if self._form == 'auto':
            free, _ = torch.cuda.mem_get_info(dev)
            self._form = stages.auto_form(len(self), self._accum, free, self._dtype)
        n = len(self)
        if n < 1:
            raise ValueError('hip1 needs at least one body')
        # 1, 2 or 3 dimensions, as the reference's dimension-generic kernels allow
        # (src/gravitation/kernel/py1.py:57-64, np1.py:61): the missing coordinates are held
        # at exactly 0, which adds exactly 0 to every pair term and keeps them 0 in stage 2.
        dim = len(self._mass_list[0]._r)
        if dim not in (1, 2, 3):
            raise ValueError('hip1 simulates 1, 2 or 3 dimensions (got %d)' % dim)
        if any(len(pm._r) != dim or len(pm._v) != dim for pm in self._mass_list):
            raise ValueError('every body needs r and v of %d components' % dim)
        self._dim = dim

        r = np.zeros((n, 3), dtype=np.float64)
        v = np.zeros((n, 3), dtype=np.float64)
        m = np.empty((n,), dtype=np.float64)
        for k, pm in enumerate(self._mass_list):
            r[k, :dim] = pm._r
            v[k, :dim] = pm._v
            m[k] = pm._m

        if self._dtype == 'float64':
            self._mode = 'f64'
        elif self._stage2_dtype == 'float64':
            self._mode = 'f64m'
        else:
            self._mode = 'f32'
        f64 = torch.float64
        f32 = torch.float32
        # What the host views show lives in one device buffer, `_dev_state`, so that refreshing
        # them after a step is one contiguous copy: positions (x, y, z, m) rows then the SoA
        # velocities in the 'f32' and 'f64' modes, the FP64 masters r then v in 'f64m'.
        if self._mode == 'f64m':
            self._dev_state = torch.empty((6 * n,), dtype=f64, device=dev)
            self.r64 = self._dev_state[:3 * n].view(3, n)
            self.vel = self._dev_state[3 * n:].view(3, n)
            self.r64.copy_(torch.from_numpy(r.T.copy()))
            self.vel.copy_(torch.from_numpy(v.T.copy()))
            self.posm = torch.empty((n, 4), dtype=f32, device=dev)
        else:
            dt_ = f64 if self._mode == 'f64' else f32
            self._dev_state = torch.empty((7 * n,), dtype=dt_, device=dev)
            self.posm = self._dev_state[:4 * n].view(n, 4)
            self.vel = self._dev_state[4 * n:].view(3, n)
            self.vel.copy_(torch.from_numpy(v.T.copy()))
        self.posm[:, :3] = torch.from_numpy(r).to(dev, self.posm.dtype)
        self.posm[:, 3] = torch.from_numpy(m).to(dev, self.posm.dtype)
        self.acc = torch.zeros((3, n), dtype=self.posm.dtype, device=dev)

        if self._form == 'tune':
            # every form and shape timed on this universe's own state (stage 1 only reads it)
            from gravitation_amd import tune
            (self._form, self._cfg), _ = tune.tune_stage1_form(self.posm, self._G, self._eps)
            stages.release_workspaces()

        # The symmetric and split forms' workspace belongs to this universe and goes with it
        # at stop().
        self._ws = None
        if self._form == 'symmetric':
            size = stages.sym_workspace_size(n, self._cfg, dtype=self._dtype, device=dev)
            self._ws = torch.empty((max(256, size),), dtype=torch.uint8, device=dev)
        elif self._form == 'split':
            stages.check_split_cfg(self._cfg, self._dtype)
            need = stages.split_workspace_bytes(n, self._cfg)
            self._ws = torch.empty((max(256, need),), dtype=torch.uint8, device=dev)

        # The host mirror of `_dev_state`, laid out the same: mapped, coherent pinned memory
        # that stage 2 writes directly when sync_host is on.  The per-body views are rows of
        # the (n, 4) positions or of the transposes of the (3, n) SoA arrays; they keep the
        # mirror alive after stop().
        hs, self._mirror = _host_mirror(self._dev_state.numel(), self._dev_state.dtype)
        self._host = hs
        if self._mode == 'f64m':
            rview, vview = hs[:3 * n].reshape(3, n).T, hs[3 * n:].reshape(3, n).T
        else:
            rview, vview = hs[:4 * n].reshape(n, 4), hs[4 * n:].reshape(3, n).T
        self._host_a = np.zeros((n, 3), dtype=hs.dtype, order='F')
        self.sync_to_host()
        for k, pm in enumerate(self._mass_list):
            pm._r = rview[k, :dim]
            pm._v = vview[k, :dim]
            pm._a = self._host_a[k, :dim]

        self._prepare()
        self._graphs = None
        if bool(meta.get('graph', False)):
            try:
                self._graphs = self._capture()
            except RuntimeError as e:
                # a capture the runtime refuses leaves the eager path, which is equivalent
                import warnings
                torch.cuda.synchronize(dev)
                self._graphs = None
                warnings.warn('hip1: graph capture failed (%s); stepping eagerly' % e)

    def _copy_to_host(self):
        import torch
        torch.from_numpy(self._host).copy_(self._dev_state)

    def sync_to_host(self):
        """Copies the device r, v into the host views and waits for the copy."""
        import torch
        self._copy_to_host()
        torch.cuda.current_stream(self._device).synchronize()

    def _prepare(self):
        """The stage calls as (C entry point, arguments) with this universe's fixed buffers.
        start_kernel creates those buffers with the shapes, dtypes and device the stages need,
        so a step skips the torch wrappers' per-call checks and calls the C ABI directly (on
        the universe's device, every status checked); ``_stage1`` / ``_stage2`` are the same
        calls through the wrappers, used for graph capture."""
        from gravitation_amd import _native
        lib = _native.load()
        n = len(self)
        p, a, v = self.posm.data_ptr(), self.acc.data_ptr(), self.vel.data_ptr()
        G, eps, T = float(self._G), float(self._eps), float(self._T)
        ws = (self._ws.data_ptr(), self._ws.numel()) if self._ws is not None else None
        cfg = self._cfg if self._cfg is not None else (0, 0)
        if self._mode == 'f64':
            if self._form == 'symmetric':
                c1 = ('grav_stage1_sym_f64', (p, a, n, G, eps, cfg[0], cfg[1]) + ws)
            elif self._form == 'split':
                k, sw, ns = self._cfg if self._cfg is not None else (0, 0, 0)
                c1 = ('grav_stage1_f64_split', (p, a, n, G, eps, int(k), int(sw), int(ns)) + ws)
            else:
                c1 = ('grav_stage1_f64', (p, a, n, G, eps))
            c2 = (('grav_stage2_f64_mirror', (p, v, a, n, T, self._mirror)) if self._sync_host
                  else ('grav_stage2_f64', (p, v, a, n, T)))
        else:
            if self._form == 'symmetric':
                c1 = ('grav_stage1_sym_f32', (p, a, n, G, eps, cfg[0], cfg[1]) + ws)
            elif self._form == 'split':
                k, sw, ns = self._cfg if self._cfg is not None else (0, 0, 0)
                c1 = ('grav_stage1_f32_split', (p, a, n, G, eps, int(k), int(sw), int(ns)) + ws)
            elif self._cfg is not None:
                k, sw, jb = self._cfg
                c1 = ('grav_stage1_f32_cfg', (p, a, n, G, eps, self._stages.accum_code(self._accum),
                                              int(k), int(sw), int(jb)))
            else:
                c1 = ('grav_stage1_f32', (p, a, n, G, eps, self._stages.accum_code(self._accum)))
            r64 = self.r64.data_ptr() if self._mode == 'f64m' else None
            if self._mode == 'f64m' and self._sync_host:
                c2 = ('grav_stage2_f64m_mirror', (p, r64, v, a, n, T, self._mirror))
            elif self._mode == 'f64m':
                c2 = ('grav_stage2_f64m', (p, r64, v, a, n, T))
            elif self._sync_host:
                c2 = ('grav_stage2_f32_mirror', (p, v, a, n, T, self._mirror))
            else:
                c2 = ('grav_stage2_f32', (p, v, a, n, T))
        self._direct = [(getattr(lib, name), name, args) for name, args in (c1, c2)]

    def _call(self, k):
        import torch
        fn, name, args = self._direct[k]
        with self._stages.on_device(self._device):
            rc = fn(*args, torch.cuda.current_stream(self._device).cuda_stream)
        if rc:
            from gravitation_amd import _native
            _native.check(name, rc)

    def _stage1(self):
        s = self._stages
        if self._mode == 'f64':
            if self._form == 'symmetric':
                s.stage1_sym_f64(self.posm, self.acc, self._G, self._eps, self._cfg, self._ws)
            elif self._form == 'split':
                s.stage1_split_f64(self.posm, self.acc, self._G, self._eps, self._cfg, self._ws)
            else:
                s.stage1_f64(self.posm, self.acc, self._G, self._eps)
        elif self._form == 'symmetric':
            s.stage1_sym_f32(self.posm, self.acc, self._G, self._eps, self._cfg, self._ws)
        elif self._form == 'split':
            s.stage1_split_f32(self.posm, self.acc, self._G, self._eps, self._cfg, self._ws)
        else:
            s.stage1_f32(self.posm, self.acc, self._G, self._eps, self._accum, self._cfg)

    def _stage2(self):
        s = self._stages
        if self._sync_host:
            # the kernel that also writes the host mirror, called as in a direct step
            self._call(1)
        elif self._mode == 'f64':
            s.stage2_f64(self.posm, self.vel, self.acc, self._T)
        elif self._mode == 'f64m':
            s.stage2_f64m(self.posm, self.r64, self.vel, self.acc, self._T)
        else:
            s.stage2_f32(self.posm, self.vel, self.acc, self._T)

    def _capture(self):
        """HIP graphs of the two stages (the second writes the host mirror when sync_host is
        set), captured on a side stream; their arguments are this universe's fixed buffers.
        They hold kernels only."""
        import torch
        dev = self._device
        cur = torch.cuda.current_stream(dev)
        side = torch.cuda.Stream(dev)
        side.wait_stream(cur)
        g1, g2 = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            with torch.cuda.graph(g1, stream=side):
                self._stage1()
            with torch.cuda.graph(g2, stream=side):
                self._stage2()
        cur.wait_stream(side)
        return g1, g2

    def step_stage1(self):
        if self._graphs is not None:
            self._graphs[0].replay()
        else:
            self._call(0)

    def step_stage2(self):
        import torch
        if self._graphs is not None:
            self._graphs[1].replay()
        else:
            self._call(1)   # with sync_host, the stage-2 kernel also writes the host mirror
        if self._sync_host:
            torch.cuda.current_stream(self._device).synchronize()

    def stop_kernel(self):
        import torch
        torch.cuda.current_stream(self._device).synchronize()
        self.sync_to_host()
        self._graphs = None
        for name in ('posm', 'vel', 'acc', 'r64', '_ws', '_dev_state'):
            if hasattr(self, name):
                delattr(self, name)
