"""ctypes bindings of the compiled reference stage 1 and of the C oracle -- TEST INFRASTRUCTURE
ONLY (imported by tests/ and bench.py's cpu_baseline leg).

* ``oracle/_ref/lib1.so``: src/gravitation/kernel/_lib1_/lib.c (single-thread FP32, flattened
  unique-pair loop, Newton's third law, lib.c:52-126), called through the reference's own
  ``struct univ`` layout (lib.c:37-49), as src/gravitation/kernel/c1a.py:65-91 does.
* ``oracle/_ref/lib4.so``: src/gravitation/kernel/_lib4_/lib.c (SSE + OpenMP, lib.c:76-380),
  with the extended struct of lib.c:50-66 / c4a.py:68-90.  Its pair term uses
  ``_mm_rsqrt_ps`` (about 12 bits, lib.c:268), so it agrees with the exact formula only to
  about 1e-3; its segmentation (lib.c:76-144) needs at least 2 threads and N divisible by 4.
* ``oracle/c/liboracle.so``: the FP64 compensated all-pairs oracle (oracle/c/oracle.c).
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, '_ref')
LIB1 = os.path.join(REF_DIR, 'lib1.so')
LIB4 = os.path.join(REF_DIR, 'lib4.so')
LIBORACLE = os.path.join(HERE, 'c', 'liboracle.so')

_fp = ctypes.POINTER(ctypes.c_float)
_lp = ctypes.POINTER(ctypes.c_long)


class _univ1(ctypes.Structure):
    # struct univ of src/gravitation/kernel/_lib1_/lib.c:37-49
    _fields_ = [('X', _fp), ('Y', _fp), ('Z', _fp), ('AX', _fp), ('AY', _fp), ('AZ', _fp),
                ('M', _fp), ('G', ctypes.c_float), ('N', ctypes.c_long)]


class _univ4(ctypes.Structure):
    # struct univ of src/gravitation/kernel/_lib4_/lib.c:50-66
    _fields_ = _univ1._fields_ + [
        ('AXmp', _fp), ('AYmp', _fp), ('AZmp', _fp), ('j_min', _lp), ('j_max', _lp),
        ('seg_len', ctypes.c_long), ('OPENMP_threadsmax', ctypes.c_long)]


def available(name='lib1'):
    return os.path.exists({'lib1': LIB1, 'lib4': LIB4, 'oracle': LIBORACLE}[name])


_cache = {}


def _load(path):
    if path not in _cache:
        if not os.path.exists(path):
            raise FileNotFoundError('%s missing: run `make -C oracle`' % path)
        _cache[path] = ctypes.CDLL(path)
    return _cache[path]


def _f32ptr(a):
    return a.ctypes.data_as(_fp)


class RefLib1:
    """Reference C stage 1, single thread (lib1)."""

    def __init__(self, pos, m, G):
        lib = _load(LIB1)
        self._fn = lib.step_stage1
        self._fn.argtypes = [ctypes.POINTER(_univ1)]
        self._fn.restype = None
        pos = np.asarray(pos, dtype=np.float32)
        self.n = pos.shape[0]
        self.cols = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
        self.acc = [np.zeros(self.n, dtype=np.float32) for _ in range(3)]
        self.m = np.ascontiguousarray(m, dtype=np.float32)
        self.u = _univ1(*[_f32ptr(c) for c in self.cols + self.acc + [self.m]],
                        ctypes.c_float(G), self.n)

    def accel(self):
        for a in self.acc:
            a[:] = 0.0
        self._fn(ctypes.byref(self.u))
        return np.stack(self.acc, axis=1)


class RefLib4:
    """Reference C stage 1, SSE + OpenMP (lib4), on ``threads`` >= 2 OpenMP threads."""

    def __init__(self, pos, m, G, threads):
        if threads < 2:
            raise ValueError('lib4 segmentation needs >= 2 threads (lib.c:76-144)')
        pos = np.asarray(pos, dtype=np.float32)
        n = pos.shape[0]
        if n % 4 or n < 16 * threads:
            raise ValueError('lib4 needs N divisible by 4 and N >= 16*threads')
        gomp = ctypes.CDLL('libgomp.so.1')
        gomp.omp_set_num_threads(ctypes.c_int(threads))
        lib = _load(LIB4)
        self._seg = lib.step_stage1_segmentation
        self._fn = lib.step_stage1
        for f in (self._seg, self._fn):
            f.argtypes = [ctypes.POINTER(_univ4)]
            f.restype = None
        self.n = n
        self.threads = threads
        self.cols = [np.ascontiguousarray(pos[:, k]) for k in range(3)]
        self.acc = [np.zeros(n, dtype=np.float32) for _ in range(3)]
        self.m = np.ascontiguousarray(m, dtype=np.float32)
        self.mp = [np.zeros(n * threads, dtype=np.float32) for _ in range(3)]
        self.u = _univ4(*[_f32ptr(c) for c in self.cols + self.acc + [self.m]],
                        ctypes.c_float(G), n, *[_f32ptr(c) for c in self.mp],
                        None, None, 0, 0)
        self._seg(ctypes.byref(self.u))
        if self.u.OPENMP_threadsmax != threads:
            raise RuntimeError('OpenMP gave %d threads, wanted %d'
                               % (self.u.OPENMP_threadsmax, threads))

    def accel(self):
        for a in self.acc:
            a[:] = 0.0
        self._fn(ctypes.byref(self.u))
        return np.stack(self.acc, axis=1)


def oracle_accel_c(pos, m, G, eps=0.0, rows=None):
    """FP64 compensated all-pairs accelerations from the C oracle (OpenMP over rows)."""
    lib = _load(LIBORACLE)
    fn = lib.oracle_accel_f64
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [dp, dp, ctypes.c_long, ctypes.c_double, ctypes.c_double, _lp, ctypes.c_long,
                   dp]
    fn.restype = None
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    m = np.ascontiguousarray(m, dtype=np.float64)
    n = pos.shape[0]
    if rows is None:
        rr = None
        nrows = n
    else:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        rr = rows.ctypes.data_as(_lp)
        nrows = rows.shape[0]
    out = np.empty((nrows, 3), dtype=np.float64)
    fn(pos.ctypes.data_as(dp), m.ctypes.data_as(dp), n, G, eps, rr, nrows,
       out.ctypes.data_as(dp))
    return out


def oracle_vjp_c(pos, m, mu, G, eps=0.0):
    """(J_r^T mu (n, 3), J_m^T mu (n,)) from the C oracle (OpenMP over rows), the formula of
    oracle.vjp_accel_f64."""
    lib = _load(LIBORACLE)
    fn = lib.oracle_vjp_f64
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [dp, dp, dp, ctypes.c_long, ctypes.c_double, ctypes.c_double, dp, dp]
    fn.restype = None
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    m = np.ascontiguousarray(m, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    n = pos.shape[0]
    jr = np.empty((n, 3), dtype=np.float64)
    jm = np.empty((n,), dtype=np.float64)
    fn(pos.ctypes.data_as(dp), m.ctypes.data_as(dp), mu.ctypes.data_as(dp), n, G, eps,
       jr.ctypes.data_as(dp), jm.ctypes.data_as(dp))
    return jr, jm
