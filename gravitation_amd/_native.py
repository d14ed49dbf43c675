"""ctypes binding of libgravhip.so (the C ABI declared in include/gravitation_hip.h).

The reference binds its native stage 1 the same way: ``ctypes.cdll.LoadLibrary`` on a
``.so`` that sits beside the kernel module (src/gravitation/kernel/c1a.py:73-76).  Unlike the
reference's ``void`` functions, every entry point here returns a status code; a non-zero code
raises ``GravError`` carrying ``grav_last_error()``.

There is no fallback: if the library is missing, loading raises ``ImportError`` naming the
build command.
"""

import ctypes
import os

LIB_NAME = 'libgravhip.so'
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

GRAV_OK = 0
GRAV_EINVAL = 1
GRAV_EHIP = 2
GRAV_ENOTSUP = 3

ACCUM_PLAIN = 0
ACCUM_KAHAN = 1

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_i = ctypes.c_int

# name -> argtypes; every function returns int.  Order and types mirror the header.
SIGNATURES = {
    'grav_version': [],
    'grav_device_count': [ctypes.POINTER(_i)],
    'grav_stage1_f32': [_vp, _vp, _i64, _d, _d, _i, _vp],
    'grav_stage1_f32_cfg': [_vp, _vp, _i64, _d, _d, _i, _i, _i, _i, _vp],
    'grav_stage1_f64': [_vp, _vp, _i64, _d, _d, _vp],
    'grav_stage1_sym_workspace': [_i64, _i, _i, ctypes.POINTER(_i64)],
    'grav_stage1_sym_f32': [_vp, _vp, _i64, _d, _d, _i, _i, _vp, _i64, _vp],
    'grav_stage1_sym_f64_workspace': [_i64, _i, _i, ctypes.POINTER(_i64)],
    'grav_stage1_sym_f64': [_vp, _vp, _i64, _d, _d, _i, _i, _vp, _i64, _vp],
    'grav_stage2_f32': [_vp, _vp, _vp, _i64, _d, _vp],
    'grav_stage2_f64m': [_vp, _vp, _vp, _vp, _i64, _d, _vp],
    'grav_stage2_f64': [_vp, _vp, _vp, _i64, _d, _vp],
    'grav_stage2_f32_mirror': [_vp, _vp, _vp, _i64, _d, _vp, _vp],
    'grav_stage2_f64m_mirror': [_vp, _vp, _vp, _vp, _i64, _d, _vp, _vp],
    'grav_stage2_f64_mirror': [_vp, _vp, _vp, _i64, _d, _vp, _vp],
    'grav_host_mirror_alloc': [_i64, ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    'grav_host_mirror_free': [_vp],
    'grav_batched_plan_create': [ctypes.POINTER(_i64), _i, _i, _i, _i, _i,
                                 ctypes.POINTER(_vp)],
    'grav_batched_plan_destroy': [_vp],
    'grav_batched_plan_info': [_vp, ctypes.POINTER(_i64), _i],
    'grav_batched_step_f32': [_vp, _vp, _vp, _vp, _d, _d, _d, _i, _vp, _vp],
    'grav_adjoint_mu': [_vp, _vp, _vp, _i64, _d, _vp],
    'grav_stage1_vjp_f32': [_vp, _vp, _vp, _vp, _i64, _d, _d, _d, _vp],
    'grav_batched_adjoint_f32': [_vp, _vp, _vp, _vp, _vp, _d, _d, _d, _i, _vp],
    'grav_batched_stage1_f32': [_vp, _vp, _vp, _d, _d, _i, _vp],
    'grav_batched_adjoint_mu': [_vp, _vp, _vp, _vp, _d, _vp],
    'grav_batched_sym_prepare': [_vp, _i, _i],
    'grav_batched_sym_workspace': [_vp, ctypes.POINTER(_i64)],
    'grav_batched_stage1_sym_f32': [_vp, _vp, _vp, _d, _d, _vp, _i64, _vp],
    'grav_stage1_split_workspace': [_i64, _i, ctypes.POINTER(_i64)],
    'grav_stage1_f32_split': [_vp, _vp, _i64, _d, _d, _i, _i, _i, _vp, _i64, _vp],
    'grav_stage1_sym_share_plan': [_i64, _i, _i, _i, _i, _vp, _vp],
    'grav_stage1_sym_f32_share': [_vp, _vp, _i64, _d, _i, _i, _i, _i, _vp, _i64, _vp],
    'grav_stage1_sym_shares_combine': [_vp, _i, _i64, _d, _vp, _vp],
    'grav_stage1_vjp_f32_shard': [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _d, _d, _d, _i, _i, _i,
                                  _vp, _i64, _vp],
    'grav_stage1_vjp_shard_cfg': [_i64, _i64, _vp, _vp, _vp],
    'grav_stage1_f32_shard': [_vp, _vp, _i64, _i64, _i64, _d, _d, _i, _i, _i, _vp, _i64, _vp],
    'grav_stage1_shard_cfg': [_i64, _i64, _vp, _vp, _vp],
    'grav_stage1_vjp_f32_split': [_vp, _vp, _vp, _vp, _i64, _d, _d, _d, _i, _i, _i, _vp, _i64,
                                  _vp],
    'grav_stage1_f64_split': [_vp, _vp, _i64, _d, _d, _i, _i, _i, _vp, _i64, _vp],
    'grav_step_split_f64m': [_vp, _vp, _vp, _i64, _d, _d, _d, _i, _i, _i, _vp, _i64, _vp],
    'grav_adjoint_mu_ws': [_vp, _vp, _vp, _i64, _d, _vp, _i64, _vp],
    'grav_batched_stage1_vjp_f32': [_vp, _vp, _vp, _vp, _vp, _d, _d, _d, _i, _vp],
    'grav_stage1_f32_split_noalloc': [_vp, _vp, _i64, _d, _d, _i, _i, _i, _vp],
    'grav_stage1_sym_workspace_min': [_i64, _i, _i, ctypes.POINTER(_i64)],
    'grav_stage1_sym_f64_workspace_min': [_i64, _i, _i, ctypes.POINTER(_i64)],
    'grav_step_split_f64m_io': [_vp, _vp, _vp, _vp, _i64, _d, _d, _d, _i, _i, _i, _vp, _i64,
                                _vp],
    'grav_stage1_vjp_f32_split_chain': [_vp, _vp, _vp, _vp, _vp, _i64, _d, _d, _d, _i, _i, _i,
                                        _i, _vp, _i64, _vp, _vp],
    'grav_backward_pair_workspace': [_i64, _i, _i, ctypes.POINTER(_i64)],
    'grav_backward_pair_split': [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _d, _d,
                                 _d, _i, _i, _i, _i, _i, _vp, _i64, _vp],
}


class GravError(RuntimeError):
    """A libgravhip call returned a non-zero status."""

    def __init__(self, fn, code, msg):
        super().__init__('%s failed (code %d): %s' % (fn, code, msg))
        self.code = code
        self.fn = fn


class GravInvalidArgument(GravError, ValueError):
    """GRAV_EINVAL: the arguments were rejected before any HIP call."""


_lib = None


def load(path=None):
    """Loads (once) and returns the CDLL with argtypes/restype set."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError('%s not found; build it with `python -m gravitation_amd.build` '
                          '(or __graft_entry__.build())' % p)
    lib = ctypes.CDLL(p)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = _i
    lib.grav_last_error.argtypes = []
    lib.grav_last_error.restype = ctypes.c_char_p
    if path is None:
        _lib = lib
    return lib


def call(name, *args):
    """Calls ``name`` and raises GravError on a non-zero status."""
    lib = load()
    return check(name, getattr(lib, name)(*args))


def check(name, rc):
    """Raises GravError for a non-zero status ``rc`` returned by ``name``."""
    if rc != GRAV_OK:
        msg = load().grav_last_error().decode('utf-8', 'replace')
        cls = GravInvalidArgument if rc == GRAV_EINVAL else GravError
        raise cls(name, rc, msg)
    return rc


def device_count():
    c = _i(0)
    call('grav_device_count', ctypes.byref(c))
    return c.value
