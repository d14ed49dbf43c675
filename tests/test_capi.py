"""The C ABI on the CPU: the library loads, exports every symbol include/gravitation_hip.h
declares, binds them with matching arity, and rejects bad arguments before any HIP call."""

import ctypes
import os
import re

import numpy as np
import pytest

from gravitation_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, 'include'))):
        if fn.endswith('.h'):
            src = open(os.path.join(ROOT, 'include', fn)).read()
            src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
            names += re.findall(r'^\s*(?:int|const\s+char\s*\*)\s*(grav_\w+)\s*\(', src, re.M)
    return names


def test_header_declares_the_bound_symbols():
    declared = set(_declared())
    assert 'grav_stage1_f32' in declared and 'grav_last_error' in declared
    assert declared == set(_native.SIGNATURES) | {'grav_last_error'}


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    for name in _declared():
        assert hasattr(lib, name), name


def _arity(name):
    src = ''.join(open(os.path.join(ROOT, 'include', f)).read()
                  for f in os.listdir(os.path.join(ROOT, 'include')) if f.endswith('.h'))
    m = re.search(r'\b%s\s*\(([^)]*)\)' % name, src)
    args = m.group(1).strip()
    return 0 if args in ('', 'void') else args.count(',') + 1


def test_binding_arity_matches_header():
    for name, argtypes in _native.SIGNATURES.items():
        assert _arity(name) == len(argtypes), name


def test_version():
    assert _native.load().grav_version() >= 100


def _expect_einval(name, *args, match):
    with pytest.raises(_native.GravInvalidArgument, match=match):
        _native.call(name, *args)


def test_argument_validation_without_gpu():
    # All of these fail in the validation that precedes any HIP call.  The fake pointers are
    # never dereferenced; the test still refuses to run where a device could be touched.
    if _native.device_count() > 0:
        pytest.skip('fake-pointer validation runs only on a host without a HIP device')
    fake = ctypes.c_void_p(4096)  # never dereferenced: validation rejects first
    _expect_einval('grav_stage1_f32', None, fake, 16, 1.0, 0.0, 1, None, match='posm is NULL')
    _expect_einval('grav_stage1_f32', fake, fake, 0, 1.0, 0.0, 1, None, match='n must be')
    _expect_einval('grav_stage1_f32', fake, fake, (1 << 29) + 1, 1.0, 0.0, 1, None,
                   match='n must be')
    _expect_einval('grav_stage1_f32', ctypes.c_void_p(4100), fake, 16, 1.0, 0.0, 1, None,
                   match='aligned')
    _expect_einval('grav_stage1_f32', fake, fake, 16, 1.0, -1.0, 1, None, match='eps')
    _expect_einval('grav_stage1_f32', fake, fake, 16, float('nan'), 0.0, 1, None, match='G')
    _expect_einval('grav_stage1_f32', fake, fake, 16, 1.0, 0.0, 7, None, match='accum')
    _expect_einval('grav_stage1_f32_cfg', fake, fake, 16, 1.0, 0.0, 1, 3, 1, 256, None,
                   match='i_per_lane')
    _expect_einval('grav_stage1_f32_cfg', fake, fake, 16, 1.0, 0.0, 1, 1, 3, 256, None,
                   match='j_waves')
    _expect_einval('grav_stage1_f32_cfg', fake, fake, 16, 1.0, 0.0, 1, 1, 1, 12, None,
                   match='j_block')
    _expect_einval('grav_stage1_f64', ctypes.c_void_p(4112), fake, 16, 1.0, 0.0, None,
                   match='32-byte')
    _expect_einval('grav_stage2_f32', fake, None, fake, 16, 0.1, None, match='NULL')
    _expect_einval('grav_stage2_f32', fake, fake, fake, 16, float('inf'), None, match='dt')
    _expect_einval('grav_stage2_f64m', ctypes.c_void_p(4104), fake, fake, fake, 16, 0.1, None,
                   match='16-byte')
    _expect_einval('grav_stage2_f64', fake, fake, fake, -5, 0.1, None, match='n must be')
    _expect_einval('grav_stage1_f32_split', fake, fake, 4096, 1.0, 0.0, 4, 8, 4, fake, 1 << 30,
                   None, match='splits')
    _expect_einval('grav_stage1_f32_split', fake, fake, 4096, 1.0, 0.0, 1, 8, 65, fake, 1 << 30,
                   None, match='splits')
    _expect_einval('grav_stage1_f32_split', fake, fake, 4096, 1.0, 0.0, 0, 0, 0, fake, 1000,
                   None, match='workspace')
    _expect_einval('grav_stage1_vjp_f32_split', fake, fake, fake, fake, 4096, 1.0, 0.0, 1.0, 1,
                   3, 4, fake, 1 << 30, None, match='splits')
    _expect_einval('grav_stage1_vjp_f32_split', fake, fake, fake, fake, 4096, 1.0, 0.0, 1.0, 0,
                   0, 0, ctypes.c_void_p(4096 + 64), 1 << 30, None, match='workspace')
    _expect_einval('grav_stage1_f64_split', ctypes.c_void_p(4096 + 16), fake, 4096, 1.0, 0.0,
                   0, 0, 0, fake, 1 << 30, None, match='32-byte')
    _expect_einval('grav_stage1_f64_split', fake, fake, 4096, 1.0, 0.0, 1, 2, 8, fake, 1 << 30,
                   None, match='splits')
    _expect_einval('grav_step_split_f64m', fake, fake, None, 4096, 1.0, 0.0, 1e-3, 0, 0, 0,
                   fake, 1 << 30, None, match='NULL')
    _expect_einval('grav_step_split_f64m', fake, fake, fake, 4096, 1.0, 0.0, float('inf'), 0,
                   0, 0, fake, 1 << 30, None, match='dt')
    _expect_einval('grav_step_split_f64m', fake, fake, fake, 4096, 1.0, 0.0, 1e-3, 1, 8, 0,
                   fake, 1 << 30, None, match='splits')
    _expect_einval('grav_batched_sym_prepare', None, 4, 2, match='plan is NULL')
    _expect_einval('grav_batched_sym_workspace', None, ctypes.byref(ctypes.c_int64()),
                   match='NULL')
    _expect_einval('grav_batched_stage1_sym_f32', None, fake, fake, 1.0, 0.0, fake, 1 << 20,
                   None, match='plan is NULL')


def test_batched_engine_mode_checks_without_gpu():
    from gravitation_amd import batched
    off = [0, 10, 30]
    with pytest.raises(ValueError, match='mode'):
        batched.Engine(off, ('owner', 8, 1, 8, 0, 4))
    with pytest.raises(ValueError, match='symmetric shape'):
        batched.Engine(off, ('perstep-sym', 8, 1, 8, 0, 4, (2, 2)))
    with pytest.raises(ValueError, match='config'):
        batched.Engine(off, ('perstep', 8, 1, 8))


def test_error_is_thread_local_message():
    lib = _native.load()
    rc = lib.grav_stage1_f32(None, None, 16, 1.0, 0.0, 1, None)
    assert rc == _native.GRAV_EINVAL
    assert b'NULL' in lib.grav_last_error()


def _sym_bytes(n, k, w):
    b = ctypes.c_int64(0)
    _native.call('grav_stage1_sym_workspace', n, k, w, ctypes.byref(b))
    return b.value


def _sym_expect(n, ib, elem):
    """(one-slab, minimum) workspace bytes of the symmetric form by its stated layout: rows of
    block pairs cut into segments of L = the largest of 8, 4, 2, 1 that leaves at least 2^18
    segments' worth of pairs (pairs // L >= 2^18), segments counted row by row; per segment
    3 * IB values of i-side partials and L * 3 * IB of j-side ones; block-diagonal sums of
    3 n values (rounded up to 256 bytes) first, and below one slab 24 n bytes of FP64 sums."""
    nb = -(-n // ib)
    pairs = nb * (nb - 1) // 2
    seglen = next(x for x in (8, 4, 2, 1) if pairs // x >= 1 << 18 or x == 1)
    nseg = sum(-(-(nb - 1 - b) // seglen) for b in range(nb - 1))
    seg = 3 * ib * elem * (1 + seglen)
    up = lambda x: -(-x // 256) * 256
    full = up(3 * n * elem) + nseg * seg
    row0 = -(-(nb - 1) // seglen) if nb > 1 else 0
    return full, min(full, up(3 * n * elem) + up(24 * n) + row0 * seg)


def test_symmetric_default_shape_without_gpu():
    # (0, 0) is the library's shape by n: blocks of 512 (4, 2) below 57344 bodies, blocks of
    # 1024 (8, 2) from there
    def both(n, k, w):
        lo = ctypes.c_int64(0)
        _native.call('grav_stage1_sym_workspace_min', n, k, w, ctypes.byref(lo))
        return _sym_bytes(n, k, w), lo.value
    for n in (1, 1000, 24576, 49152, 57343):
        assert both(n, 0, 0) == both(n, 4, 2) == _sym_expect(n, 512, 4)
    for n in (57344, 262144, 1 << 20):
        assert both(n, 0, 0) == both(n, 8, 2) == _sym_expect(n, 1024, 4)
    assert both(262144, 8, 4) == _sym_expect(262144, 2048, 4)
    # at 2^18 one slab (single block pairs) needs about 0.8 GB, the smallest workspace about
    # 16 MB; at 2^22 (segments of 8 block pairs) the form runs in 208 MB
    assert 0.7e9 < _sym_bytes(262144, 0, 0) < 0.9e9
    assert both(1 << 22, 0, 0)[1] < 0.25e9
    for k, w in ((0, 2), (4, 0), (2, 2), (8, 8)):
        with pytest.raises(_native.GravInvalidArgument):
            _sym_bytes(1000, k, w)


def test_symmetric_f64_workspace_without_gpu():
    def f64(n, k, w, fn='grav_stage1_sym_f64_workspace'):
        b = ctypes.c_int64(0)
        _native.call(fn, n, k, w, ctypes.byref(b))
        return b.value

    def both(n, k, w):
        return f64(n, k, w), f64(n, k, w, 'grav_stage1_sym_f64_workspace_min')
    assert both(1000, 0, 0) == both(1000, 2, 2) == _sym_expect(1000, 256, 8)
    assert both(30000, 0, 0) == both(30000, 4, 2) == _sym_expect(30000, 512, 8)
    assert both(262144, 0, 0) == both(262144, 4, 4) == _sym_expect(262144, 1024, 8)
    assert both(262144, 4, 2) == _sym_expect(262144, 512, 8)
    for k, w in ((8, 2), (4, 8), (0, 2)):
        with pytest.raises(_native.GravInvalidArgument):
            f64(1000, k, w)


def test_split_workspace_without_gpu():
    from gravitation_amd import stages
    b = ctypes.c_int64(0)
    _native.call('grav_stage1_split_workspace', 4096, 5, ctypes.byref(b))
    assert b.value == 5 * 4 * 4096 * 8
    # the default covers the largest of the FP32 stage-1, VJP and FP64 defaults: 16 ranges
    assert stages.split_workspace_bytes(2048) == 16 * 32 * 2048
    assert stages.split_workspace_bytes(4096) == 16 * 32 * 4096
    _expect_einval('grav_stage1_split_workspace', 4096, 65, ctypes.byref(b), match='splits')
    stages.check_split_cfg((2, 8, 64))
    stages.check_split_cfg((1, 2, 8))
    stages.check_split_cfg((2, 4, 8), 'float64')
    for bad in ((4, 8, 4), (1, 3, 4), (1, 8, 0), (1, 8)):
        with pytest.raises(ValueError):
            stages.check_split_cfg(bad)
    with pytest.raises(ValueError):
        stages.check_split_cfg((1, 2, 8), 'float64')


def test_chain_io_and_noalloc_validation_without_gpu():
    # the step with separate output, the chained backward step and the allocation-free stage 1
    # reject bad arguments before any HIP call (the fake pointers are never dereferenced)
    if _native.device_count() > 0:
        pytest.skip('fake-pointer validation runs only on a host without a HIP device')
    fake, other = ctypes.c_void_p(4096), ctypes.c_void_p(4096 + 16 * 64)
    ws = ctypes.c_void_p(1 << 20)
    _expect_einval('grav_step_split_f64m_io', fake, ctypes.c_void_p(4096 + 16), fake, fake, 64,
                   1.0, 0.0, 1e-3, 0, 0, 0, ws, 1 << 30, None, match='overlap')
    _expect_einval('grav_step_split_f64m_io', fake, None, fake, fake, 64, 1.0, 0.0, 1e-3,
                   0, 0, 0, ws, 1 << 30, None, match='NULL')
    _expect_einval('grav_step_split_f64m_io', fake, other, fake, fake, 64, 1.0, 0.0, 1e-3,
                   0, 0, 0, ws, 16, None, match='workspace')
    _expect_einval('grav_step_split_f64m_io', fake, other, fake, fake, 64, 1.0, 0.0, 1e-3,
                   3, 8, 4, ws, 1 << 30, None, match='i_per_lane')
    _expect_einval('grav_stage1_vjp_f32_split_chain', fake, fake, fake, fake, fake, 64, 1.0,
                   0.0, 1e-3, 0, 0, 0, 2, ws, 1 << 30, fake, None, match='next_mu')
    _expect_einval('grav_stage1_vjp_f32_split_chain', fake, fake, fake, fake, fake, 64, 1.0,
                   0.0, 1e-3, 0, 0, 0, 1, ws, 1 << 30, None, None, match='NULL')
    _expect_einval('grav_stage1_vjp_f32_split_chain', fake, fake, fake, fake, fake, 64, 1.0,
                   0.0, 1e-3, 0, 0, 0, 1, ws, 64, fake, None, match='workspace')
    # the stage-2 calls that write a host mirror: the mirror address is checked first
    _expect_einval('grav_stage2_f32_mirror', fake, fake, fake, 64, 1e-3, None, None,
                   match='mirror')
    _expect_einval('grav_stage2_f64m_mirror', fake, fake, fake, fake, 64, 1e-3,
                   ctypes.c_void_p(4096 + 4), None, match='mirror')
    _expect_einval('grav_stage2_f64_mirror', fake, fake, fake, 64, 1e-3,
                   ctypes.c_void_p(4096 + 16), None, match='mirror')
    h, d = ctypes.c_void_p(), ctypes.c_void_p()
    _expect_einval('grav_host_mirror_alloc', 0, ctypes.byref(h), ctypes.byref(d), match='bytes')
    # the paired backward step: shapes, workspace, overlapping positions
    b = ctypes.c_int64(0)
    _native.call('grav_backward_pair_workspace', 4096, 8, 16, ctypes.byref(b))
    assert b.value == (8 + 16) * 4 * 4096 * 8
    _native.call('grav_backward_pair_workspace', 4096, 0, 0, ctypes.byref(b))
    assert b.value == (8 + 8) * 4 * 4096 * 8
    args = [fake] * 10 + [64, 1.0, 0.0, 1e-3]
    _expect_einval('grav_backward_pair_split', *args, 1, 8, 1, 8, 2, ws, 1 << 30, None,
                   match='j_waves')
    _expect_einval('grav_backward_pair_split', *args, 3, 8, 1, 8, 8, ws, 1 << 30, None,
                   match='j_waves')
    _expect_einval('grav_backward_pair_split', *args, 0, 0, 0, 0, 0, ws, 1024, None,
                   match='workspace')
    over = [fake] * 6 + [fake, ctypes.c_void_p(4096 + 32)] + [fake] * 2 + [64, 1.0, 0.0, 1e-3]
    _expect_einval('grav_backward_pair_split', *over, 0, 0, 0, 0, 0, ws, 1 << 30, None,
                   match='overlap')
    _expect_einval('grav_backward_pair_split', *([fake] * 5 + [None] + [fake] * 4),
                   64, 1.0, 0.0, 1e-3, 0, 0, 0, 0, 0, ws, 1 << 30, None, match='NULL')
    _expect_einval('grav_stage1_f32_split_noalloc', fake, fake, (1 << 20) + 1, 1.0, 0.0, 1, 8,
                   1, None, match='array')
    _expect_einval('grav_stage1_f32_split_noalloc', fake, fake, 64, 1.0, -1.0, 0, 0, 0, None,
                   match='eps')
    lo = ctypes.c_int64(0)
    _native.call('grav_stage1_sym_workspace_min', 1 << 18, 0, 0, ctypes.byref(lo))
    _expect_einval('grav_stage1_sym_f32', fake, fake, 1 << 18, 1.0, 1e-3, 0, 0, ws,
                   lo.value - 256, None, match='at least')


def test_shard_validation_without_gpu():
    # one GPU's share of an i-range divided universe: the range, the shapes of both forms and
    # the j-split workspace are checked before any HIP call
    if _native.device_count() > 0:
        pytest.skip('fake-pointer validation runs only on a host without a HIP device')
    fake, ws = ctypes.c_void_p(4096), ctypes.c_void_p(1 << 20)
    n = 1000
    for i0, m in ((900, 101), (0, 0), (-1, 10), (1000, 1)):
        _expect_einval('grav_stage1_f32_shard', fake, fake, n, i0, m, 1.0, 1e-3, 1, 8, 1,
                       None, 0, None, match='shard')
    _expect_einval('grav_stage1_f32_shard', fake, fake, n, 0, 100, 1.0, 1e-3, 4, 16, 1, None,
                   0, None, match='j_waves')
    _expect_einval('grav_stage1_f32_shard', fake, fake, n, 0, 100, 1.0, 1e-3, 4, 8, 4, ws,
                   1 << 30, None, match='i_per_lane')
    # the j-split partials of a share are splits x 4 x m doubles, not of all n bodies
    need = 4 * 4 * 100 * 8
    _expect_einval('grav_stage1_f32_shard', fake, fake, n, 0, 100, 1.0, 1e-3, 1, 8, 4, ws,
                   need - 8, None, match='workspace')
    _expect_einval('grav_stage1_f32_shard', fake, fake, n, 0, 100, 1.0, 1e-3, 1, 8, 4, None,
                   need, None, match='workspace')
    k, s, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    for nn, mm in ((1 << 18, 1 << 15), (1 << 18, 1 << 18), (4096, 512), (100, 1)):
        _native.call('grav_stage1_shard_cfg', nn, mm, ctypes.byref(k), ctypes.byref(s),
                     ctypes.byref(p))
        if p.value == 1:
            assert k.value in (1, 2, 4, 8) and s.value in (1, 2, 4, 8, 16)
        else:
            assert k.value in (1, 2) and s.value in (1, 2, 4, 8) and 2 <= p.value <= 64
    _expect_einval('grav_stage1_shard_cfg', 100, 101, ctypes.byref(k), ctypes.byref(s),
                   ctypes.byref(p), match='m <= n')


def test_symmetric_share_plan_without_gpu():
    # the shares' rows of block pairs (a run of rows and their partner rows nb - 2 - beta)
    # cover every row once with about equal pair counts, exactly equal when the row couples
    # divide evenly, and their diagonal blocks cover every block once
    rows = (ctypes.c_int * 6)()
    b = (ctypes.c_int64 * 2)()
    for n, shares in ((1 << 18, 8), (1 << 18, 3), (70001, 5), (1000, 3), (100, 4),
                      (1 << 20, 64), (5000, 400), (1 << 18, 255), (3 << 16, 7)):
        ib = 1024 if n >= 57344 else 512
        nb = -(-n // ib)
        seen = np.zeros(max(nb - 1, 0), dtype=int)
        diag_end = 0
        loads = []
        for k in range(shares):
            _native.call('grav_stage1_sym_share_plan', n, k, shares, 0, 0, rows, b)
            r0, r1, s0, s1, d0, d1 = list(rows)
            assert r1 >= r0 and s1 >= s0 and d0 == diag_end and d1 >= d0
            diag_end = d1
            assert b[0] >= b[1] >= 256
            mine = list(range(r0, r1)) + list(range(s0, s1))
            seen[mine] += 1
            loads.append(sum(nb - 1 - beta for beta in mine))
        assert np.all(seen == 1) and diag_end == nb
        total = nb * (nb - 1) // 2
        assert sum(loads) == total
        # at most the fair part plus one row couple (nb pairs)
        assert max(loads) <= total / shares + nb
        units = nb // 2
        if nb % 2 == 1 and units % shares == 0:
            assert len(set(loads)) == 1
    for share, shares in ((3, 3), (-1, 2), (0, 0), (0, 65537)):
        _expect_einval('grav_stage1_sym_share_plan', 4096, share, shares, 0, 0, rows, b,
                       match='share')
    if _native.device_count() > 0:
        return
    fake, ws = ctypes.c_void_p(4096), ctypes.c_void_p(1 << 20)
    _native.call('grav_stage1_sym_share_plan', 1 << 18, 2, 8, 0, 0, rows, b)
    _expect_einval('grav_stage1_sym_f32_share', fake, fake, 1 << 18, 1e-3, 2, 8, 0, 0, ws,
                   b[1] - 256, None, match='at least')
    _expect_einval('grav_stage1_sym_f32_share', fake, ctypes.c_void_p(4096 + 4), 1 << 18, 1e-3,
                   2, 8, 0, 0, ws, b[0], None, match='misaligned')
    _expect_einval('grav_stage1_sym_f32_share', fake, fake, 1 << 18, 1e-3, 8, 8, 0, 0, ws, b[0],
                   None, match='share')
    _expect_einval('grav_stage1_sym_shares_combine', fake, 0, 100, 1.0, fake, None,
                   match='shares')
    _expect_einval('grav_stage1_sym_shares_combine', fake, 2, 0, 1.0, fake, None, match='count')


def test_vjp_shard_validation_without_gpu():
    if _native.device_count() > 0:
        pytest.skip('fake-pointer validation runs only on a host without a HIP device')
    fake, ws = ctypes.c_void_p(4096), ctypes.c_void_p(1 << 20)
    for i0, m in ((900, 101), (0, 0), (-1, 10)):
        _expect_einval('grav_stage1_vjp_f32_shard', fake, fake, fake, fake, 1000, i0, m, 1.0,
                       1e-3, 1e-3, 1, 8, 1, None, 0, None, match='shard')
    _expect_einval('grav_stage1_vjp_f32_shard', fake, fake, fake, fake, 1000, 0, 100, 1.0, 1e-3,
                   1e-3, 4, 8, 1, None, 0, None, match='j_waves')
    _expect_einval('grav_stage1_vjp_f32_shard', fake, fake, fake, fake, 1000, 0, 100, 1.0, 1e-3,
                   1e-3, 1, 8, 4, ws, 4 * 4 * 100 * 8 - 8, None, match='workspace')
    k, s, p = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    for nn, mm in ((1 << 18, 1 << 15), (4096, 512), (100, 1), (1 << 20, 1 << 20)):
        _native.call('grav_stage1_vjp_shard_cfg', nn, mm, ctypes.byref(k), ctypes.byref(s),
                     ctypes.byref(p))
        assert k.value in (1, 2) and s.value in (1, 2, 4, 8) and 1 <= p.value <= 64
