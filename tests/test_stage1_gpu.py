"""GPU parity of stage 1 and stage 2 against the FP64 oracle (through the C ABI).

Bars: FP32 pair terms (compensated) within 1e-5 normwise relative error of the FP64 oracle on
the identical FP32 inputs; FP64 within 1e-12; stage 2 bit-identical to the same NumPy
operations; repeated runs bit-identical.
"""

import numpy as np
import pytest
import torch

from gravitation_amd import _native, stages
from gravitation_amd.lib import scenarios as S
from oracle import oracle as O

from conftest import oracle_rows

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-12


def _posm(pos, m, dev, dtype=torch.float32):
    p = torch.empty((pos.shape[0], 4), dtype=dtype)
    p[:, :3] = torch.from_numpy(np.asarray(pos, dtype=np.float64))
    p[:, 3] = torch.from_numpy(np.asarray(m, dtype=np.float64))
    return p.to(dev)


def _run32(posm, G, eps, accum='kahan', cfg=None):
    n = posm.shape[0]
    acc = torch.full((3, n), float('nan'), dtype=torch.float32, device=posm.device)
    stages.stage1_f32(posm, acc, G, eps, accum, cfg)
    torch.cuda.synchronize()
    return acc.cpu().numpy().T.astype(np.float64)


def _run64(posm64, G, eps):
    n = posm64.shape[0]
    acc = torch.full((3, n), float('nan'), dtype=torch.float64, device=posm64.device)
    stages.stage1_f64(posm64, acc, G, eps)
    torch.cuda.synchronize()
    return acc.cpu().numpy().T


def _inputs(p):
    h = p.cpu().numpy().astype(np.float64)
    return h[:, :3], h[:, 3]


@pytest.mark.parametrize('n', [1, 2, 3, 7, 63, 64, 65, 255, 257, 1000, 1024, 4097])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_stage1_f32_ragged_sizes(cuda, n, eps):
    print('stage1_f32 n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    pos, m = _inputs(posm)
    ref = O.accel_f64(pos, m, 0.9, eps)
    for accum in ('kahan', 'plain'):
        a = _run32(posm, 0.9, eps, accum)
        assert np.all(np.isfinite(a))
        if n == 1:
            assert np.all(a == 0.0)
        else:
            assert O.rel_err(a, ref) < TOL32, accum


@pytest.mark.parametrize('n', [1, 2, 65, 1000, 4096])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_stage1_f64_ragged_sizes(cuda, n, eps):
    print('stage1_f64 n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 7)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda, torch.float64)
    pos, m = _inputs(posm)
    a = _run64(posm, 0.9, eps)
    ref = oracle_rows(pos, m, 0.9, eps)
    if n == 1:
        assert np.all(a == 0.0)
    else:
        assert O.rel_err(a, ref) < TOL64
        np.testing.assert_allclose(a, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_stage1_f32_config_sweep(cuda):
    n = 3001
    rng = np.random.default_rng(3)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    pos, m = _inputs(posm)
    for eps in (0.0, 1e-2):
        ref = O.accel_f64(pos, m, 1.0, eps)
        for k in (1, 2, 4, 8):
            for s in (1, 2, 4, 8, 16):
                if s == 16 and k > 2:
                    continue   # 16 wavefronts only with 1 or 2 bodies per lane (LDS)
                for jb in (8, 64, 256):
                    print('cfg k=%d s=%d jb=%d eps=%g' % (k, s, jb, eps), flush=True)
                    a = _run32(posm, 1.0, eps, 'kahan', (k, s, jb))
                    assert O.rel_err(a, ref) < TOL32, (k, s, jb)


def test_c1_uniform_cube(cuda):
    # 1024 bodies, cube side 1e3, masses U[1e-3, 1e-1], G = 6.6743e-11, FP32.
    r, v, m = S.uniform_cube(1024, seed=0)
    posm = _posm(r, m, cuda)
    pos, mm = _inputs(posm)
    a = _run32(posm, 6.6743e-11, 0.0)
    assert O.rel_err(a, O.accel_f64(pos, mm, 6.6743e-11)) < 1e-6


@pytest.mark.parametrize('logn', [14, 16, 18, 20])
def test_c2_plummer(cuda, logn):
    # Plummer sphere, m = 1/N, G = 1, eps = 1e-3; FP32-compensated and FP64 vs the oracle.
    n = 1 << logn
    print('plummer N=2^%d' % logn, flush=True)
    r, v, m = S.plummer(n, seed=logn)
    posm = _posm(r, m, cuda)
    pos, mm = _inputs(posm)
    rows = None if logn <= 14 else np.sort(np.random.default_rng(1).choice(n, 512, False))
    ref = oracle_rows(pos, mm, 1.0, 1e-3, rows)
    a = _run32(posm, 1.0, 1e-3)
    a = a if rows is None else a[rows]
    err32 = O.rel_err(a, ref)
    print('  fp32 kahan rel err %.3e' % err32, flush=True)
    assert err32 < TOL32
    posm64 = _posm(pos, mm, cuda, torch.float64)
    a64 = _run64(posm64, 1.0, 1e-3)
    a64 = a64 if rows is None else a64[rows]
    err64 = O.rel_err(a64, ref)
    print('  fp64 rel err %.3e' % err64, flush=True)
    assert err64 < TOL64


def test_momentum_conservation_full_size(cuda):
    # sum_i m_i a_i = 0 exactly in exact arithmetic; a size-independent check at N = 2^18.
    n = 1 << 18
    r, v, m = S.plummer(n, seed=11)
    posm = _posm(r, m, cuda)
    a = _run32(posm, 1.0, 1e-3)
    mm = posm[:, 3].double().cpu().numpy()
    f = mm[:, None] * a
    assert np.abs(f.sum(axis=0)).max() < 1e-6 * np.abs(f).sum(axis=0).max()


def test_deterministic_bitwise(cuda):
    n = 20000
    r, v, m = S.plummer(n, seed=2)
    posm = _posm(r, m, cuda)
    a1 = _run32(posm, 1.0, 1e-3)
    a2 = _run32(posm, 1.0, 1e-3)
    assert np.array_equal(a1, a2)


def test_stage2_bitwise_vs_numpy(cuda):
    n = 10007
    rng = np.random.default_rng(9)
    dt = 1e-3
    r = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    a = rng.normal(size=(n, 3)).astype(np.float32)
    m = rng.uniform(0.1, 1.0, n)
    # FP32 state
    posm = _posm(r, m, cuda)
    vel = torch.from_numpy(v.T.astype(np.float32).copy()).to(cuda)
    acc = torch.from_numpy(a.T.copy()).to(cuda)
    r32 = posm[:, :3].cpu().numpy()
    m32 = posm[:, 3].cpu().numpy().copy()
    stages.stage2_f32(posm, vel, acc, dt)
    torch.cuda.synchronize()
    rn, vn = O.stage2(r32, v.astype(np.float32), a, np.float32(dt))
    assert np.array_equal(vel.cpu().numpy().T, vn)
    assert np.array_equal(posm[:, :3].cpu().numpy(), rn)
    assert np.array_equal(posm[:, 3].cpu().numpy(), m32)
    # FP64 masters with FP32 accelerations
    posm = _posm(r, m, cuda)
    r64 = torch.from_numpy(r.T.copy()).to(cuda)
    v64 = torch.from_numpy(v.T.copy()).to(cuda)
    stages.stage2_f64m(posm, r64, v64, acc, dt)
    torch.cuda.synchronize()
    rn, vn = O.stage2(r, v, a.astype(np.float64), dt)
    assert np.array_equal(v64.cpu().numpy().T, vn)
    assert np.array_equal(r64.cpu().numpy().T, rn)
    assert np.array_equal(posm[:, :3].cpu().numpy(), rn.astype(np.float32))
    assert np.array_equal(posm[:, 3].cpu().numpy(), m32)
    # all FP64
    posm64 = _posm(r, m, cuda, torch.float64)
    v64 = torch.from_numpy(v.T.copy()).to(cuda)
    a64 = torch.from_numpy(a.T.astype(np.float64).copy()).to(cuda)
    stages.stage2_f64(posm64, v64, a64, dt)
    torch.cuda.synchronize()
    assert np.array_equal(v64.cpu().numpy().T, vn)
    assert np.array_equal(posm64[:, :3].cpu().numpy(), rn)


def test_python_side_checks(cuda):
    posm = torch.zeros((8, 4), dtype=torch.float32, device=cuda)
    with pytest.raises(TypeError):
        stages.stage1_f32(posm.double(), torch.zeros((3, 8), device=cuda), 1.0)
    with pytest.raises(ValueError):
        stages.stage1_f32(posm, torch.zeros((8, 3), device=cuda), 1.0)
    with pytest.raises(ValueError):
        stages.stage1_f32(posm.cpu(), torch.zeros((3, 8)), 1.0)
    with pytest.raises(ValueError):
        stages.stage1_f32(posm, torch.zeros((3, 8), device=cuda), 1.0, accum='fast')
    with pytest.raises(_native.GravInvalidArgument):
        stages.stage1_f32(posm, torch.zeros((3, 8), device=cuda), 1.0, cfg=(3, 1, 256))
    with pytest.raises(_native.GravInvalidArgument, match='16'):
        stages.stage1_f32(posm, torch.zeros((3, 8), device=cuda), 1.0, cfg=(4, 16, 256))


@pytest.mark.parametrize('name', ['c1_lib1.npz', 'solarsystem_lib1.npz'])
def test_against_reference_golden_vectors(cuda, name):
    # HIP FP32 stage 1 against the reference's own C stage 1 on the same FP32 inputs (the
    # committed fixtures of tests/golden/make_golden.py).
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', name))
    posm = _posm(d['pos'].astype(np.float64), d['m'].astype(np.float64), cuda)
    for accum in ('kahan', 'plain'):
        a = _run32(posm, float(d['G']), 0.0, accum)
        assert O.rel_err(a, d['acc']) < 2e-6
    if name.startswith('solarsystem'):
        assert abs(a[1, 1] - 5.931412e-03) < 1e-8


@pytest.mark.parametrize('n', [1, 2, 65, 511, 512, 1023, 1024, 3000, 4096, 10001])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_stage1_symmetric_vs_oracle(cuda, n, eps):
    # blocks of 512 (4, 2) ... 2048 (8, 4) bodies: sizes below, at and across block edges
    print('sym n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 3)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    pos, m = _inputs(posm)
    ref = O.accel_f64(pos, m, 0.9, eps)
    for cfg in ((4, 2), (4, 4), (8, 2), (8, 4)):
        acc = torch.full((3, n), float('nan'), dtype=torch.float32, device=cuda)
        stages.stage1_sym_f32(posm, acc, 0.9, eps, cfg)
        torch.cuda.synchronize()
        a = acc.cpu().numpy().T.astype(np.float64)
        assert np.all(np.isfinite(a)), cfg
        if n == 1:
            assert np.all(a == 0.0)
        else:
            assert O.rel_err(a, ref) < TOL32, cfg


@pytest.mark.parametrize('logn', [16, 18])
def test_stage1_symmetric_plummer_and_determinism(cuda, logn):
    n = 1 << logn
    r, v, m = S.plummer(n, seed=logn)
    posm = _posm(r, m, cuda)
    pos, mm = _inputs(posm)
    rows = np.sort(np.random.default_rng(2).choice(n, 512, False))
    ref = oracle_rows(pos, mm, 1.0, 1e-3, rows)
    acc = torch.empty((3, n), dtype=torch.float32, device=cuda)
    stages.stage1_sym_f32(posm, acc, 1.0, 1e-3, (8, 4))
    a1 = acc.cpu().numpy().T.astype(np.float64)
    err = O.rel_err(a1[rows], ref)
    print('  sym plummer 2^%d rel err %.3e' % (logn, err), flush=True)
    assert err < TOL32
    stages.stage1_sym_f32(posm, acc, 1.0, 1e-3, (8, 4))
    assert np.array_equal(acc.cpu().numpy().T.astype(np.float64), a1)
    f = mm[:, None] * a1
    assert np.abs(f.sum(axis=0)).max() < 1e-6 * np.abs(f).sum(axis=0).max()


def test_dpp_wave_rotate_lane_layout(cuda):
    # The symmetric kernel's correctness rests on 64 wave_ror:1 steps being the identity and
    # visiting every lane once; both hold for any rotation direction.  Checked end to end:
    # two bodies in different blocks see each other exactly once (a = G m / d^2 each).
    n = 2048 * 2
    posm = torch.zeros((n, 4), dtype=torch.float32)
    posm[:, 0] = torch.arange(n, dtype=torch.float32) * 1e4    # far apart, no other pairs matter
    posm[:, 3] = 0.0
    posm[5, 3] = 2.0
    posm[n - 7, 3] = 3.0
    posm = posm.to(cuda)
    acc = torch.zeros((3, n), dtype=torch.float32, device=cuda)
    stages.stage1_sym_f32(posm, acc, 1.0, 0.0, (8, 4))
    a = acc.cpu().numpy()
    d = float(posm[n - 7, 0] - posm[5, 0])
    np.testing.assert_allclose(a[0, 5], 3.0 / d ** 2, rtol=1e-6)
    np.testing.assert_allclose(a[0, n - 7], -2.0 / d ** 2, rtol=1e-6)


def test_symmetric_default_shape_is_the_explicit_one(cuda):
    # cfg None (the library's (0, 0)) is (4, 2) below 57344 bodies and (8, 2) from there,
    # bit for bit.
    for n, cfg in ((30000, (4, 2)), (57344, (8, 2))):
        r, _, m = S.plummer(n, seed=12)
        posm = _posm(r, m, cuda)
        a = torch.empty((3, n), dtype=torch.float32, device=cuda)
        b = torch.empty_like(a)
        stages.stage1_sym_f32(posm, a, 1.0, 1e-3)
        stages.stage1_sym_f32(posm, b, 1.0, 1e-3, cfg)
        torch.cuda.synchronize()
        assert torch.equal(a, b), n
    stages.release_workspaces()


@pytest.mark.parametrize('n', [1, 2, 65, 255, 256, 257, 511, 512, 513, 3000, 10001])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_stage1_symmetric_f64_vs_oracle(cuda, n, eps):
    # blocks of 256 (2, 2) ... 1024 (4, 4) bodies: sizes below, at and across block edges
    print('sym f64 n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 5)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda, torch.float64)
    pos, m = _inputs(posm)
    ref = O.accel_f64(pos, m, 0.9, eps)
    own = torch.empty((3, n), dtype=torch.float64, device=cuda)
    stages.stage1_f64(posm, own, 0.9, eps)
    for cfg in ((2, 2), (2, 4), (4, 2), (4, 4)):
        acc = torch.full((3, n), float('nan'), dtype=torch.float64, device=cuda)
        stages.stage1_sym_f64(posm, acc, 0.9, eps, cfg)
        torch.cuda.synchronize()
        a = acc.cpu().numpy().T
        assert np.all(np.isfinite(a)), cfg
        if n == 1:
            assert np.all(a == 0.0)
        else:
            assert O.rel_err(a, ref) < TOL64, cfg
            assert O.rel_err(a, own.cpu().numpy().T) < TOL64, cfg
    stages.release_workspaces()


@pytest.mark.parametrize('logn', [16, 18])
def test_stage1_symmetric_f64_plummer_and_determinism(cuda, logn):
    n = 1 << logn
    r, v, m = S.plummer(n, seed=logn)
    posm = _posm(r, m, cuda, torch.float64)
    pos, mm = _inputs(posm)
    rows = np.sort(np.random.default_rng(4).choice(n, 256, False))
    ref = oracle_rows(pos, mm, 1.0, 1e-3, rows)
    acc = torch.empty((3, n), dtype=torch.float64, device=cuda)
    stages.stage1_sym_f64(posm, acc, 1.0, 1e-3)
    a1 = acc.cpu().numpy().T
    err = O.rel_err(a1[rows], ref)
    print('  sym f64 plummer 2^%d rel err %.3e' % (logn, err), flush=True)
    assert err < TOL64
    stages.stage1_sym_f64(posm, acc, 1.0, 1e-3)
    assert np.array_equal(acc.cpu().numpy().T, a1)
    f = mm[:, None] * a1
    assert np.abs(f.sum(axis=0)).max() < 1e-12 * np.abs(f).sum(axis=0).max()
    stages.release_workspaces()


@pytest.mark.parametrize('dtype, n', [(torch.float32, 30000), (torch.float64, 5000)])
def test_tune_stage1_form(cuda, dtype, n):
    from gravitation_amd import tune
    r, _, m = S.plummer(n, seed=13)
    posm = _posm(r, m, cuda, dtype)
    (form, cfg), table = tune.tune_stage1_form(posm, 1.0, 1e-3, reps=1)
    forms = {f for (f, _), _ in table}
    assert forms == {'owner', 'split', 'symmetric'} and form in forms
    acc = torch.empty((3, n), dtype=dtype, device=cuda)
    if form == 'owner':
        (stages.stage1_f64(posm, acc, 1.0, 1e-3) if dtype == torch.float64
         else stages.stage1_f32(posm, acc, 1.0, 1e-3, 'kahan', cfg))
    elif form == 'split':
        (stages.stage1_split_f64 if dtype == torch.float64 else stages.stage1_split_f32)(
            posm, acc, 1.0, 1e-3, cfg)
    else:
        (stages.stage1_sym_f64 if dtype == torch.float64 else stages.stage1_sym_f32)(
            posm, acc, 1.0, 1e-3, cfg)
    pos, mm = _inputs(posm)
    rows = np.arange(0, n, n // 97)
    ref = oracle_rows(pos, mm, 1.0, 1e-3, rows)
    a = acc.cpu().numpy().T.astype(np.float64)
    assert O.rel_err(a[rows], ref) < (TOL64 if dtype == torch.float64 else TOL32)
    stages.release_workspaces()


@pytest.mark.parametrize('n', [1, 2, 63, 64, 65, 1000, 1024, 4096, 5000])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_stage1_split_vs_oracle_and_owner(cuda, n, eps):
    # every i-tile also split over j-ranges: shapes with empty ranges (64 splits of small n),
    # uneven ranges and the library's default
    print('split n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 9)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    pos, m = _inputs(posm)
    ref = O.accel_f64(pos, m, 0.9, eps)
    own = _run32(posm, 0.9, eps)
    for cfg in (None, (1, 1, 1), (1, 4, 8), (2, 2, 3), (1, 8, 64), (2, 8, 5)):
        acc = torch.full((3, n), float('nan'), dtype=torch.float32, device=cuda)
        stages.stage1_split_f32(posm, acc, 0.9, eps, cfg)
        a = acc.cpu().numpy().T.astype(np.float64)
        assert np.all(np.isfinite(a)), cfg
        if n == 1:
            assert np.all(a == 0.0), cfg
            continue
        assert O.rel_err(a, ref) < TOL32, cfg
        assert O.rel_err(a, own) < TOL32, cfg
        stages.stage1_split_f32(posm, acc, 0.9, eps, cfg)
        assert np.array_equal(acc.cpu().numpy().T.astype(np.float64), a), cfg


@pytest.mark.parametrize('n', [1, 2, 63, 65, 1000, 4096, 5000])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_stage1_split_f64_vs_oracle(cuda, n, eps):
    print('split f64 n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 19)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda, torch.float64)
    pos, m = _inputs(posm)
    ref = O.accel_f64(pos, m, 0.9, eps)
    for cfg in (None, (1, 4, 1), (1, 8, 5), (2, 4, 16), (2, 8, 64)):
        acc = torch.full((3, n), float('nan'), dtype=torch.float64, device=cuda)
        stages.stage1_split_f64(posm, acc, 0.9, eps, cfg)
        a = acc.cpu().numpy().T
        assert np.all(np.isfinite(a)), cfg
        if n == 1:
            assert np.all(a == 0.0), cfg
            continue
        assert O.rel_err(a, ref) < TOL64, cfg
        stages.stage1_split_f64(posm, acc, 0.9, eps, cfg)
        assert np.array_equal(acc.cpu().numpy().T, a), cfg


def test_owner_default_shape_is_the_explicit_one(cuda):
    # grav_stage1_f32's default is (i_per_lane, j_waves) = (1, 16) from 8192 bodies, (1, 8)
    # below, bit for bit
    for n, cfg in ((1000, (1, 8, 256)), (8191, (1, 8, 256)), (8192, (1, 16, 256)),
                   (1 << 16, (1, 16, 256))):
        r, _, m = S.plummer(n, seed=21)
        posm = _posm(r, m, cuda)
        assert np.array_equal(_run32(posm, 1.0, 1e-3), _run32(posm, 1.0, 1e-3, 'kahan', cfg)), n


def test_split_default_shape_is_the_explicit_one(cuda):
    # grav_stage1_f32_split's default is (i_per_lane, j_waves, splits) = (1, 8, 16) from 8192
    # bodies, (1, 8, 8) below, bit for bit
    for n, cfg in ((4000, (1, 8, 8)), (8191, (1, 8, 8)), (8192, (1, 8, 16)),
                   (33001, (1, 8, 16))):
        r, _, m = S.plummer(n, seed=23)
        posm = _posm(r, m, cuda)
        out = []
        for c in (None, cfg):
            acc = torch.full((3, n), float('nan'), dtype=torch.float32, device=cuda)
            stages.stage1_split_f32(posm, acc, 1.0, 1e-3, c)
            out.append(acc.cpu().numpy())
        assert np.array_equal(out[0], out[1]), n


@pytest.mark.parametrize('n', [65, 4096, 5000])
def test_fused_split_step_is_bitwise_split_plus_stage2(cuda, n):
    # grav_step_split_f64m == grav_stage1_f32_split + grav_stage2_f64m, bit for bit
    r, v, m = S.plummer(n, seed=n + 2)
    v = np.random.default_rng(n).normal(0, 0.1, v.shape)
    for cfg in (None, (2, 4, 3)):
        states = []
        for fused in (False, True):
            posm = _posm(r, m, cuda)
            r64 = torch.from_numpy(np.ascontiguousarray(r.T)).to(cuda)
            v64 = torch.from_numpy(np.ascontiguousarray(v.T)).to(cuda)
            acc = torch.empty((3, n), dtype=torch.float32, device=cuda)
            for _ in range(3):
                if fused:
                    stages.step_split_f64m(posm, r64, v64, 1.0, 1e-3, 1e-3, cfg)
                else:
                    stages.stage1_split_f32(posm, acc, 1.0, 1e-3, cfg)
                    stages.stage2_f64m(posm, r64, v64, acc, 1e-3)
            states.append([x.cpu().numpy() for x in (posm, r64, v64)])
        for a, b in zip(*states):
            assert np.array_equal(a, b), (n, cfg)


@pytest.mark.parametrize('n', [8193, 12288, 16384, 33001])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_stage1_split_large_vs_oracle(cuda, n, eps):
    # the j-split form where hip1 takes it by default (2048 <= n < 65536), eps = 0 (self term
    # masked in the j-range that holds the tile's own bodies) included: the default shape
    # switches to (1, 8, 16) at 8192; rows sampled across tiles plus every tile edge of the
    # first and last tile
    print('split large n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 29)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    pos, m = _inputs(posm)
    rows = np.unique(np.concatenate([rng.choice(n, 600, replace=False),
                                     [0, 63, 64, 127, 128, n - 129, n - 65, n - 64, n - 1]]))
    ref = oracle_rows(pos, m, 0.9, eps, rows)
    for cfg in (None, (1, 8, 8), (2, 8, 16), (2, 4, 7)):
        acc = torch.full((3, n), float('nan'), dtype=torch.float32, device=cuda)
        stages.stage1_split_f32(posm, acc, 0.9, eps, cfg)
        a = acc.cpu().numpy().T.astype(np.float64)
        assert np.all(np.isfinite(a)), cfg
        assert O.rel_err(a[rows], ref) < TOL32, cfg
    stages.release_workspaces()


@pytest.mark.parametrize('dtype,n,cfg', [(torch.float32, 3000, (4, 2)),
                                         (torch.float32, 100000, (4, 2)),
                                         (torch.float32, 140000, (4, 2)),
                                         (torch.float32, 70001, (8, 4)),
                                         (torch.float64, 20000, (2, 2)),
                                         (torch.float64, 70000, (4, 4))])
def test_stage1_symmetric_slabs(cuda, dtype, n, cfg):
    # Row segments of 1, 2, 4 and 8 block pairs (the library's choice by the block-pair count:
    # 2 at 100000 bodies in blocks of 512, 8 at 140000) and workspaces from the one-slab size
    # down to the minimum (one row of block pairs per slab, FP64 sums between slabs): every
    # size meets the oracle's bar, agrees with the one-slab result to the FP64 fold's rounding
    # and repeats bit for bit.
    f64 = dtype == torch.float64
    print('sym slabs n=%d cfg=%s %s' % (n, cfg, 'f64' if f64 else 'f32'), flush=True)
    rng = np.random.default_rng(n + 41)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda, dtype)
    pos, m = _inputs(posm)
    rows = np.unique(np.concatenate([rng.choice(n, 400, replace=False), [0, n - 1]]))
    eps = 0.0 if n == 70001 else 1e-3
    ref = oracle_rows(pos, m, 0.9, eps, rows)
    of = stages.sym_f64_workspace_bytes if f64 else stages.sym_workspace_bytes
    run = stages.stage1_sym_f64 if f64 else stages.stage1_sym_f32
    full, least = of(n, cfg), of(n, cfg, minimum=True)
    assert least < full
    sizes = [full, least, (least + full) // 2 // 256 * 256, least + 4096]
    first = None
    for size in sizes:
        ws = torch.empty((size,), dtype=torch.uint8, device=cuda)
        acc = torch.full((3, n), float('nan'), dtype=dtype, device=cuda)
        run(posm, acc, 0.9, eps, cfg, ws)
        a = acc.cpu().numpy().T.astype(np.float64)
        assert np.all(np.isfinite(a)), size
        assert O.rel_err(a[rows], ref) < (TOL64 if f64 else TOL32), size
        if first is None:
            first = a
        else:
            assert O.rel_err(a, first) < (1e-14 if f64 else 1e-6), size
        run(posm, acc, 0.9, eps, cfg, ws)
        assert np.array_equal(acc.cpu().numpy().T.astype(np.float64), a), size
        del ws
    small = torch.empty((least - 256,), dtype=torch.uint8, device=cuda)
    with pytest.raises(ValueError):
        run(posm, torch.empty((3, n), dtype=dtype, device=cuda), 0.9, eps, cfg, small)
    stages.release_workspaces()


@pytest.mark.parametrize('n', [65, 4096, 5000, 20000])
def test_split_step_io_is_bitwise_in_place(cuda, n):
    # grav_step_split_f64m_io (new positions to another buffer, as the recomputed segments of
    # the adjoint write one record slot from the previous) == grav_step_split_f64m in place,
    # bit for bit, over three steps; a partly overlapping output is refused
    r, v, m = S.plummer(n, seed=n + 7)
    v = np.random.default_rng(n).normal(0, 0.1, v.shape)
    for cfg in (None, (2, 4, 3), (1, 8, 16)):
        print('io step n=%d cfg=%s' % (n, cfg), flush=True)
        states = []
        for io in (False, True):
            posm = _posm(r, m, cuda)
            alt = torch.full_like(posm, float('nan'))
            r64 = torch.from_numpy(np.ascontiguousarray(r.T)).to(cuda)
            v64 = torch.from_numpy(np.ascontiguousarray(v.T)).to(cuda)
            for _ in range(3):
                if io:
                    stages.step_split_f64m(posm, r64, v64, 1.0, 1e-3, 1e-3, cfg, out=alt)
                    posm, alt = alt, posm
                else:
                    stages.step_split_f64m(posm, r64, v64, 1.0, 1e-3, 1e-3, cfg)
            states.append([x.cpu().numpy() for x in (posm, r64, v64)])
        for a, b in zip(*states):
            assert np.array_equal(a, b), (n, cfg)
    big = torch.zeros((n + 1, 4), dtype=torch.float32, device=cuda)
    with pytest.raises(_native.GravError, match='overlap'):
        stages.step_split_f64m(big[:n], r64, v64, 1.0, 1e-3, 1e-3, None, out=big[1:])


@pytest.mark.parametrize('n, eps', [(1000, 0.0), (5000, 1e-3), (70001, 0.0), (70001, 1e-3)])
def test_stage1_shard_rows_equal_whole_universe(cuda, n, eps):
    # a universe divided by i-range into uneven shares, some starting inside an i-tile: each
    # share's rows equal the whole-universe launch of the same shape bit for bit, in the
    # owner-computes and the j-split form; the default shape against the oracle
    print('shard n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 41)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    cuts = sorted(set([0, n, n // 8, n // 2] + [int(x) for x in rng.integers(1, n, 5)]))
    shares = list(zip(cuts[:-1], cuts[1:]))
    for cfg in ((2, 16, 1), (1, 8, 1), (4, 4, 1), (1, 8, 8), (2, 8, 16), (2, 4, 7)):
        whole = torch.full((3, n), float('nan'), dtype=torch.float32, device=cuda)
        if cfg[2] == 1:
            stages.stage1_f32(posm, whole, 0.9, eps, 'kahan', (cfg[0], cfg[1], 256))
        else:
            stages.stage1_split_f32(posm, whole, 0.9, eps, cfg)
        w = whole.cpu()
        assert torch.isfinite(w).all(), cfg
        for i0, i1 in shares:
            acc = torch.full((3, i1 - i0), float('nan'), dtype=torch.float32, device=cuda)
            stages.stage1_shard_f32(posm, acc, i0, 0.9, eps, cfg)
            assert torch.equal(acc.cpu(), w[:, i0:i1]), (cfg, i0, i1)
    pos, m = _inputs(posm)
    i0, i1 = n // 8, n // 2
    rows = np.unique(np.concatenate([rng.integers(i0, i1, 300), [i0, i0 + 63, i1 - 1]]))
    ref = oracle_rows(pos, m, 0.9, eps, rows)
    acc = torch.full((3, i1 - i0), float('nan'), dtype=torch.float32, device=cuda)
    stages.stage1_shard_f32(posm, acc, i0, 0.9, eps)
    a = acc.cpu().numpy().T.astype(np.float64)
    assert O.rel_err(a[rows - i0], ref) < TOL32
    stages.release_workspaces()


@pytest.mark.parametrize('n, eps, shares', [(1000, 0.0, 3), (70001, 1e-3, 8), (70001, 0.0, 3),
                                            (140000, 1e-3, 5)])
def test_stage1_symmetric_shares_combine_to_the_whole(cuda, n, eps, shares):
    # the symmetric form divided among GPUs by rows of block pairs: the shares' FP64 sums,
    # combined in share order, against the one-GPU form (same terms, another FP64 order), the
    # oracle and themselves (deterministic); one share is the one-GPU form bit for bit; a
    # share in slabs of its minimum workspace against its one-slab result
    print('sym shares n=%d eps=%g shares=%d' % (n, eps, shares), flush=True)
    rng = np.random.default_rng(n + 53)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    whole = torch.empty((3, n), dtype=torch.float32, device=cuda)
    stages.stage1_sym_f32(posm, whole, 0.9, eps)
    sums = torch.full((shares, 3, n), float('nan'), dtype=torch.float64, device=cuda)
    for k in range(shares):
        stages.stage1_sym_f32_share(posm, sums[k], k, shares, eps)
    acc = torch.empty((3, n), dtype=torch.float32, device=cuda)
    stages.sym_shares_combine(sums, 0.9, acc)
    a, w = acc.cpu().numpy().T.astype(np.float64), whole.cpu().numpy().T.astype(np.float64)
    assert np.all(np.isfinite(a))
    assert O.rel_err(a, w) < 1e-7
    pos, m = _inputs(posm)
    rows = np.unique(np.concatenate([rng.integers(0, n, 400), [0, n - 1]]))
    assert O.rel_err(a[rows], oracle_rows(pos, m, 0.9, eps, rows)) < TOL32
    again = torch.empty_like(acc)
    for k in range(shares):
        stages.stage1_sym_f32_share(posm, sums[k], k, shares, eps)
    stages.sym_shares_combine(sums, 0.9, again)
    assert torch.equal(again, acc)
    one = torch.empty((1, 3, n), dtype=torch.float64, device=cuda)
    stages.stage1_sym_f32_share(posm, one[0], 0, 1, eps)
    stages.sym_shares_combine(one, 0.9, again)
    assert torch.equal(again, whole)
    k = shares // 2
    _, full, low = stages.sym_share_plan(n, k, shares)
    small = torch.empty((3, n), dtype=torch.float64, device=cuda)
    ws = torch.empty((low,), dtype=torch.uint8, device=cuda)
    stages.stage1_sym_f32_share(posm, small, k, shares, eps, workspace=ws)
    ref = sums[k].cpu().numpy()
    if np.any(ref):
        assert O.rel_err(small.cpu().numpy(), ref) < 1e-13
    else:   # a share with no rows and no diagonal block adds nothing
        assert not np.any(small.cpu().numpy())
    stages.release_workspaces()
