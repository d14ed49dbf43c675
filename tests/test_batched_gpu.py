"""Batched universes (one fused launch for all universes and steps) against the FP64 oracle
per universe, against the plain per-universe loop, across every launch shape, under hipGraph
replay, and with the trajectory record."""

import numpy as np
import pytest
import torch

from gravitation_amd import _native, batched, tune
from gravitation_amd.lib import scenarios as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 5, 63, 64, 65, 127, 128, 129, 300, 511, 1000, 2048]


def _batch_inputs(sizes, seed=0, sigma=0.1):
    rng = np.random.default_rng(seed)
    rs, vs, ms = [], [], []
    for n in sizes:
        r, v, m = S.random_universe(int(n), rng, sigma)
        rs.append(r)
        vs.append(v)
        ms.append(m)
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    return np.concatenate(rs), np.concatenate(vs), np.concatenate(ms), off


def _oracle(r, v, m, off, G, eps, dt, steps):
    rr, vv = np.empty_like(r), np.empty_like(v)
    for u in range(off.shape[0] - 1):
        a, b = off[u], off[u + 1]
        ru, vu, _ = O.simulate(r[a:b], v[a:b], m[a:b], G, eps, dt, steps, dtype=np.float32)
        rr[a:b], vv[a:b] = ru, vu
    return rr, vv


def _check_vs_oracle(batch, r0, v0, m, off, steps, tol=2e-5):
    r, v = batch.state()
    rr, vv = _oracle(r0, v0, m, off, batch.G, batch.eps, batch.dt, steps)
    # the change over the run carries the physics; compare it normwise
    assert O.rel_err(v - v0, vv - v0) < tol
    assert O.rel_err(r - r0, rr - r0) < tol


@pytest.mark.parametrize('eps', [1e-3, 0.0])
def test_batched_vs_oracle(cuda, eps):
    r, v, m, off = _batch_inputs(SIZES, seed=1)
    b = batched.Batch(r, v, m, off, G=1.0, eps=eps, dt=1e-3, device=cuda)
    b.step(5)
    _check_vs_oracle(b, r, v, m, off, 5)


def test_batched_every_launch_shape(cuda):
    r, v, m, off = _batch_inputs(SIZES, seed=2)
    rr, vv = _oracle(r, v, m, off, 1.0, 1e-3, 1e-3, 3)
    shapes = [(w, kt, sm, g) for w in (1, 4, 8)
              for kt, sm in ((1, 1), (1, 2), (1, 4), (1, 8), (2, 1), (2, 2), (2, 4), (4, 1),
                             (4, 2))
              for g in (0, 3)]
    ran = 0
    for cfg in shapes:
        print('batched shape', cfg, flush=True)
        b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda, config=cfg)
        if not b.plan.info()['fused_ok']:
            continue  # this shape cannot hold the 2048-body universe in one workgroup
        b.step(3)
        rg, vg = b.state()
        assert O.rel_err(vg - v, vv - v) < 2e-5, cfg
        assert O.rel_err(rg - r, rr - r) < 2e-5, cfg
        ran += 1
    assert ran >= 12


def test_batched_matches_plain_loop(cuda):
    r, v, m, off = _batch_inputs([40, 700, 64, 1500, 9, 256], seed=3)
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda)
    p = b.plain()
    b.step(10)
    p.step(10)
    rb, vb = b.state()
    rp, vp = p.state()
    assert O.rel_err(vb - v, vp - v) < 1e-5
    assert O.rel_err(rb - r, rp - r) < 1e-5


def test_graph_replay_equals_eager(cuda):
    r, v, m, off = _batch_inputs(SIZES, seed=4)
    b1 = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda)
    b2 = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda)
    g = b2.capture(4)
    for rep in range(3):
        b1.step(4)
        g.replay()
        torch.cuda.synchronize()
        r1, v1 = b1.state()
        r2, v2 = b2.state()
        assert np.array_equal(r1, r2) and np.array_equal(v1, v2), rep


def test_trajectory_record(cuda):
    r, v, m, off = _batch_inputs([10, 100, 65], seed=5)
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda)
    c = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda)
    traj = torch.zeros((4, b.n, 4), dtype=torch.float32, device=cuda)
    b.step(4, traj=traj)
    for s in range(4):
        assert torch.equal(traj[s], c.posm), s
        c.step(1)
    assert torch.equal(b.posm, c.posm)


def test_single_body_universes_move_straight(cuda):
    r, v, m, off = _batch_inputs([1] * 17, seed=6)
    b = batched.Batch(r, v, m, off, G=1.0, eps=0.0, dt=0.25, device=cuda)
    b.step(4)
    rg, vg = b.state()
    assert np.array_equal(vg, v)
    rr = r.copy()
    for _ in range(4):
        rr = v * 0.25 + rr
    assert np.array_equal(rg, rr)


def test_plan_errors(cuda):
    # a universe too large for one workgroup: no fused schedule, the per-step one works
    p = batched.BatchedPlan(np.array([0, 6000]), 8, 1, 8, 0)
    assert p.info()['fused_ok'] == 0 and p.info()['n_items'] == (6000 + 63) // 64
    e = batched.Engine.wrap(p, 'fused')
    z = torch.zeros((6000, 4), dtype=torch.float32, device=cuda)
    r = torch.zeros((3, 6000), dtype=torch.float64, device=cuda)
    with pytest.raises(_native.GravError, match='does not fit'):
        e.forward(z, r, r.clone(), 1.0, 1e-3, 1e-3, 1)
    with pytest.raises(_native.GravInvalidArgument, match='unsupported'):
        batched.BatchedPlan(np.array([0, 10]), 8, 4, 4, 0)
    with pytest.raises(_native.GravInvalidArgument, match='empty'):
        batched.BatchedPlan(np.array([0, 10, 10]), 8, 1, 8, 0)
    p = batched.BatchedPlan(np.array([0, 10, 300, 301]), 4, 1, 2, 2)
    info = p.info()
    assert info['n_bodies'] == 301 and info['n_universes'] == 3 and info['groups'] == 2
    assert info['total_cost'] == 10 * 10 + 290 * 290 + 1
    assert info['fused_ok'] == 1 and info['n_items'] == 1 + 5 + 1
    p.close()


@pytest.mark.parametrize('eps', [1e-3, 0.0])
def test_perstep_vs_oracle_and_fused(cuda, eps):
    r, v, m, off = _batch_inputs(SIZES, seed=11)
    for kt, jw in ((1, 1), (2, 4), (4, 8), (1, 8)):
        print('perstep kt=%d jw=%d' % (kt, jw), flush=True)
        b = batched.Batch(r, v, m, off, G=1.0, eps=eps, dt=1e-3, device=cuda,
                          config=('perstep', 8, kt, 8 // kt, 0, jw))
        b.step(5)
        _check_vs_oracle(b, r, v, m, off, 5)
    f = batched.Batch(r, v, m, off, G=1.0, eps=eps, dt=1e-3, device=cuda)
    f.step(5)
    rb, vb = b.state()
    rf, vf = f.state()
    assert O.rel_err(vb - v, vf - v) < 1e-5


def test_perstep_large_universe_and_graph(cuda):
    r, v, m, off = _batch_inputs([7000, 3, 100, 2049], seed=12)
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda,
                      config=('perstep', 8, 2, 4, 0, 4))
    assert b.plan.info()['fused_ok'] == 0
    b.step(2)
    _check_vs_oracle(b, r, v, m, off, 2)
    c = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda,
                      config=('perstep', 8, 2, 4, 0, 4))
    g = c.capture(2)
    g.replay()
    torch.cuda.synchronize()
    rb, vb = b.state()
    rc, vc = c.state()
    assert np.array_equal(rb, rc) and np.array_equal(vb, vc)


def test_autotuner_picks_a_valid_shape(cuda):
    r, v, m, off = _batch_inputs([100, 2000, 30, 500] * 8, seed=7)
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda)
    r0, v0 = b.state()
    cands = [('fused', 8, 1, 8, 16, 4), ('fused', 4, 2, 4, 32, 4), ('perstep', 8, 2, 4, 0, 4),
             ('perstep', 8, 1, 8, 0, 8)]
    best, table = tune.tune_batched(b, steps=2, candidates=cands)
    assert best in [c for c, _ in table] and len(table) >= 2
    r1, v1 = b.state()
    assert np.array_equal(r0, r1) and np.array_equal(v0, v1)   # state restored
    best1, table1 = tune.tune_stage1(b.posm, candidates=[(1, 8, 256), (4, 4, 256)])
    assert best1 in [(1, 8, 256), (4, 4, 256)]


def test_engine_rejects_tensors_off_its_device(cuda):
    off = np.array([0, 10, 30], dtype=np.int64)
    eng = batched.Engine(off, ('perstep', 4, 1, 4, 0, 2))
    assert eng.plan.device == torch.device('cuda', torch.cuda.current_device())
    posm = torch.zeros((30, 4), dtype=torch.float32)
    r64 = torch.zeros((3, 30), dtype=torch.float64)
    with pytest.raises(ValueError):
        eng.forward(posm, r64, r64.clone(), 1.0, 1e-3, 1e-3, 1)


# sizes below, at and across the symmetric blocks of 256 (4, 1) ... 2048 (8, 4) bodies
SYM_SIZES = [1, 2, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025, 2047, 2048, 2049, 3000, 70]


def _sym_acc(eng, posm, G, eps):
    acc = torch.full((3, eng.n), float('nan'), dtype=torch.float32, device=posm.device)
    _native.call('grav_batched_stage1_sym_f32', eng.plan.handle, posm.data_ptr(),
                 acc.data_ptr(), G, eps, eng._ws.data_ptr(), eng._ws.numel(),
                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return acc


@pytest.mark.parametrize('eps', [1e-3, 0.0])
def test_sym_stage1_vs_oracle_and_owner(cuda, eps):
    r, v, m, off = _batch_inputs(SYM_SIZES, seed=21)
    n = int(off[-1])
    posm = torch.tensor(np.concatenate([r, m[:, None]], axis=1), dtype=torch.float32,
                        device=cuda)
    pn = posm.cpu().numpy().astype(np.float64)
    ref = np.concatenate([O.accel_f64(pn[a:b, :3], pn[a:b, 3], 0.9, eps)
                          for a, b in zip(off[:-1], off[1:])])
    own = torch.empty((3, n), dtype=torch.float32, device=cuda)
    peng = batched.Engine(off, ('perstep', 8, 1, 8, 0, 4))
    _native.call('grav_batched_stage1_f32', peng.plan.handle, posm.data_ptr(), own.data_ptr(),
                 0.9, eps, 4, torch.cuda.current_stream().cuda_stream)
    own = own.cpu().numpy().T.astype(np.float64)
    for sym in batched.SYM_SHAPES:
        print('batched sym', sym, 'eps', eps, flush=True)
        eng = batched.Engine(off, ('perstep-sym', 8, 1, 8, 0, 4, sym))
        a = _sym_acc(eng, posm, 0.9, eps).cpu().numpy().T.astype(np.float64)
        assert np.all(np.isfinite(a)), sym
        for u, (lo, hi) in enumerate(zip(off[:-1], off[1:])):
            if hi - lo == 1:
                assert np.all(a[lo:hi] == 0.0), (sym, u)
                continue
            assert O.rel_err(a[lo:hi], ref[lo:hi]) < 1e-5, (sym, u, hi - lo)
            assert O.rel_err(a[lo:hi], own[lo:hi]) < 1e-5, (sym, u, hi - lo)
        # fixed reduction order: a second call is bitwise the same
        b = _sym_acc(eng, posm, 0.9, eps).cpu().numpy().T.astype(np.float64)
        assert np.array_equal(a, b), sym


def test_sym_perstep_vs_oracle_and_graph(cuda):
    r, v, m, off = _batch_inputs([3000, 5, 700, 1300, 64, 2049], seed=22)
    cfg = ('perstep-sym', 8, 2, 4, 0, 4, (4, 1))
    b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda, config=cfg)
    assert b.config == cfg
    b.step(4)
    _check_vs_oracle(b, r, v, m, off, 4)
    c = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda, config=cfg)
    g = c.capture(2)
    for k in range(2):
        g.replay()
        torch.cuda.synchronize()
        print('graph replay %d' % k, flush=True)
    rb, vb = b.state()
    rc, vc = c.state()
    assert np.array_equal(rb, rc) and np.array_equal(vb, vc)


def test_sym_perstep_adjoint_equals_perstep(cuda):
    # same VJP launches as 'perstep': gradients over the same stored trajectory are equal
    r, v, m, off = _batch_inputs([900, 40, 1600], seed=23)
    steps = 3
    out = {}
    for cfg in (('perstep', 8, 1, 8, 0, 4), ('perstep-sym', 8, 1, 8, 0, 4, (8, 1))):
        b = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda, config=cfg)
        traj = torch.empty((steps, b.n, 4), dtype=torch.float32, device=cuda)
        b.step(steps, traj=traj)
        rho = torch.ones((3, b.n), dtype=torch.float64, device=cuda)
        nu = torch.zeros_like(rho)
        gam = torch.zeros((b.n,), dtype=torch.float64, device=cuda)
        b.engine.adjoint(traj, rho, nu, gam, b.G, b.eps, b.dt, steps)
        torch.cuda.synchronize()
        out[cfg[0]] = (traj.cpu().numpy(), rho.cpu().numpy(), gam.cpu().numpy())
    tp, rp, gp = out['perstep']
    ts, rs, gs = out['perstep-sym']
    assert np.array_equal(tp[0], ts[0])
    assert O.rel_err(ts.astype(np.float64), tp.astype(np.float64)) < 1e-6
    assert O.rel_err(rs, rp) < 1e-5 and O.rel_err(gs, gp) < 1e-5


@pytest.mark.parametrize('cfg', [batched.DEFAULT_CONFIG, ('perstep', 8, 2, 4, 0, 4),
                                 ('perstep-sym', 8, 1, 8, 0, 4, (4, 2))])
def test_shares_step_like_the_whole_batch(cuda, cfg):
    # a batch divided among GPUs by whole universes: each share, stepped alone with the same
    # schedule, gives its universes' state bit for bit as the whole batch does
    print('shares %s' % (cfg,), flush=True)
    sizes = SIZES + [int(x) for x in S.batch_sizes(40, 32, 1500, seed=21)]
    r, v, m, off = _batch_inputs(sizes, seed=22)
    whole = batched.Batch(r, v, m, off, G=1.0, eps=1e-3, dt=1e-3, device=cuda, config=cfg)
    whole.step(4)
    rw, vw = whole.state()
    for share in batched.shard_universes(off, 3):
        rs, vs, ms, os_ = batched.take_universes(r, v, m, off, share)
        b = batched.Batch(rs, vs, ms, os_, G=1.0, eps=1e-3, dt=1e-3, device=cuda, config=cfg)
        b.step(4)
        rb, vb = b.state()
        want_r, want_v, _, _ = batched.take_universes(rw, vw, m, off, share)
        assert np.array_equal(rb, want_r) and np.array_equal(vb, want_v), share
