"""Adjoint kernels and differentiable trajectories on the GPU.

* grav_stage1_vjp_f32 against the oracle's closed-form VJP (oracle.vjp_accel_f64);
* trajectory gradients (single universe and batched) against the oracle adjoint;
* the 4096-body Plummer case (FP32 forward + FP64 stage 2, T = 200, dt = 1e-4, checkpoints
  every 20 steps) against FP64 central differences (h = 1e-6) on 64 random coordinates; the
  FD trajectories run on the FP64 GPU path, which is pinned to the oracle at 1e-12 by
  test_stage1_gpu.py;
* checkpoint-and-recompute equal to the recorded trajectory, bit for bit.
"""

import numpy as np
import pytest
import torch

from gravitation_amd import _native, adjoint, batched, stages
from gravitation_amd.lib import scenarios as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _posm(pos, m, dev):
    p = torch.empty((pos.shape[0], 4), dtype=torch.float32)
    p[:, :3] = torch.from_numpy(pos)
    p[:, 3] = torch.from_numpy(m)
    return p.to(dev)


def _unscaled(mu4):
    """The FP32 cotangent a scaled mu4 (x, y, z, E) stands for: (x, y, z) * 2^E, in FP64;
    checks that one universe has one exponent and that |x, y, z| < 1."""
    h = mu4.cpu()
    e = h[:, 3].contiguous().view(torch.int32).numpy()
    xyz = h[:, :3].numpy().astype(np.float64)
    assert np.all(e == e[0]) and np.abs(xyz).max() < 1.0
    return np.ldexp(xyz, e[:, None])


@pytest.mark.parametrize('n', [1, 2, 65, 500, 3000])
@pytest.mark.parametrize('eps', [0.0, 1e-2])
def test_vjp_kernel_vs_oracle(cuda, n, eps):
    print('vjp n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    mu = rng.normal(size=(n, 3))
    nu = torch.from_numpy(mu.T.copy()).to(cuda)
    rho = torch.zeros((3, n), dtype=torch.float64, device=cuda)
    gamma = torch.zeros((n,), dtype=torch.float64, device=cuda)
    mu4 = torch.empty((n, 4), dtype=torch.float32, device=cuda)
    adjoint.adjoint_mu(rho, nu, mu4, 1.0)           # mu4 = (float) mu 2^-E, nu unchanged
    adjoint.stage1_vjp_f32(posm, mu4, rho, gamma, 0.7, eps, 1.0)
    torch.cuda.synchronize()
    p = posm.cpu().numpy().astype(np.float64)
    mu32 = _unscaled(mu4)
    jr, jm = O.vjp_accel_f64(p[:, :3], p[:, 3], mu32, 0.7, eps)
    gr = rho.cpu().numpy().T
    gm = gamma.cpu().numpy()
    assert np.array_equal(nu.cpu().numpy().T, mu)
    if n == 1:
        assert np.all(gr == 0) and np.all(gm == 0)
        return
    assert O.rel_err(gr, jr) < 2e-5
    assert O.rel_err(gm, jm) < 2e-5


def _oracle_grad(r0, v0, m, G, eps, dt, steps, target):
    return O.trajectory_with_grad(r0, v0, m, G, eps, dt, steps, target, dtype=np.float32)


def _gpu_grad(fn, r0, v0, m, target, dev):
    tr = torch.tensor(r0, device=dev, requires_grad=True)
    tv = torch.tensor(v0, device=dev, requires_grad=True)
    tm = torch.tensor(m, device=dev, requires_grad=True)
    rT, vT = fn(tr, tv, tm)
    loss = ((rT - torch.tensor(target, device=dev)) ** 2).sum()
    loss.backward()
    return (loss.item(), tr.grad.cpu().numpy(), tv.grad.cpu().numpy(), tm.grad.cpu().numpy())


@pytest.mark.parametrize('eps', [0.0, 1e-2])
def test_single_trajectory_gradient_vs_oracle(cuda, eps):
    n, steps, G, dt = 300, 16, 1.0, 1e-3
    rng = np.random.default_rng(5)
    r0, v0, m = S.random_universe(n, rng)
    target = r0 + rng.normal(0, 1e-2, r0.shape)
    L, _, _, gr, gv, gm = _oracle_grad(r0, v0, m, G, eps, dt, steps, target)
    for interval in (None, 5):
        print('single grad eps=%g interval=%s' % (eps, interval), flush=True)
        res = _gpu_grad(lambda a, b, c: adjoint.trajectory(a, b, c, steps, G, eps, dt,
                                                           interval=interval),
                        r0, v0, m, target, cuda)
        assert abs(res[0] - L) < 1e-6 * L
        assert O.rel_err(res[1], gr) < 1e-4
        assert O.rel_err(res[2], gv) < 1e-4
        assert O.rel_err(res[3], gm) < 1e-4


def test_checkpointing_equals_recording_bitwise(cuda):
    n, steps = 200, 23
    rng = np.random.default_rng(6)
    r0, v0, m = S.random_universe(n, rng)
    target = r0 + rng.normal(0, 1e-2, r0.shape)
    f = lambda iv: (lambda a, b, c: adjoint.trajectory(a, b, c, steps, 1.0, 1e-3, 1e-3,
                                                       interval=iv))
    rec = _gpu_grad(lambda a, b, c: adjoint.trajectory(a, b, c, steps, 1.0, 1e-3, 1e-3),
                    r0, v0, m, target, cuda)
    for iv in (1, 4, 23):
        res = _gpu_grad(f(iv), r0, v0, m, target, cuda)
        for x, y in zip(res[1:], rec[1:]):
            assert np.array_equal(x, y), iv


def _fp64_loss(r0, v0, m, G, eps, dt, steps, target, dev):
    n = r0.shape[0]
    posm = torch.empty((n, 4), dtype=torch.float64, device=dev)
    posm[:, :3] = torch.from_numpy(r0).to(dev)
    posm[:, 3] = torch.from_numpy(m).to(dev)
    vel = torch.from_numpy(np.ascontiguousarray(v0.T)).to(dev)
    acc = torch.empty((3, n), dtype=torch.float64, device=dev)
    for _ in range(steps):
        stages.stage1_f64(posm, acc, G, eps)
        stages.stage2_f64(posm, vel, acc, dt)
    return float(((posm[:, :3].cpu().numpy() - target) ** 2).sum())


def test_c4_plummer_gradient_vs_fp64_central_differences(cuda):
    n, steps, G, eps, dt = 4096, 200, 1.0, 1e-3, 1e-4
    r0, v0, m = S.plummer(n, seed=4)
    rng = np.random.default_rng(44)
    # target: the state perturbed by 1e-3 Gaussian noise, integrated (FP64 GPU path)
    rp = r0 + rng.normal(0, 1e-3, r0.shape)
    vp = v0 + rng.normal(0, 1e-3, v0.shape)
    posm = torch.empty((n, 4), dtype=torch.float64, device=cuda)
    posm[:, :3] = torch.from_numpy(rp).to(cuda)
    posm[:, 3] = torch.from_numpy(m).to(cuda)
    vel = torch.from_numpy(np.ascontiguousarray(vp.T)).to(cuda)
    acc = torch.empty((3, n), dtype=torch.float64, device=cuda)
    for _ in range(steps):
        stages.stage1_f64(posm, acc, G, eps)
        stages.stage2_f64(posm, vel, acc, dt)
    target = posm[:, :3].cpu().numpy()
    L, gr, gv, gm = _gpu_grad(lambda a, b, c: adjoint.trajectory(a, b, c, steps, G, eps, dt,
                                                                 interval=20),
                              r0, v0, m, target, cuda)
    h = 1e-6
    fd, ad = [], []
    for q in range(64):
        which = q % 3
        k = int(rng.integers(n))
        c = int(rng.integers(3))
        args = [r0.copy(), v0.copy(), m.copy()]
        plus = [a.copy() for a in args]
        minus = [a.copy() for a in args]
        if which == 2:
            plus[2][k] += h
            minus[2][k] -= h
            ad.append(gm[k])
        else:
            plus[which][k, c] += h
            minus[which][k, c] -= h
            ad.append((gr, gv)[which][k, c])
        fd.append((_fp64_loss(*plus, G, eps, dt, steps, target, cuda)
                   - _fp64_loss(*minus, G, eps, dt, steps, target, cuda)) / (2 * h))
    fd, ad = np.array(fd), np.array(ad)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print('C4 gradient vs FP64 central differences (64 coords): rel err %.3e' % err, flush=True)
    assert err < 1e-4


def _batch(sizes, seed):
    rng = np.random.default_rng(seed)
    parts = [S.random_universe(int(s), rng) for s in sizes]
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    return (np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]),
            np.concatenate([p[2] for p in parts]), off)


@pytest.mark.parametrize('eps', [0.0, 1e-3])
@pytest.mark.parametrize('mode', ['fused', 'perstep'])
def test_batched_gradient_vs_oracle_and_single(cuda, eps, mode):
    sizes = [1, 3, 64, 65, 200, 700, 33]
    r0, v0, m, off = _batch(sizes, 7)
    rng = np.random.default_rng(8)
    target = r0 + rng.normal(0, 1e-2, r0.shape)
    steps, G, dt = 12, 1.0, 1e-3
    engine = batched.Engine(off, (mode, 4, 1, 4, 3, 4))
    for interval in (None, 5):
        print('batched grad %s eps=%g interval=%s' % (mode, eps, interval), flush=True)
        res = _gpu_grad(lambda a, b, c: adjoint.batched_trajectory(engine, a, b, c, steps, G,
                                                                   eps, dt, interval=interval),
                        r0, v0, m, target, cuda)
        for u in range(len(sizes)):
            a, b = off[u], off[u + 1]
            _, _, _, gr, gv, gm = _oracle_grad(r0[a:b], v0[a:b], m[a:b], G, eps, dt, steps,
                                               target[a:b])
            if np.linalg.norm(gr) > 0:
                assert O.rel_err(res[1][a:b], gr) < 1e-4, u
            assert O.rel_err(res[2][a:b], gv) < 1e-4, u
            if np.linalg.norm(gm) > 0:
                assert O.rel_err(res[3][a:b], gm) < 1e-4, u
            one = _gpu_grad(lambda x, y, z: adjoint.trajectory(x, y, z, steps, G, eps, dt),
                            r0[a:b], v0[a:b], m[a:b], target[a:b], cuda)
            assert O.rel_err(res[2][a:b], one[2]) < 2e-5, u


@pytest.mark.parametrize('mode', ['fused', 'perstep'])
def test_batched_checkpointing_equals_recording_bitwise(cuda, mode):
    r0, v0, m, off = _batch([100, 40, 300], 9)
    target = r0 + 0.01
    plan = batched.Engine(off, (mode, 4, 1, 4, 2, 2))
    rec = _gpu_grad(lambda a, b, c: adjoint.batched_trajectory(plan, a, b, c, 17, 1.0, 1e-3,
                                                               1e-3), r0, v0, m, target, cuda)
    for iv in (1, 6):
        res = _gpu_grad(lambda a, b, c: adjoint.batched_trajectory(plan, a, b, c, 17, 1.0, 1e-3,
                                                                   1e-3, interval=iv),
                        r0, v0, m, target, cuda)
        for x, y in zip(res[1:], rec[1:]):
            assert np.array_equal(x, y), iv


def test_orbit_fit_reduces_loss_and_matches_oracle_gradient(cuda):
    # Batched orbit fitting (gradient descent on v0) at reduced size: the loss never rises,
    # falls twentyfold, and the first gradient equals the oracle adjoint's.
    sizes = [64, 100, 300, 77]
    r0, v0, m, off = _batch(sizes, 21)
    steps, dt = 20, 1e-3
    rng = np.random.default_rng(22)
    v_true = v0 + rng.normal(0, 0.05, v0.shape)
    target, _, _ = O.simulate(r0, v_true, m, 1.0, 1e-3, dt, steps, dtype=np.float32)
    eng = batched.Engine(off, ('perstep', 8, 1, 8, 0, 4))
    tr = torch.tensor(r0, device=cuda)
    tm = torch.tensor(m, device=cuda)
    stats = {}
    v_fit, losses = adjoint.fit_initial_velocities(
        eng, tr, torch.tensor(v0, device=cuda), tm, torch.tensor(target, device=cuda), steps,
        iters=6, G=1.0, eps=1e-3, dt=dt, stats=stats)
    assert all(b <= a for a, b in zip(losses, losses[1:])), losses
    assert losses[-1] < 0.05 * losses[0]
    assert np.all(stats['best'] <= stats['first']) and stats['best'].shape == (len(sizes),)
    assert abs(stats['best'].sum() - losses[-1]) <= 1e-12 * losses[-1]
    # gradient of the first iteration against the oracle adjoint, per universe
    res = _gpu_grad(lambda a, b, c: adjoint.batched_trajectory(eng, a, b, c, steps, 1.0, 1e-3, dt),
                    r0, v0, m, target, cuda)
    for u in range(len(sizes)):
        a, e = off[u], off[u + 1]
        _, _, _, _, gv, _ = _oracle_grad(r0[a:e], v0[a:e], m[a:e], 1.0, 1e-3, dt, steps,
                                         target[a:e])
        assert O.rel_err(res[2][a:e], gv) < 1e-4, u


def _vjp_once(posm, mu, G, eps):
    n = posm.shape[0]
    dev = posm.device
    nu = torch.from_numpy(np.ascontiguousarray(mu.T)).to(dev)
    rho = torch.zeros((3, n), dtype=torch.float64, device=dev)
    gamma = torch.zeros((n,), dtype=torch.float64, device=dev)
    mu4 = torch.empty((n, 4), dtype=torch.float32, device=dev)
    adjoint.adjoint_mu(rho, nu, mu4, 1.0)
    adjoint.stage1_vjp_f32(posm, mu4, rho, gamma, G, eps, 1.0)
    torch.cuda.synchronize()
    return rho.cpu().numpy().T, gamma.cpu().numpy(), mu4


def test_vjp_scale_is_exact_and_extends_the_range(cuda):
    # The cotangent is scaled by a power of two per universe: results are homogeneous
    # bit for bit, and a cotangent of 1e35 (whose FP32 pair terms s (m_k mu_j - m_j mu_k)
    # would overflow unscaled) still matches the oracle.
    n, G, eps = 400, 0.7, 1e-3
    rng = np.random.default_rng(31)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    mu = rng.normal(size=(n, 3))
    r1, g1, _ = _vjp_once(posm, mu, G, eps)
    r2, g2, _ = _vjp_once(posm, np.ldexp(mu, 60), G, eps)
    r3, g3, _ = _vjp_once(posm, np.ldexp(mu, -60), G, eps)
    assert np.array_equal(r2, np.ldexp(r1, 60)) and np.array_equal(g2, np.ldexp(g1, 60))
    assert np.array_equal(r3, np.ldexp(r1, -60)) and np.array_equal(g3, np.ldexp(g1, -60))
    rb, gb, mu4 = _vjp_once(posm, mu * 1e35, G, eps)
    assert np.all(np.isfinite(rb)) and np.all(np.isfinite(gb))
    p = posm.cpu().numpy().astype(np.float64)
    jr, jm = O.vjp_accel_f64(p[:, :3], p[:, 3], _unscaled(mu4), G, eps)
    assert O.rel_err(rb, jr) < 2e-5 and O.rel_err(gb, jm) < 2e-5


def test_batched_adjoint_mu_scales_each_universe(cuda):
    # One exponent per universe, taken from that universe's own largest |mu|; the batched VJP
    # then equals the single-universe VJP of each universe.
    sizes = [50, 70, 30, 1]
    r0, _, m, off = _batch(sizes, 32)
    n = r0.shape[0]
    rng = np.random.default_rng(33)
    mu = rng.normal(size=(n, 3))
    for u, f in enumerate((1e-30, 1.0, 3e30, 0.0)):
        mu[off[u]:off[u + 1]] *= f
    posm = _posm(r0, m, cuda)
    nu = torch.from_numpy(np.ascontiguousarray(mu.T)).to(cuda)
    rho = torch.zeros((3, n), dtype=torch.float64, device=cuda)
    gamma = torch.zeros((n,), dtype=torch.float64, device=cuda)
    mu4 = torch.empty((n, 4), dtype=torch.float32, device=cuda)
    plan = batched.BatchedPlan(off, 4, 1, 4, 0)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    _native.call('grav_batched_adjoint_mu', plan.handle, rho.data_ptr(), nu.data_ptr(),
                 mu4.data_ptr(), 1.0, stream)
    _native.call('grav_batched_stage1_vjp_f32', plan.handle, posm.data_ptr(), mu4.data_ptr(),
                 rho.data_ptr(), gamma.data_ptr(), 0.7, 1e-3, 1.0, 2, stream)
    torch.cuda.synchronize()
    e = mu4[:, 3].cpu().contiguous().view(torch.int32).numpy()
    gr = rho.cpu().numpy().T
    gm = gamma.cpu().numpy()
    for u in range(len(sizes)):
        a, b = off[u], off[u + 1]
        big = np.abs(mu[a:b]).max()
        want = np.frexp(big)[1] if big > 0 else 0
        assert np.all(e[a:b] == want), (u, e[a:b], want)
        if b - a < 2 or big == 0:
            assert np.all(gr[a:b] == 0) and np.all(gm[a:b] == 0)
            continue
        r1, g1, _ = _vjp_once(posm[a:b].contiguous(), mu[a:b], 0.7, 1e-3)
        assert O.rel_err(gr[a:b], r1) < 1e-6, u
        assert O.rel_err(gm[a:b], g1) < 1e-6, u


def test_fused_adjoint_scales_like_perstep(cuda):
    # The fused backward (exponents in LDS) and the per-step backward agree on cotangents of
    # 1e35, where unscaled FP32 pair terms would overflow, and the fused one is homogeneous
    # bit for bit under a power-of-two change of the cotangent.
    sizes = [64, 200, 31, 129]
    r0, v0, m, off = _batch(sizes, 34)
    n = r0.shape[0]
    steps, G, eps, dt = 6, 1.0, 1e-3, 1e-3
    rng = np.random.default_rng(35)
    rho0 = rng.normal(size=(3, n)) * 1e35
    nu0 = rng.normal(size=(3, n)) * 1e35
    out = {}
    for mode, shift in (('fused', 0), ('fused', 50), ('perstep', 0)):
        eng = batched.Engine(off, (mode, 4, 1, 4, 0, 2))
        posm = _posm(r0, m, cuda)
        r64 = torch.from_numpy(np.ascontiguousarray(r0.T)).to(cuda)
        v64 = torch.from_numpy(np.ascontiguousarray(v0.T)).to(cuda)
        traj = torch.empty((steps, n, 4), dtype=torch.float32, device=cuda)
        eng.forward(posm, r64, v64, G, eps, dt, steps, traj)
        rho = torch.from_numpy(np.ldexp(rho0, shift)).to(cuda)
        nu = torch.from_numpy(np.ldexp(nu0, shift)).to(cuda)
        gamma = torch.zeros((n,), dtype=torch.float64, device=cuda)
        eng.adjoint(traj, rho, nu, gamma, G, eps, dt, steps)
        torch.cuda.synchronize()
        out[(mode, shift)] = (rho.cpu().numpy(), nu.cpu().numpy(), gamma.cpu().numpy())
    f0, f50, p0 = out[('fused', 0)], out[('fused', 50)], out[('perstep', 0)]
    for x, y, z in zip(f0, f50, p0):
        assert np.all(np.isfinite(x))
        assert np.array_equal(np.ldexp(x, 50), y)
        assert O.rel_err(x, z) < 1e-4


SPLIT_CFGS = (None, (1, 1, 1), (1, 4, 8), (2, 2, 3), (1, 8, 64), (2, 8, 5))


@pytest.mark.parametrize('n', [1, 2, 65, 1000, 4096])
@pytest.mark.parametrize('eps', [0.0, 1e-2])
def test_vjp_split_vs_oracle_and_unsplit(cuda, n, eps):
    print('vjp split n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 5)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    mu = rng.normal(size=(n, 3))
    rho0 = torch.from_numpy(rng.normal(size=(3, n))).to(cuda)
    gam0 = torch.from_numpy(rng.normal(size=(n,))).to(cuda)
    nu = torch.from_numpy(mu.T.copy()).to(cuda)
    mu4 = torch.empty((n, 4), dtype=torch.float32, device=cuda)
    adjoint.adjoint_mu(torch.zeros_like(rho0), nu, mu4, 1.0)
    ru, gu = rho0.clone(), gam0.clone()
    adjoint.stage1_vjp_f32(posm, mu4, ru, gu, 0.7, eps, 1.0)
    p = posm.cpu().numpy().astype(np.float64)
    jr, jm = O.vjp_accel_f64(p[:, :3], p[:, 3], _unscaled(mu4), 0.7, eps)
    du = (ru - rho0).cpu().numpy().T, (gu - gam0).cpu().numpy()
    for cfg in SPLIT_CFGS:
        rs, gs = rho0.clone(), gam0.clone()
        adjoint.stage1_vjp_split_f32(posm, mu4, rs, gs, 0.7, eps, 1.0, cfg)
        dr, dg = (rs - rho0).cpu().numpy().T, (gs - gam0).cpu().numpy()
        if n == 1:
            assert np.all(dr == 0) and np.all(dg == 0), cfg
            continue
        assert O.rel_err(dr, jr) < 2e-5 and O.rel_err(dg, jm) < 2e-5, cfg
        assert O.rel_err(dr, du[0]) < 1e-6 and O.rel_err(dg, du[1]) < 1e-6, cfg
        r2, g2 = rho0.clone(), gam0.clone()
        adjoint.stage1_vjp_split_f32(posm, mu4, r2, g2, 0.7, eps, 1.0, cfg)
        assert torch.equal(r2, rs) and torch.equal(g2, gs), cfg


def test_single_trajectory_split_forms_match_owner(cuda):
    # 3000 bodies: 'auto' runs the j-split stage 1 and VJP, 'owner' the one-launch kernels;
    # the gradients agree to the FP32 pair terms' rounding
    n, steps, G, eps, dt = 3000, 12, 1.0, 1e-3, 1e-4
    r0, v0, m = S.plummer(n, seed=31)
    target = r0 + np.random.default_rng(32).normal(0, 1e-3, r0.shape)
    out = {}
    for form in ('auto', 'owner'):
        out[form] = _gpu_grad(lambda a, b, c: adjoint.trajectory(a, b, c, steps, G, eps, dt,
                                                                 interval=4, form=form),
                              r0, v0, m, target, cuda)
    (la, *ga), (lo, *go) = out['auto'], out['owner']
    assert abs(la - lo) <= 1e-6 * abs(lo)
    for a, o in zip(ga, go):
        assert O.rel_err(a, o) < 1e-5


@pytest.mark.parametrize('steps', [50, 47])
def test_single_trajectory_graph_replay_is_bitwise_eager(cuda, steps):
    # checkpoint segments after the first replay as HIP graphs (forward, recompute, adjoint),
    # with eager copies between the replays and a partial last segment for 47 steps: every
    # output equals eager stepping bit for bit
    n, G, eps, dt = 3000, 1.0, 1e-3, 1e-4
    r0, v0, m = S.plummer(n, seed=41)
    target = r0 + np.random.default_rng(42).normal(0, 1e-3, r0.shape)
    out = {}
    for graph in (False, True):
        print('graph=%s steps=%d' % (graph, steps), flush=True)
        out[graph] = _gpu_grad(lambda a, b, c: adjoint.trajectory(a, b, c, steps, G, eps, dt,
                                                                  interval=10, graph=graph),
                               r0, v0, m, target, cuda)
    assert out[True][0] == out[False][0]
    for a, b in zip(out[True][1:], out[False][1:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize('n', [2048, 4097, 100000])
def test_adjoint_mu_tiles_is_bitwise_one_workgroup(cuda, n):
    # grav_adjoint_mu_ws (ceil(n/1024) workgroups and a workspace for the exponents) writes
    # the same nu and mu4 as grav_adjoint_mu (one workgroup), bit for bit
    rng = np.random.default_rng(n)
    rho = torch.from_numpy(rng.normal(size=(3, n))).to(cuda)
    nu0 = torch.from_numpy(rng.normal(size=(3, n))).to(cuda)
    nu0[1, n // 3] = 3.5e40   # the universe's exponent comes from one body
    out = []
    for ws in (None, torch.empty((4 * ((n + 1023) // 1024),), dtype=torch.uint8, device=cuda)):
        nu = nu0.clone()
        mu4 = torch.full((n, 4), float('nan'), dtype=torch.float32, device=cuda)
        adjoint.adjoint_mu(rho, nu, mu4, 0.37, ws)
        out.append((nu.cpu().numpy(), mu4.cpu().numpy().view(np.uint32)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def _gpu_grad_rv(fn, r0, v0, m, target, target_v, dev):
    """Loss sum |v_T - target_v|^2 (+ sum |r_T - target|^2 unless target is None) through
    fn; returns (L, dL/dr0, dL/dv0, dL/dm)."""
    tr = torch.tensor(r0, device=dev, requires_grad=True)
    tv = torch.tensor(v0, device=dev, requires_grad=True)
    tm = torch.tensor(m, device=dev, requires_grad=True)
    rT, vT = fn(tr, tv, tm)
    loss = ((vT - torch.tensor(target_v, device=dev)) ** 2).sum()
    if target is not None:
        loss = loss + ((rT - torch.tensor(target, device=dev)) ** 2).sum()
    loss.backward()
    return (loss.item(), tr.grad.cpu().numpy(), tv.grad.cpu().numpy(), tm.grad.cpu().numpy())


@pytest.mark.parametrize('with_r', [False, True])
@pytest.mark.parametrize('eps', [0.0, 1e-2])
def test_velocity_loss_single_vs_oracle(cuda, eps, with_r):
    # A loss on v_T seeds the backward with nu = dL/dv_T: recorded, checkpointed eager and
    # checkpointed graph-replayed trajectories all against the oracle adjoint
    n, steps, G, dt = 300, 16, 1.0, 1e-3
    rng = np.random.default_rng(51)
    r0, v0, m = S.random_universe(n, rng)
    target = r0 + rng.normal(0, 1e-2, r0.shape) if with_r else None
    target_v = v0 + rng.normal(0, 5e-2, v0.shape)
    L, _, _, gr, gv, gm = O.trajectory_with_grad(r0, v0, m, G, eps, dt, steps, target,
                                                 dtype=np.float32, target_v=target_v)
    for interval, graph in ((None, True), (5, False), (5, True)):
        print('v_T loss eps=%g with_r=%s interval=%s graph=%s' % (eps, with_r, interval, graph),
              flush=True)
        res = _gpu_grad_rv(lambda a, b, c: adjoint.trajectory(a, b, c, steps, G, eps, dt,
                                                              interval=interval, graph=graph),
                           r0, v0, m, target, target_v, cuda)
        assert abs(res[0] - L) < 1e-6 * L
        assert O.rel_err(res[1], gr) < 1e-4
        assert O.rel_err(res[2], gv) < 1e-4
        assert O.rel_err(res[3], gm) < 1e-4


def test_velocity_loss_split_forms_vs_oracle(cuda):
    # 3000 bodies: the j-split stage 1, fused split step and j-split VJP, with a loss on both
    # r_T and v_T, against the oracle adjoint (C oracle: same contract, pinned on the CPU)
    from oracle import ref as R
    n, steps, G, eps, dt = 3000, 12, 1.0, 1e-3, 1e-4
    r0, v0, m = S.plummer(n, seed=52)
    rng = np.random.default_rng(53)
    v0 = v0 + rng.normal(0, 0.1, v0.shape)
    target = r0 + rng.normal(0, 1e-3, r0.shape)
    target_v = v0 + rng.normal(0, 1e-2, v0.shape)
    kw = dict(accel=R.oracle_accel_c, vjp=R.oracle_vjp_c) if R.available('oracle') else {}
    L, _, _, gr, gv, gm = O.trajectory_with_grad(r0, v0, m, G, eps, dt, steps, target,
                                                 dtype=np.float32, target_v=target_v, **kw)
    res = _gpu_grad_rv(lambda a, b, c: adjoint.trajectory(a, b, c, steps, G, eps, dt,
                                                          interval=4),
                       r0, v0, m, target, target_v, cuda)
    assert abs(res[0] - L) < 1e-6 * L
    for got, want in zip(res[1:], (gr, gv, gm)):
        assert O.rel_err(got, want) < 1e-4


@pytest.mark.parametrize('mode', ['fused', 'perstep'])
def test_velocity_loss_batched_vs_oracle(cuda, mode):
    sizes = [1, 3, 64, 65, 200, 700, 33]
    r0, v0, m, off = _batch(sizes, 54)
    rng = np.random.default_rng(55)
    target = r0 + rng.normal(0, 1e-2, r0.shape)
    target_v = v0 + rng.normal(0, 5e-2, v0.shape)
    steps, G, eps, dt = 12, 1.0, 1e-3, 1e-3
    engine = batched.Engine(off, (mode, 4, 1, 4, 3, 4))
    for interval in (None, 5):
        print('batched v_T loss %s interval=%s' % (mode, interval), flush=True)
        res = _gpu_grad_rv(lambda a, b, c: adjoint.batched_trajectory(engine, a, b, c, steps, G,
                                                                      eps, dt, interval=interval),
                           r0, v0, m, target, target_v, cuda)
        for u in range(len(sizes)):
            a, b = off[u], off[u + 1]
            _, _, _, gr, gv, gm = O.trajectory_with_grad(
                r0[a:b], v0[a:b], m[a:b], G, eps, dt, steps, target[a:b], dtype=np.float32,
                target_v=target_v[a:b])
            assert O.rel_err(res[1][a:b], gr) < 1e-4, u
            assert O.rel_err(res[2][a:b], gv) < 1e-4, u
            if np.linalg.norm(gm) > 0:
                assert O.rel_err(res[3][a:b], gm) < 1e-4, u


@pytest.mark.parametrize('n', [1024, 4097, 20000])
def test_chained_vjp_step_is_bitwise_separate(cuda, n):
    # grav_stage1_vjp_f32_split_chain (the VJP, then its reduce forming the next step's
    # cotangent with the previous step's exponent) == grav_adjoint_mu + grav_stage1_vjp_f32_split
    # in rho, nu and gamma, bit for bit, over six backward steps (the last without a next
    # cotangent).  With rho of 1e30, mu = nu + dt rho grows every step (by 1e27 plus what the
    # VJP adds to rho, which close pairs make large), so the lagged scale leaves the largest
    # |mu4| at or above 1 where the exact one keeps it below 1: the lag is exercised, and
    # mu4 * 2^E equals the exactly scaled copy
    rng = np.random.default_rng(n + 61)
    posm = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    rho0 = rng.normal(size=(3, n)) * 1e30
    nu0 = rng.normal(size=(3, n))
    dt, G, eps = 1e-3, 0.7, 1e-3
    for cfg in (None, (1, 8, 16), (2, 4, 5)):
        print('chained vjp n=%d cfg=%s' % (n, cfg), flush=True)
        out, lagged, amax = [], [], []
        for chain in (False, True):
            rho = torch.from_numpy(rho0.copy()).to(cuda)
            nu = torch.from_numpy(nu0.copy()).to(cuda)
            gamma = torch.zeros((n,), dtype=torch.float64, device=cuda)
            mu4 = torch.empty((n, 4), dtype=torch.float32, device=cuda)
            eslot = torch.zeros((2,), dtype=torch.int32, device=cuda)
            if chain:
                adjoint.adjoint_mu(rho, nu, mu4, dt)
            for t in range(6):
                if chain:
                    adjoint.stage1_vjp_split_chain(posm, mu4, rho, nu, gamma, G, eps, dt, t < 5,
                                                   eslot, cfg)
                    if t < 5:
                        m = mu4.cpu().numpy()
                        amax.append(np.abs(m[:, :3]).max())
                        lagged.append(np.ldexp(m[:, :3].astype(np.float64),
                                               m[:, 3].view(np.int32)[:, None]))
                else:
                    adjoint.adjoint_mu(rho, nu, mu4, dt)
                    if t > 0:
                        m = mu4.cpu().numpy()
                        exact = np.ldexp(m[:, :3].astype(np.float64),
                                         m[:, 3].view(np.int32)[:, None])
                        lagged.append(exact)
                    adjoint.stage1_vjp_split_f32(posm, mu4, rho, gamma, G, eps, dt, cfg)
            out.append([x.cpu().numpy() for x in (rho, nu, gamma)])
        for a, b in zip(*out):
            assert np.array_equal(a, b), (n, cfg)
        exact, chained = lagged[:5], lagged[5:]
        for a, b in zip(exact, chained):
            assert np.array_equal(a, b), (n, cfg)
        assert max(amax) >= 1.0 and np.isfinite(amax).all(), amax


@pytest.mark.parametrize('steps', [50, 47])
def test_single_trajectory_chain_and_overlap_are_bitwise_plain(cuda, steps):
    # the single-universe adjoint with chained backward steps and the recomputation of one
    # segment overlapping the adjoint of the next (graphs, eager, and a partial segment of odd
    # length) equals the plain sequence of separate kernels on one stream bit for bit
    n, G, eps, dt = 3000, 0.7, 1e-3, 1e-4   # G not a float: the FP32 and FP64 scalings differ
    r0, v0, m = S.plummer(n, seed=71)
    rng = np.random.default_rng(72)
    target = r0 + rng.normal(0, 1e-3, r0.shape)
    target_v = v0 + rng.normal(0, 1e-3, v0.shape)
    out = {}
    saved = adjoint.CHAIN, adjoint.OVERLAP
    try:
        for chain in (False, True):
            for overlap in (False, True):
                adjoint.CHAIN, adjoint.OVERLAP = chain, overlap
                for graph in (True, False):
                    print('chain=%s overlap=%s graph=%s steps=%d' % (chain, overlap, graph,
                                                                     steps), flush=True)
                    out[(chain, overlap, graph)] = _gpu_grad_rv(
                        lambda a, b, c: adjoint.trajectory(a, b, c, steps, G, eps, dt,
                                                           interval=10, graph=graph),
                        r0, v0, m, target, target_v, cuda)
    finally:
        adjoint.CHAIN, adjoint.OVERLAP = saved
    base = out[(False, False, False)]
    for key, res in out.items():
        assert res[0] == base[0], key
        for a, b in zip(res[1:], base[1:]):
            assert np.array_equal(a, b), key


@pytest.mark.parametrize('n', [1000, 4097, 20000])
def test_backward_pair_is_bitwise_separate(cuda, n):
    # grav_backward_pair_split (a chained backward step and an independent forward step in
    # the same two launches) == grav_stage1_vjp_f32_split_chain + grav_step_split_f64m_io, bit
    # for bit, over four steps of each, for the library's shapes and two others
    rng = np.random.default_rng(n + 83)
    pa = _posm(rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n), cuda)
    r, v, m = S.plummer(n, seed=n + 84)
    rho0 = rng.normal(size=(3, n)) * 1e20
    nu0 = rng.normal(size=(3, n))
    G, eps, dt = 0.9, 1e-3, 1e-3
    for cfg in (None, (1, 8, 2, 3, 8), (2, 5, 1, 16, 4)):
        print('backward pair n=%d cfg=%s' % (n, cfg), flush=True)
        vcfg = (cfg[0], cfg[4], cfg[1]) if cfg else None
        scfg = (cfg[2], cfg[4], cfg[3]) if cfg else None
        out = []
        for paired in (False, True):
            rho = torch.from_numpy(rho0.copy()).to(cuda)
            nu = torch.from_numpy(nu0.copy()).to(cuda)
            gamma = torch.zeros((n,), dtype=torch.float64, device=cuda)
            mu4 = torch.empty((n, 4), dtype=torch.float32, device=cuda)
            eslot = torch.zeros((2,), dtype=torch.int32, device=cuda)
            traj = torch.full((5, n, 4), float('nan'), dtype=torch.float32, device=cuda)
            traj[0] = _posm(r, m, cuda)
            r64 = torch.from_numpy(np.ascontiguousarray(r.T)).to(cuda)
            v64 = torch.from_numpy(np.ascontiguousarray(v.T)).to(cuda)
            ws = torch.empty((adjoint.backward_pair_workspace_bytes(n, cfg),), dtype=torch.uint8,
                             device=cuda)
            adjoint.adjoint_mu(rho, nu, mu4, dt)
            for t in range(4):
                if paired:
                    adjoint.backward_pair_split(pa, mu4, rho, nu, gamma, eslot, traj[t],
                                                traj[t + 1], r64, v64, G, eps, dt, ws, cfg)
                else:
                    adjoint.stage1_vjp_split_chain(pa, mu4, rho, nu, gamma, G, eps, dt, True,
                                                   eslot, vcfg)
                    stages.step_split_f64m(traj[t], r64, v64, G, eps, dt, scfg, out=traj[t + 1])
            out.append([x.cpu().numpy() for x in (rho, nu, gamma, mu4, traj, r64, v64)])
        for a, b in zip(*out):
            assert np.array_equal(a, b, equal_nan=True), (n, cfg)


def test_recorded_last_segment_only_reused_by_its_own_backward(cuda):
    # The forward pass leaves its last segment's positions in a record buffer that its own
    # backward pass reuses.  Two trajectories of one configuration, both run forward before
    # either backward, must not pick up each other's record: their gradients equal those of
    # each trajectory run alone, bit for bit
    n, G, eps, dt, steps = 3000, 0.7, 1e-3, 1e-4, 47
    runs = []
    for seed in (81, 82):
        r0, v0, m = S.plummer(n, seed=seed)
        rng = np.random.default_rng(seed)
        runs.append((r0, v0, m, r0 + rng.normal(0, 1e-3, r0.shape)))

    def leaves(r0, v0, m):
        return tuple(torch.tensor(x, device=cuda, requires_grad=True) for x in (r0, v0, m))

    def loss(ls, target):
        rT, _ = adjoint.trajectory(*ls, steps, G, eps, dt, interval=10)
        return ((rT - torch.tensor(target, device=cuda)) ** 2).sum()

    alone = []
    for r0, v0, m, target in runs:
        ls = leaves(r0, v0, m)
        loss(ls, target).backward()
        alone.append([x.grad.cpu().numpy() for x in ls])
    both = [leaves(r0, v0, m) for r0, v0, m, _ in runs]
    total = loss(both[0], runs[0][3]) + loss(both[1], runs[1][3])
    total.backward()
    for ls, ref in zip(both, alone):
        for x, g in zip(ls, ref):
            assert np.array_equal(x.grad.cpu().numpy(), g)


@pytest.mark.parametrize('mode', ['fused', 'perstep'])
def test_batched_adjoint_shares_equal_the_whole(cuda, mode):
    # a batch divided among GPUs by whole universes: each share's forward + adjoint, alone,
    # gives its universes' loss terms and gradients bit for bit as the whole batch does
    sizes = [1, 3, 64, 65, 200, 700, 33, 129, 450, 5]
    r0, v0, m, off = _batch(sizes, 31)
    rng = np.random.default_rng(32)
    target = r0 + rng.normal(0, 1e-2, r0.shape)
    steps, G, eps, dt = 14, 1.0, 1e-3, 1e-3
    cfg = (mode, 4, 1, 4, 0, 4)
    for interval in (None, 5):
        print('batched adjoint shares %s interval=%s' % (mode, interval), flush=True)
        whole = _gpu_grad(lambda a, b, c: adjoint.batched_trajectory(
            batched.Engine(off, cfg), a, b, c, steps, G, eps, dt, interval=interval),
            r0, v0, m, target, cuda)
        for share in batched.shard_universes(off, 3):
            rs, vs, ms, os_ = batched.take_universes(r0, v0, m, off, share)
            ts = batched.take_universes(target, target, m, off, share)[0]
            res = _gpu_grad(lambda a, b, c: adjoint.batched_trajectory(
                batched.Engine(os_, cfg), a, b, c, steps, G, eps, dt, interval=interval),
                rs, vs, ms, ts, cuda)
            for k in (1, 2, 3):
                want = batched.take_universes(whole[k], whole[k], whole[3], off, share)
                got = res[k]
                exp = want[2] if k == 3 else want[0]
                assert np.array_equal(got, exp), (mode, interval, share, k)


@pytest.mark.parametrize('n, eps', [(1000, 0.0), (20000, 1e-3), (70001, 0.0)])
def test_vjp_shard_rows_equal_whole_universe(cuda, n, eps):
    # the VJP of one GPU's share of the rows, uneven shares starting on multiples of 256: rho
    # and gamma rows equal the whole-universe VJP's of the same shape bit for bit,
    # owner-computes (the library's shape for n) and j-split; a share starting inside an
    # i-tile equals them to rounding (its self-masked pairs fall in other compensation blocks)
    print('vjp shard n=%d eps=%g' % (n, eps), flush=True)
    rng = np.random.default_rng(n + 61)
    r, _, m = S.random_universe(n, rng)
    posm = torch.empty((n, 4), dtype=torch.float32)
    posm[:, :3] = torch.from_numpy(r)
    posm[:, 3] = torch.from_numpy(m)
    posm = posm.to(cuda)
    mu4 = torch.empty((n, 4), dtype=torch.float32, device=cuda)
    nu = torch.from_numpy(rng.normal(size=(3, n)) * 1e3).to(cuda)
    adjoint.adjoint_mu(torch.zeros_like(nu), nu, mu4, 1.0)
    rho0 = torch.from_numpy(rng.normal(size=(3, n))).to(cuda)
    gam0 = torch.from_numpy(rng.normal(size=n)).to(cuda)
    cuts = sorted(set([0, n] + [256 * int(x) for x in rng.integers(1, n // 256, 4)]))
    owner = (4, 4) if n >= 1 << 16 else ((2, 4) if n >= 1 << 13 else (1, 8))
    for cfg in (owner + (1,), (1, 8, 8), (2, 8, 16), (2, 4, 7)):
        rho, gam = rho0.clone(), gam0.clone()
        if cfg[2] == 1:
            adjoint.stage1_vjp_f32(posm, mu4, rho, gam, 0.9, eps, 1e-3)
        else:
            adjoint.stage1_vjp_split_f32(posm, mu4, rho, gam, 0.9, eps, 1e-3, cfg)
        rw, gw = rho.cpu(), gam.cpu()
        assert torch.isfinite(rw).all() and not torch.equal(rw, rho0.cpu())
        for i0, i1 in zip(cuts[:-1], cuts[1:]):
            rs, gs = rho0[:, i0:i1].clone(), gam0[i0:i1].clone()
            adjoint.stage1_vjp_shard_f32(posm, mu4, rs, gs, i0, 0.9, eps, 1e-3, cfg)
            assert torch.equal(rs.cpu(), rw[:, i0:i1]), (cfg, i0, i1)
            assert torch.equal(gs.cpu(), gw[i0:i1]), (cfg, i0, i1)
        i0 = 77
        rs, gs = rho0[:, i0:].clone(), gam0[i0:].clone()
        adjoint.stage1_vjp_shard_f32(posm, mu4, rs, gs, i0, 0.9, eps, 1e-3, cfg)
        for got, ref, start in ((rs.cpu(), rw[:, i0:], rho0[:, i0:].cpu()),
                                (gs.cpu(), gw[i0:], gam0[i0:].cpu())):
            assert O.rel_err((got - start).numpy(), (ref - start).numpy()) < 1e-6, cfg
    rs, gs = rho0[:, :n // 2].clone(), gam0[:n // 2].clone()
    adjoint.stage1_vjp_shard_f32(posm, mu4, rs, gs, 0, 0.9, eps, 1e-3)
    assert torch.isfinite(rs).all() and not torch.equal(rs, rho0[:, :n // 2])
    stages.release_workspaces()
