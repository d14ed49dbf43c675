"""Pins the oracle's adjoint (oracle.vjp_accel_f64, oracle.trajectory_with_grad) on the CPU:
FP64 central finite differences, torch autograd in FP64 on the CPU, and a two-body closed
form for dL/dm."""

import numpy as np
import pytest
import torch

from oracle import oracle as O


def _setup(n=6, seed=0, steps=4):
    rng = np.random.default_rng(seed)
    r0 = rng.normal(size=(n, 3))
    v0 = rng.normal(0, 0.1, size=(n, 3))
    m = rng.uniform(0.5, 1.5, n)
    target = r0 + rng.normal(0, 1e-2, size=(n, 3))
    return r0, v0, m, target


@pytest.mark.parametrize('eps', [0.0, 0.05])
def test_vjp_accel_against_finite_differences(eps):
    rng = np.random.default_rng(1)
    n = 7
    pos = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n)
    mu = rng.normal(size=(n, 3))
    jr, jm = O.vjp_accel_f64(pos, m, mu, 1.3, eps)
    h = 1e-6
    for k in range(n):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[k, c] += h
            p2[k, c] -= h
            fd = ((O.accel_f64(p1, m, 1.3, eps) - O.accel_f64(p2, m, 1.3, eps)) * mu).sum()
            assert abs(fd / (2 * h) - jr[k, c]) < 1e-6 * (1 + abs(jr[k, c]))
        m1, m2 = m.copy(), m.copy()
        m1[k] += h
        m2[k] -= h
        fd = ((O.accel_f64(pos, m1, 1.3, eps) - O.accel_f64(pos, m2, 1.3, eps)) * mu).sum()
        assert abs(fd / (2 * h) - jm[k]) < 1e-6 * (1 + abs(jm[k]))


@pytest.mark.parametrize('eps', [0.0, 0.05])
def test_trajectory_gradient_against_finite_differences(eps):
    r0, v0, m, target = _setup()
    G, dt, steps = 1.0, 0.01, 5
    _, _, _, gr, gv, gm = O.trajectory_with_grad(r0, v0, m, G, eps, dt, steps, target)
    h = 1e-6
    rng = np.random.default_rng(2)
    for _ in range(12):
        which = rng.integers(3)
        k = rng.integers(r0.shape[0])
        c = rng.integers(3)
        args = [r0.copy(), v0.copy(), m.copy()]
        plus = [a.copy() for a in args]
        minus = [a.copy() for a in args]
        if which == 2:
            plus[2][k] += h
            minus[2][k] -= h
            ref = gm[k]
        else:
            plus[which][k, c] += h
            minus[which][k, c] -= h
            ref = (gr, gv)[which][k, c]
        fd = (O.trajectory_loss(*plus, G, eps, dt, steps, target)
              - O.trajectory_loss(*minus, G, eps, dt, steps, target)) / (2 * h)
        assert abs(fd - ref) < 1e-7 + 1e-6 * abs(ref), (which, k, c, fd, ref)


def _torch_loss(r0, v0, m, G, eps, dt, steps, target):
    r, v = r0, v0
    n = r.shape[0]
    eye = torch.eye(n, dtype=torch.bool)
    for _ in range(steps):
        d = r[None, :, :] - r[:, None, :]                  # r_j - r_i
        q = (d * d).sum(-1) + eps * eps
        q = torch.where(eye, torch.ones_like(q), q)
        s = torch.where(eye, torch.zeros_like(q), q ** -1.5)
        a = G * (m[None, :, None] * s[:, :, None] * d).sum(1)
        v = a * dt + v
        r = v * dt + r
    return ((r - target) ** 2).sum()


@pytest.mark.parametrize('eps', [0.0, 0.05])
def test_trajectory_gradient_against_torch_autograd(eps):
    r0, v0, m, target = _setup(n=9, seed=3)
    G, dt, steps = 0.8, 0.02, 6
    loss, _, _, gr, gv, gm = O.trajectory_with_grad(r0, v0, m, G, eps, dt, steps, target)
    tr = torch.tensor(r0, requires_grad=True)
    tv = torch.tensor(v0, requires_grad=True)
    tm = torch.tensor(m, requires_grad=True)
    tl = _torch_loss(tr, tv, tm, G, eps, dt, steps, torch.tensor(target))
    tl.backward()
    assert abs(tl.item() - loss) < 1e-12 * abs(loss) + 1e-15
    np.testing.assert_allclose(gr, tr.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gv, tv.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gm, tm.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_mass_gradient_two_body_closed_form():
    # one step: r1' = r1 + dt v1 + dt^2 G m2 d s, with d = r2 - r1, s = |d|^-3.
    r0 = np.array([[0.0, 0.0, 0.0], [1.5, 0.5, -0.25]])
    v0 = np.zeros((2, 3))
    m = np.array([2.0, 3.0])
    target = np.array([[0.1, 0.0, 0.0], [1.0, 0.0, 0.0]])
    G, dt = 1.0, 0.1
    _, r1, _, _, _, gm = O.trajectory_with_grad(r0, v0, m, G, 0.0, dt, 1, target)
    d = r0[1] - r0[0]
    s = np.linalg.norm(d) ** -3
    dm2 = 2 * (r1[0] - target[0]) @ (dt * dt * G * d * s)
    dm1 = 2 * (r1[1] - target[1]) @ (dt * dt * G * (-d) * s)
    np.testing.assert_allclose(gm, [dm1, dm2], rtol=1e-13)


def _torch_loss_rv(r0, v0, m, G, eps, dt, steps, target, target_v):
    r, v = r0, v0
    n = r.shape[0]
    eye = torch.eye(n, dtype=torch.bool)
    for _ in range(steps):
        d = r[None, :, :] - r[:, None, :]
        q = (d * d).sum(-1) + eps * eps
        q = torch.where(eye, torch.ones_like(q), q)
        s = torch.where(eye, torch.zeros_like(q), q ** -1.5)
        a = G * (m[None, :, None] * s[:, :, None] * d).sum(1)
        v = a * dt + v
        r = v * dt + r
    loss = ((v - target_v) ** 2).sum()
    if target is not None:
        loss = loss + ((r - target) ** 2).sum()
    return loss


@pytest.mark.parametrize('with_r', [False, True])
@pytest.mark.parametrize('eps', [0.0, 0.05])
def test_velocity_loss_against_torch_autograd(eps, with_r):
    # A loss on v_T seeds nu = 2 (v_T - target_v): checked against torch autograd of the same
    # FP64 step (alone and beside the r_T term).
    r0, v0, m, target = _setup(n=8, seed=7)
    target_v = v0 + np.random.default_rng(8).normal(0, 0.05, v0.shape)
    G, dt, steps = 0.9, 0.02, 6
    tgt = target if with_r else None
    loss, _, _, gr, gv, gm = O.trajectory_with_grad(r0, v0, m, G, eps, dt, steps, tgt,
                                                    target_v=target_v)
    tr = torch.tensor(r0, requires_grad=True)
    tv = torch.tensor(v0, requires_grad=True)
    tm = torch.tensor(m, requires_grad=True)
    tl = _torch_loss_rv(tr, tv, tm, G, eps, dt, steps,
                        None if tgt is None else torch.tensor(tgt), torch.tensor(target_v))
    tl.backward()
    assert abs(tl.item() - loss) < 1e-12 * abs(loss) + 1e-15
    np.testing.assert_allclose(gr, tr.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gv, tv.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gm, tm.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_velocity_loss_against_finite_differences():
    r0, v0, m, target = _setup(n=6, seed=9)
    target_v = v0 + 0.03
    G, eps, dt, steps = 1.0, 0.02, 0.01, 5
    _, _, _, gr, gv, gm = O.trajectory_with_grad(r0, v0, m, G, eps, dt, steps, target,
                                                 target_v=target_v)
    h = 1e-6
    for which, k, c in ((0, 1, 2), (1, 3, 0), (1, 0, 1), (2, 4, 0), (0, 5, 1)):
        plus = [r0.copy(), v0.copy(), m.copy()]
        minus = [r0.copy(), v0.copy(), m.copy()]
        if which == 2:
            plus[2][k] += h
            minus[2][k] -= h
            ref = gm[k]
        else:
            plus[which][k, c] += h
            minus[which][k, c] -= h
            ref = (gr, gv)[which][k, c]
        fd = (O.trajectory_loss(*plus, G, eps, dt, steps, target, target_v)
              - O.trajectory_loss(*minus, G, eps, dt, steps, target, target_v)) / (2 * h)
        assert abs(fd - ref) < 1e-7 + 1e-6 * abs(ref), (which, k, c, fd, ref)


def test_velocity_loss_free_motion_closed_form():
    # G = 0: v_T = v0 and r_T = r0 + T dt v0, so dL/dr0 = 2 (r_T - r*) and
    # dL/dv0 = 2 T dt (r_T - r*) + 2 (v_T - v*); dL/dm = 0.
    r0, v0, m, target = _setup(n=5, seed=11)
    target_v = np.full_like(v0, 0.25)
    dt, steps = 0.05, 7
    _, r_t, v_t, gr, gv, gm = O.trajectory_with_grad(r0, v0, m, 0.0, 0.0, dt, steps, target,
                                                     target_v=target_v)
    np.testing.assert_allclose(v_t, v0, rtol=0, atol=0)
    np.testing.assert_allclose(gr, 2 * (r_t - target), rtol=1e-14)
    np.testing.assert_allclose(gv, 2 * steps * dt * (r_t - target) + 2 * (v0 - target_v),
                               rtol=1e-13, atol=1e-15)
    assert np.all(gm == 0.0)


@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_c_oracle_vjp_matches_numpy(eps):
    # The C oracle's VJP (used by the GPU tests for universes of thousands of bodies) against
    # the NumPy one that finite differences and torch autograd pin above; and a trajectory
    # gradient through both.
    from oracle import ref as R
    if not R.available('oracle'):
        pytest.skip('oracle/c/liboracle.so not built')
    rng = np.random.default_rng(12)
    n = 300
    pos = rng.random((n, 3))
    m = rng.uniform(1e-4, 1e-2, n)
    mu = rng.normal(size=(n, 3))
    jr, jm = R.oracle_vjp_c(pos, m, mu, 0.9, eps)
    nr, nm = O.vjp_accel_f64(pos, m, mu, 0.9, eps)
    assert O.rel_err(jr, nr) < 1e-13 and O.rel_err(jm, nm) < 1e-13
    r0, v0, m, target = _setup(n=40, seed=13)
    a = O.trajectory_with_grad(r0, v0, m, 1.0, eps, 0.01, 5, target, target_v=v0)
    b = O.trajectory_with_grad(r0, v0, m, 1.0, eps, 0.01, 5, target, target_v=v0,
                               accel=R.oracle_accel_c, vjp=R.oracle_vjp_c)
    for x, y in zip(a[3:], b[3:]):
        assert O.rel_err(y, x) < 1e-12
