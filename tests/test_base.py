"""The plugin contract (gravitation_amd/kernel/_base_.py) against the reference's behaviour
(src/gravitation/kernel/_base_.py:39-177): lifecycle errors, G and body scaling, stage order."""

import numpy as np
import pytest

from gravitation_amd.kernel import _base_ as B
from oracle import oracle as O


class _oracle_universe(B.universe_base):
    """A CPU test kernel: FP64 oracle stage 1 into the lists, base stage 2."""

    def step_stage1(self):
        pos = np.array([pm._r for pm in self], dtype=np.float64)
        m = np.array([pm._m for pm in self])
        a = O.accel_f64(pos, m, self._G)
        for pm, ai in zip(self, a):
            pm._a[:] = list(ai)


def test_lifecycle_errors():
    u = _oracle_universe()
    with pytest.raises(SyntaxError, match='simulation was not started'):
        u.step()
    with pytest.raises(SyntaxError, match='simulation was not started'):
        u.stop()
    u.add_object(name='a', r=[0.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], m=1.0)
    u.add_object(name='b', r=[1.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], m=1.0)
    u.start()
    with pytest.raises(SyntaxError, match='simulation is running'):
        u.start()
    with pytest.raises(SyntaxError, match='simulation was started'):
        u.add_object(name='c', r=[2.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], m=1.0)
    u.step()
    u.stop()
    with pytest.raises(SyntaxError, match='simulation was stopped before'):
        u.stop()
    with pytest.raises(SyntaxError, match='simulation was stopped'):
        u.step()
    with pytest.raises(SyntaxError, match='simulation was stopped'):
        u.start()
    with pytest.raises(SyntaxError, match='simulation was stopped'):
        u.add_object(name='c', r=[2.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], m=1.0)


def test_missing_stage1():
    u = B.universe_base()
    u.add_object(name='a', r=[0.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], m=1.0)
    u.start()
    with pytest.raises(NotImplementedError):
        u.step()


def test_scaling_G_and_objects():
    u = B.universe_base(G=2.0, scale_m=1e-3, scale_r=10.0, T=5.0, t=1.0, dtype='float64',
                        threads=3, eps=0.1)
    assert u._G == pytest.approx(2.0 * 1000.0 / 1e-3)
    assert u._meta == {'eps': 0.1} and u._T == 5.0 and u._t == 1.0
    assert u._dtype == 'float64' and u._threads == 3
    r, v = [1.0, 2.0, 3.0], [0.5, 0.0, -1.0]
    u.add_object(name='x', r=r, v=v, m=4.0)
    assert r == [10.0, 20.0, 30.0] and v == [5.0, 0.0, -10.0]  # scaled in place
    pm = next(iter(u))
    assert pm._r is r and pm._m == pytest.approx(4e-3) and pm._a == [0.0, 0.0, 0.0]
    u2 = B.universe_base(G=2.0, scale_m=1e-3, scale_r=10.0, scale_off=True)
    assert u2._G == 2.0
    r2 = [1.0, 2.0, 3.0]
    u2.add_object(name='y', r=r2, v=[0.0, 0.0, 0.0], m=1.0, scale_off=True)
    assert r2 == [1.0, 2.0, 3.0]


def test_add_objects_bulk_and_alias():
    u = B.universe_base(scale_r=2.0, scale_m=3.0)
    r = np.ones((4, 3))
    v = np.zeros((4, 3))
    m = np.ones(4)
    u.add_objects(r, v, m)
    u.add_mass(name='z', r=[1.0, 1.0, 1.0], v=[0.0, 0.0, 0.0], m=1.0)
    assert len(u) == 5
    assert np.all(r == 2.0) and np.all(m == 3.0)
    assert np.shares_memory(next(iter(u))._r, r)


def test_base_stage2_and_time():
    # One step of the default stage 2 (src/gravitation/kernel/_base_.py:58-62) on a 2-body
    # system equals the oracle's semi-implicit Euler; stage 3 adds T.
    u = _oracle_universe(T=0.5, G=1.0, t=2.0)
    u.add_object(name='a', r=[0.0, 0.0, 0.0], v=[0.0, 1.0, 0.0], m=1.0)
    u.add_object(name='b', r=[2.0, 0.0, 0.0], v=[0.0, -1.0, 0.0], m=3.0)
    r0 = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    v0 = np.array([[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]])
    u.start()
    u.step()
    a = O.accel_f64(r0, np.array([1.0, 3.0]), 1.0)
    r1, v1 = O.stage2(r0, v0, a, 0.5)
    np.testing.assert_array_equal(np.array([pm._r for pm in u]), r1)
    np.testing.assert_array_equal(np.array([pm._v for pm in u]), v1)
    assert all(pm._a == [0.0, 0.0, 0.0] for pm in u)
    assert u._t == 2.5
    assert 'a | ' in str(u)
