"""Pins the CPU oracle (oracle/) before anything is checked against it.

Every check is chosen so that a plausible mistake in the oracle (a dropped term, a flipped
sign, a wrong exponent, a transposed operand, the self term kept) fails at least one of them:
brute force in the reference's own pair order, closed forms, the reference's solar-system
known answer, invariants, and the reference's compiled C stage 1.
"""

import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref as R

G_SI = 6.6740831e-11  # base-class default, src/gravitation/kernel/_base_.py:73


@pytest.mark.parametrize('dim', [1, 2, 3])
def test_brute_force_py1_order(rng, dim):
    # The reference's own loop (py1.py:57-69) in pure Python vs the vectorised oracle, in
    # every dimension py1 allows (it zips over the components).
    for n in (2, 3, 5, 8):
        pos = rng.normal(size=(n, dim))
        m = rng.uniform(0.1, 2.0, n)
        np.testing.assert_allclose(O.accel_f64(pos, m, 0.7), O.accel_py1(pos, m, 0.7),
                                   rtol=1e-13, atol=0)


def test_two_body_closed_form():
    d, m1, m2, G = 3.0, 2.0, 5.0, 1.5
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, d]])
    a = O.accel_f64(pos, np.array([m1, m2]), G)
    np.testing.assert_allclose(a[0], [0, 0, G * m2 / d ** 2], rtol=1e-15)
    np.testing.assert_allclose(a[1], [0, 0, -G * m1 / d ** 2], rtol=1e-15)


def test_regular_polygon_closed_form():
    # n equal masses on a circle of radius R: inward radial |a| = G m/(4R^2) sum 1/sin(pi k/n).
    n, R, m, G = 7, 2.5, 0.3, 1.0
    th = 2 * np.pi * np.arange(n) / n
    pos = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(n)], axis=1)
    a = O.accel_f64(pos, np.full(n, m), G)
    mag = G * m / (4 * R * R) * sum(1.0 / math.sin(math.pi * k / n) for k in range(1, n))
    np.testing.assert_allclose(a, -mag * pos / R, rtol=1e-13, atol=1e-15)


def test_solarsystem_known_answer():
    # src/gravitation/lib/simulation.py:85-97 with the base G: |a_earth| = G M_sun / AU^2.
    au = 149597870700.0
    pos = np.array([[0.0, 0.0, 0.0], [0.0, -au, 0.0]])
    m = np.array([1.98892e30, 5.97237e24])
    a = O.accel_f64(pos, m, G_SI)
    assert abs(a[1, 1] - 5.931412e-03) < 5e-9
    assert a[1, 0] == 0.0 and a[1, 2] == 0.0
    np.testing.assert_allclose(a[0, 1], -G_SI * 5.97237e24 / au ** 2, rtol=1e-15)


def test_softened_pair_closed_form():
    d, eps, m2, G = 0.01, 0.003, 2.0, 1.0
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    a = O.accel_f64(pos, np.array([1.0, m2]), G, eps=eps)
    np.testing.assert_allclose(a[0, 0], G * m2 * d / (d * d + eps * eps) ** 1.5, rtol=1e-14)
    # eps > 0: the self term is 0 * finite and a single body feels nothing
    assert np.all(O.accel_f64(pos[:1], np.array([1.0]), G, eps=eps) == 0.0)


@pytest.mark.parametrize('eps', [0.0, 0.05])
def test_invariants(rng, eps):
    n = 64
    pos = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n)
    a = O.accel_f64(pos, m, 1.0, eps)
    scale = np.abs(m[:, None] * a).sum()
    assert np.abs((m[:, None] * a).sum(axis=0)).max() < 1e-13 * scale   # momentum
    np.testing.assert_allclose(O.accel_f64(pos + 3.0, m, 1.0, eps), a, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(O.accel_f64(pos, 2.5 * m, 1.0, eps), 2.5 * a, rtol=1e-13)
    perm = rng.permutation(n)
    np.testing.assert_allclose(O.accel_f64(pos[perm], m[perm], 1.0, eps), a[perm], rtol=1e-12)
    if eps == 0.0:
        lam = 4.0
        np.testing.assert_allclose(O.accel_f64(lam * pos, m, 1.0), a / lam ** 2, rtol=1e-12)
    # continuity in eps at 0
    np.testing.assert_allclose(O.accel_f64(pos, m, 1.0, 1e-9), O.accel_f64(pos, m, 1.0, 0.0),
                               rtol=1e-8)


def test_rows_subset(rng):
    pos = rng.normal(size=(50, 3))
    m = rng.uniform(0.5, 1.5, 50)
    rows = np.array([0, 7, 49, 3])
    np.testing.assert_array_equal(O.accel_f64(pos, m, 1.0, 0.01, rows=rows),
                                  O.accel_f64(pos, m, 1.0, 0.01)[rows])


@pytest.mark.skipif(not R.available('oracle'), reason='liboracle.so not built (make -C oracle)')
def test_c_oracle_matches_numpy(rng):
    pos = rng.normal(size=(300, 3))
    m = rng.uniform(0.5, 1.5, 300)
    for eps in (0.0, 1e-3):
        np.testing.assert_allclose(R.oracle_accel_c(pos, m, 1.0, eps),
                                   O.accel_f64(pos, m, 1.0, eps), rtol=1e-11, atol=1e-13)
        rows = np.array([5, 0, 299])
        np.testing.assert_allclose(R.oracle_accel_c(pos, m, 1.0, eps, rows),
                                   O.accel_f64(pos, m, 1.0, eps, rows), rtol=1e-11, atol=1e-13)


def _c1_inputs():
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 1e3, (1024, 3)).astype(np.float32).astype(np.float64)
    m = rng.uniform(1e-3, 1e-1, 1024).astype(np.float32).astype(np.float64)
    return pos, m, 6.6743e-11


@pytest.mark.skipif(not R.available('lib1'), reason='oracle/_ref not built (reference absent)')
def test_reference_lib1_agrees():
    # The reference's own compiled C stage 1 (FP32, src/gravitation/kernel/_lib1_/lib.c).
    pos, m, G = _c1_inputs()
    assert O.rel_err(R.RefLib1(pos, m, G).accel(), O.accel_f64(pos, m, G)) < 2e-6


@pytest.mark.skipif(not R.available('lib4'), reason='oracle/_ref not built (reference absent)')
def test_reference_lib4_agrees():
    # SSE + OpenMP reference (_lib4_/lib.c); its _mm_rsqrt_ps term limits agreement to ~1e-4.
    pos, m, G = _c1_inputs()
    assert O.rel_err(R.RefLib4(pos, m, G, 4).accel(), O.accel_f64(pos, m, G)) < 1e-3


def test_np1_np4_restatements():
    pos, m, G = _c1_inputs()
    ref = O.accel_f64(pos, m, G)
    assert O.rel_err(O.np1_accel(pos, m, G), ref) < 2e-6
    assert O.rel_err(O.np4_accel(pos, m, G, 3), ref) < 2e-6
    # np4 over every row meets the same bar; np4 over a row subset equals np1 over the same
    # rows up to the order of the per-worker sums
    full = O.np4_rows(pos, m, G, range(pos.shape[0]), 3)
    assert O.rel_err(full, ref) < 2e-6
    sub = np.arange(0, pos.shape[0] - 1, 7)
    np.testing.assert_allclose(O.np4_rows(pos, m, G, sub, 4), O.np1_rows(pos, m, G, sub),
                               rtol=1e-5, atol=1e-6 * np.abs(ref).max())


def test_equal_pair_batches():
    for n, w in [(10, 4), (1024, 4), (3, 8), (2, 1), (1000, 7)]:
        b = O.equal_pair_batches(n, w)
        assert b[0][0] == 0 and b[-1][1] == n - 1
        assert all(b[k][1] == b[k + 1][0] for k in range(len(b) - 1))
        assert len(b) <= w


def test_stage2_closed_forms():
    # a = 0: straight line; constant a: r_k = r0 + k T v0 + a T^2 k(k+1)/2 (dyadic, exact).
    r0 = np.array([[1.0, -2.0, 0.5]])
    v0 = np.array([[0.25, 0.5, -1.0]])
    a = np.array([[0.125, -0.25, 2.0]])
    T = 0.5
    r, v = r0.copy(), v0.copy()
    for k in range(1, 9):
        r, v = O.stage2(r, v, a, T)
        np.testing.assert_array_equal(v, v0 + k * T * a)
        np.testing.assert_array_equal(r, r0 + k * T * v0 + a * T * T * k * (k + 1) / 2)
    r, v = O.stage2(r0, v0, np.zeros_like(a), T)
    np.testing.assert_array_equal(r, r0 + T * v0)
    # update order: v first, then r with the new v (src/gravitation/kernel/_base_.py:60-61)
    r1, v1 = O.stage2(r0, v0, a, T)
    np.testing.assert_array_equal(r1, r0 + T * (v0 + T * a))
    assert not np.array_equal(r1, r0 + T * v0)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.mark.parametrize('name', ['c1_lib1.npz', 'solarsystem_lib1.npz'])
def test_oracle_against_reference_golden_vectors(name):
    # Committed outputs of the reference's own C stage 1 (tests/golden/make_golden.py).
    d = np.load(os.path.join(GOLDEN, name))
    pos = d['pos'].astype(np.float64)
    m = d['m'].astype(np.float64)
    ref = O.accel_f64(pos, m, float(d['G']))
    assert O.rel_err(d['acc'], ref) < 2e-6
    if name.startswith('solarsystem'):
        assert abs(float(d['acc'][1, 1]) - 5.931412e-03) < 1e-8
