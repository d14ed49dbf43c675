"""Pins ``create_galaxy`` (the reference's default worker workload,
src/gravitation/lib/simulation.py:99-184, built by create_simulation :57-83) on the CPU with
facts its construction fixes, not by restating its code:

* each star's geometry, inverted (azimuth, cylindrical radius, disk thickness, elevation), gives
  back the uniform draws of a ``random.Random(seed)`` replayed in the reference's order: per
  star alpha, then r_abs, then the thickness draw (disk) or the inclination draw (bulge).  Any
  other order, scale or offset of a draw fails;
* counts: 1 + stars_len bodies, the first 4/5 of the stars (integer division) in the disk;
* bands: disk radius in [0.1, 4.6) radius, |z| <= radius/4 * (4.6 radius - r_c) / (4.6 radius);
  bulge radius in [0.1, 0.85) radius, inclination in [-pi/2, pi/2);
* circular orbits about the hole: v^2 |r| = G M, r . v = 0, and v = |v| (sin phi, -cos phi, 0)
  (clockwise seen from +z);
* the galaxy rotation (g_alpha, g_beta) is R = Rz(g_alpha) Rx(g_beta) applied to every position
  and velocity (it draws nothing, so the same seed gives R times the unrotated galaxy), and
  moves the disk's angular momentum from -z to -R z;
* the centre offsets r, v shift every star;
* create_simulation's galaxy: G * scale_r^3 / scale_m leaves G unchanged (1e-30 / 1e-30),
  positions and velocities are the unscaled galaxy times 1e-10, masses times 1e-30.
"""

import math
import random

import numpy as np
import pytest

from gravitation_amd.kernel import _base_ as B
from gravitation_amd.lib import simulation

RADIUS = 1e20
M_HOLE, M_STAR = 4e40, 2e30
OUTER = 4.6 * RADIUS


def _galaxy(stars, seed, g_alpha=0.0, g_beta=0.0, r=(0.0, 0.0, 0.0), v=(0.0, 0.0, 0.0)):
    u = B.universe_base()
    simulation.create_galaxy(u, stars, r=list(r), v=list(v), g_alpha=g_alpha, g_beta=g_beta,
                             m_hole=M_HOLE, m_star=M_STAR, radius=RADIUS,
                             rng=random.Random(seed))
    pos = np.array([pm._r for pm in u], dtype=np.float64)
    vel = np.array([pm._v for pm in u], dtype=np.float64)
    m = np.array([pm._m for pm in u])
    names = [pm._name for pm in u]
    return u, pos, vel, m, names


def _rot(g_alpha, g_beta):
    ca, sa, cb, sb = math.cos(g_alpha), math.sin(g_alpha), math.cos(g_beta), math.sin(g_beta)
    rz = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cb, -sb], [0.0, sb, cb]])
    return rz @ rx


@pytest.mark.parametrize('stars,seed', [(999, 0), (1023, 5)])
def test_draws_replay_in_reference_order(stars, seed):
    _, pos, _, m, names = _galaxy(stars, seed)
    assert len(names) == stars + 1
    assert names[0] == 'back hole' and m[0] == M_HOLE
    assert all(nm == 'star' for nm in names[1:]) and np.all(m[1:] == M_STAR)
    assert np.all(pos[0] == 0.0)
    disk = stars * 4 // 5
    rnd = random.Random(seed).random
    for k in range(stars):
        x, y, z = pos[1 + k]
        phi = math.atan2(y, x) % (2.0 * math.pi)
        u_alpha = rnd()
        # the azimuth is the first draw (u = 1 - tiny wraps to 0 under mod 2 pi)
        d_phi = abs(phi - 2.0 * math.pi * u_alpha)
        assert min(d_phi, 2.0 * math.pi - d_phi) < 1e-12, (k, phi, u_alpha)
        if k < disk:
            r_c = math.hypot(x, y)
            u_r = rnd()
            assert r_c == pytest.approx((4.5 * u_r + 0.1) * RADIUS, rel=1e-13), k
            u_z = rnd()
            z_ref = (0.5 * u_z - 0.25) * RADIUS * (OUTER - r_c) / OUTER
            assert z == pytest.approx(z_ref, rel=1e-12, abs=1e-12 * RADIUS), k
        else:
            rad = math.sqrt(x * x + y * y + z * z)
            u_r = rnd()
            assert rad == pytest.approx((0.75 * u_r + 0.1) * RADIUS, rel=1e-13), k
            u_b = rnd()
            assert math.asin(z / rad) == pytest.approx(math.pi * (u_b - 0.5), abs=1e-12), k


def test_bands_and_circular_orbits():
    u, pos, vel, _, _ = _galaxy(4999, 11)
    stars = 4999
    disk = stars * 4 // 5
    s, w = pos[1:], vel[1:]
    r_c = np.hypot(s[:, 0], s[:, 1])
    rad = np.linalg.norm(s, axis=1)
    assert np.all((r_c[:disk] >= 0.1 * RADIUS) & (r_c[:disk] < 4.6 * RADIUS))
    thick = 0.25 * RADIUS * (OUTER - r_c[:disk]) / OUTER
    assert np.all(np.abs(s[:disk, 2]) <= thick * (1 + 1e-12))
    assert np.all((rad[disk:] >= 0.1 * RADIUS) & (rad[disk:] < 0.85 * RADIUS))
    # both populations fill their bands (not only the disk's radial range)
    assert r_c[:disk].max() > 4.4 * RADIUS and rad[disk:].max() > 0.8 * RADIUS
    np.testing.assert_allclose((w ** 2).sum(1) * rad, u._G * M_HOLE, rtol=1e-12)
    speed = np.linalg.norm(w, axis=1)
    assert np.abs((s * w).sum(1)).max() <= 1e-12 * (rad * speed).max()
    phi = np.arctan2(s[:, 1], s[:, 0])
    np.testing.assert_allclose(w[:, 0], speed * np.sin(phi), atol=1e-12 * speed.max())
    np.testing.assert_allclose(w[:, 1], -speed * np.cos(phi), atol=1e-12 * speed.max())
    assert np.all(w[:, 2] == 0.0)


@pytest.mark.parametrize('g_alpha,g_beta', [(0.7, 0.4), (-2.1, 1.2), (0.0, -0.9), (3.0, 0.0)])
def test_rotation_is_rz_rx(g_alpha, g_beta):
    _, p0, v0, _, _ = _galaxy(600, 3)
    _, p1, v1, _, _ = _galaxy(600, 3, g_alpha, g_beta)
    rot = _rot(g_alpha, g_beta)
    scale_r = np.linalg.norm(p0, axis=1).max()
    scale_v = np.linalg.norm(v0, axis=1).max()
    np.testing.assert_allclose(p1, p0 @ rot.T, rtol=0, atol=1e-13 * scale_r)
    np.testing.assert_allclose(v1, v0 @ rot.T, rtol=0, atol=1e-13 * scale_v)
    # |r|, |v| and the orbit invariant survive, and the disk normal turns with R
    np.testing.assert_allclose(np.linalg.norm(p1, axis=1), np.linalg.norm(p0, axis=1),
                               rtol=1e-13)
    disk = 600 * 4 // 5
    lz = np.cross(p1[1:1 + disk], v1[1:1 + disk]).sum(0)
    normal = -rot @ np.array([0.0, 0.0, 1.0])
    assert np.dot(lz / np.linalg.norm(lz), normal) > 0.999
    if g_beta != 0.0:
        assert abs(normal[2]) < 0.999          # the normal really moved off the z axis


def test_centre_offsets_shift_every_star():
    r, v = (3e19, -2e19, 5e18), (120.0, 0.0, -75.0)
    _, p0, v0, _, _ = _galaxy(300, 9)
    _, p1, v1, _, _ = _galaxy(300, 9, r=r, v=v)
    np.testing.assert_allclose(p1, p0 + np.array(r), rtol=0, atol=1e-15 * 5 * RADIUS)
    np.testing.assert_allclose(v1, v0 + np.array(v), rtol=0, atol=1e-12)


def test_create_simulation_galaxy_scaling():
    n = 512
    u = simulation.create_simulation('galaxy', B.universe_base, {'stars_len': n}, seed=4)
    _, p0, v0, m0, _ = _galaxy(n - 1, 4)
    assert len(u) == n
    assert u._G == pytest.approx(6.6740831e-11, rel=1e-15)   # scale_r^3 / scale_m = 1
    assert u._T == 2.0e12 and u._dtype == 'float32'
    pos = np.array([pm._r for pm in u])
    vel = np.array([pm._v for pm in u])
    m = np.array([pm._m for pm in u])
    np.testing.assert_allclose(pos, p0 * 1e-10, rtol=1e-15)
    np.testing.assert_allclose(vel, v0 * 1e-10, rtol=1e-15)
    np.testing.assert_allclose(m, m0 * 1e-30, rtol=1e-15)
    assert m[0] == pytest.approx(4e10) and m[1] == pytest.approx(2.0)
