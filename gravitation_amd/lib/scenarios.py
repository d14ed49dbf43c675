"""Seeded initial conditions for the benchmark workloads.

The reference's scenario generators (src/gravitation/lib/simulation.py:41-184) draw from the
unseeded ``random`` module; every generator here takes a seed (NumPy ``default_rng``).

Workloads
---------
* ``uniform_cube``: N bodies uniform in [0, side)^3, zero velocities, masses uniform in
  [m_lo, m_hi] -- the 1024-body FP32 parity case (G = 6.6743e-11, dt = 1, 10 steps).
* ``plummer``: Plummer sphere (scale radius a, truncated at r_max) by inverse transform of the
  cumulative mass M(<r)/M = r^3 / (r^2 + a^2)^(3/2), isotropic directions, equal masses 1/N,
  zero velocities -- the single large universe (G = 1, eps = 1e-3, dt = 1e-4).
* ``batch_sizes`` / ``random_universe``: uneven batched universes, N log-uniform in
  [n_lo, n_hi], positions uniform in the unit cube, velocities N(0, sigma^2), masses
  log-uniform in [m_lo, m_hi].
"""

import numpy as np


def uniform_cube(n, seed=0, side=1e3, m_lo=1e-3, m_hi=1e-1):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, side, (n, 3))
    v = np.zeros((n, 3))
    m = rng.uniform(m_lo, m_hi, n)
    return r, v, m


def plummer_radius(u, a=1.0):
    """Inverse of M(<r)/M = r^3/(r^2+a^2)^(3/2): r = a / sqrt(u^(-2/3) - 1)."""
    u = np.asarray(u, dtype=np.float64)
    return a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)


def plummer_mass_fraction(r, a=1.0):
    r = np.asarray(r, dtype=np.float64)
    return r ** 3 / (r * r + a * a) ** 1.5


def plummer(n, seed=0, a=1.0, r_max=10.0):
    """Plummer sphere truncated at r_max: u ~ U(0, M(<r_max)/M) gives the truncation exactly
    (no rejection loop)."""
    rng = np.random.default_rng(seed)
    u_max = plummer_mass_fraction(r_max, a)
    u = rng.uniform(0.0, u_max, n)
    u = np.maximum(u, np.finfo(np.float64).tiny)
    rad = plummer_radius(u, a)
    cos_t = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sin_t = np.sqrt(1.0 - cos_t * cos_t)
    r = np.stack([rad * sin_t * np.cos(phi), rad * sin_t * np.sin(phi), rad * cos_t], axis=1)
    v = np.zeros((n, 3))
    m = np.full(n, 1.0 / n)
    return r, v, m


def batch_sizes(n_universes, n_lo, n_hi, seed=0):
    """N_u = round(exp(U[ln n_lo, ln n_hi])) for each universe."""
    rng = np.random.default_rng(seed)
    ln = rng.uniform(np.log(n_lo), np.log(n_hi), n_universes)
    return np.clip(np.rint(np.exp(ln)).astype(np.int64), n_lo, n_hi)


def random_universe(n, rng, sigma_v=0.1, m_lo=1e-4, m_hi=1e-2):
    r = rng.uniform(0.0, 1.0, (n, 3))
    v = rng.normal(0.0, sigma_v, (n, 3))
    m = np.exp(rng.uniform(np.log(m_lo), np.log(m_hi), n))
    return r, v, m


def batched(n_universes, n_lo, n_hi, seed=0, sigma_v=0.1, m_lo=1e-4, m_hi=1e-2):
    """Concatenated state of a batch of universes: (r (B,3), v (B,3), m (B,), offsets (U+1,))
    with universe u owning bodies offsets[u]:offsets[u+1]."""
    sizes = batch_sizes(n_universes, n_lo, n_hi, seed)
    rng = np.random.default_rng(seed + 1)
    rs, vs, ms = [], [], []
    for n in sizes:
        r, v, m = random_universe(int(n), rng, sigma_v, m_lo, m_hi)
        rs.append(r)
        vs.append(v)
        ms.append(m)
    offsets = np.zeros(n_universes + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    return np.concatenate(rs), np.concatenate(vs), np.concatenate(ms), offsets
