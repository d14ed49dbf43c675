"""Scenarios and state files with the reference's contract (src/gravitation/lib/
simulation.py:41-258), seeded.

* ``create_simulation(scenario, universe_class, scenario_param, threads, seed)`` builds and
  starts a universe of any kernel class.  Scenarios: ``solarsystem`` and ``galaxy`` as in the
  reference (:85-184), and the benchmark workloads ``plummer`` and ``cube``.  ``galaxy``
  draws from ``random.Random(seed)`` in the reference's order of draws, so it is reproducible
  (the reference's module-level ``random.random()`` is not seeded, :116 ff.).
* ``store_simulation`` / ``load_simulation`` keep the reference's layout: one group per
  snapshot named like ``kernel=K;len=N;step=S`` (src/gravitation/cli/worker.py:141-150) with
  datasets ``r`` (N, 3), ``v`` (N, 3), ``m`` (N,), ``name`` and the attributes ``scale_m,
  scale_r, t, T, G, dtype, threads`` (:229-256).  HDF5 is used when h5py is importable;
  otherwise an ``.npz`` mirror stores the same fields as ``<group>/r`` etc. plus
  ``<group>/attrs`` (JSON).
"""

import json
import math
import os
import random

import numpy as np

from . import scenarios

AU = 149597870700.0


def create_solarsystem(universe_obj):
    """The reference's two-body sun + earth scenario (:85-97)."""
    universe_obj.add_object(name='sun', r=[0.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], m=1.98892e30)
    universe_obj.add_object(name='earth', r=[0.0, -AU, 0.0], v=[29777.777, 0.0, 0.0],
                            m=5.97237e24)


def create_galaxy(universe_obj, stars_len, r, v, g_alpha, g_beta, m_hole, m_star, radius,
                  rng=None):
    """A central mass plus ``stars_len`` stars on circular orbits about it: 4/5 in a flat
    disk, 1/5 in a central bulge (the reference's construction, :99-184, drawing from
    ``rng``, a random.Random)."""
    rnd = (rng or random).random
    universe_obj.add_object(name='back hole', r=list(r), v=list(v), m=m_hole)
    outer = (4.5 + 0.1) * radius
    for n in range(stars_len):
        alpha = rnd() * 2.0 * math.pi
        if n < (stars_len * 4 // 5):
            r_abs = (rnd() * 4.5 + 0.1) * radius
            r_s = [r_abs * math.cos(alpha), r_abs * math.sin(alpha),
                   (0.5 * rnd() - 0.25) * radius * (outer - r_abs) / outer]
        else:
            r_abs = (rnd() * 0.75 + 0.1) * radius
            beta = math.pi * (rnd() - 0.5)
            r_s = [r_abs * math.cos(alpha) * math.cos(beta),
                   r_abs * math.sin(alpha) * math.cos(beta), r_abs * math.sin(beta)]
        v_abs = math.sqrt(universe_obj._G * m_hole / math.sqrt(sum([d ** 2 for d in r_s])))
        v_alpha = alpha - math.pi / 2
        v_s = [v_abs * math.cos(v_alpha), v_abs * math.sin(v_alpha), 0.0]
        v_s[1:] = [v_s[1] * math.cos(g_beta), v_s[1] * math.sin(g_beta)]
        v_alpha = math.atan2(v_s[1], v_s[0]) + g_alpha
        v_factor = math.sqrt(v_s[0] ** 2 + v_s[1] ** 2)
        v_s[0:2] = [v_factor * math.cos(v_alpha), v_factor * math.sin(v_alpha)]
        v_s = [d + e for d, e in zip(v_s, v)]
        r_beta = math.atan2(r_s[2], r_s[1]) + g_beta
        r_factor = math.sqrt(r_s[2] ** 2 + r_s[1] ** 2)
        r_s[1:] = [r_factor * math.cos(r_beta), r_factor * math.sin(r_beta)]
        r_alpha = math.atan2(r_s[1], r_s[0]) + g_alpha
        r_factor = math.sqrt(r_s[0] ** 2 + r_s[1] ** 2)
        r_s[0:2] = [r_factor * math.cos(r_alpha), r_factor * math.sin(r_alpha)]
        r_s = [d + e for d, e in zip(r_s, r)]
        universe_obj.add_object(name='star', r=r_s, v=v_s, m=m_star)


def create_simulation(scenario, universe_class, scenario_param=None, threads=1, seed=None):
    """Creates, fills and starts a universe of ``universe_class`` (:41-83)."""
    universe_param = {'threads': threads}
    scenario_param = dict(scenario_param or {})
    if scenario == 'solarsystem':
        universe_param.update(scenario_param)
        u = universe_class(**universe_param)
        create_solarsystem(u)
    elif scenario == 'galaxy':
        stars_len = scenario_param.get('stars_len', 2000)   # also lands in _meta, as there
        universe_param.update({'T': 2.0e12, 'scale_m': 1.0e-30, 'scale_r': 1.0e-10,
                               'dtype': 'float32'})
        universe_param.update(scenario_param)
        u = universe_class(**universe_param)
        create_galaxy(u, stars_len - 1, r=[0.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], g_alpha=0.0,
                      g_beta=0.0, m_hole=4e40, m_star=2e30, radius=1e20,
                      rng=random.Random(seed) if seed is not None else None)
    elif scenario in ('plummer', 'cube'):
        n = int(scenario_param.pop('stars_len', 1024))
        if scenario == 'plummer':
            universe_param.update({'T': 1e-4, 'G': 1.0, 'eps': 1e-3})
            r, v, m = scenarios.plummer(n, seed=seed or 0)
        else:
            universe_param.update({'T': 1.0, 'G': 6.6743e-11})
            r, v, m = scenarios.uniform_cube(n, seed=seed or 0)
        universe_param.update(scenario_param)
        u = universe_class(**universe_param)
        u.add_objects(r, v, m, names=['star'] * n)
    else:
        raise ValueError('Unknown scenario: "%s"' % scenario)
    u.start()
    return u


_ATTRS = ('scale_m', 'scale_r', 't', 'T', 'G', 'dtype', 'threads')


def _state(universe_obj):
    n = len(universe_obj)
    dim = len(universe_obj._mass_list[0]._r)
    dtype = {'float32': '<f4', 'float64': '<f8'}[universe_obj._dtype]
    r = np.empty((n, dim), dtype=dtype)
    v = np.empty((n, dim), dtype=dtype)
    m = np.empty((n,), dtype=dtype)
    names = []
    for k, pm in enumerate(universe_obj):
        r[k] = pm._r[:]
        v[k] = pm._v[:]
        m[k] = pm._m
        names.append(pm._name.encode('utf-8'))
    attrs = {a: getattr(universe_obj, '_' + a) for a in _ATTRS}
    return r, v, m, np.array(names, dtype='S%d' % max(1, max(len(x) for x in names))), attrs


def _h5py():
    try:
        import h5py
        return h5py
    except ImportError:
        return None


def store_simulation(universe_obj, fn, gn):
    """Appends group ``gn`` with the universe's state to ``fn`` (HDF5, or .npz mirror)."""
    r, v, m, names, attrs = _state(universe_obj)
    h5 = _h5py()
    if h5 is not None and not fn.endswith('.npz'):
        with h5.File(fn, 'a') as f:
            g = f.create_group(gn)
            g.create_dataset('r', data=r)
            g.create_dataset('v', data=v)
            g.create_dataset('m', data=m)
            g.create_dataset('name', data=names)
            for a, val in attrs.items():
                g.attrs[a] = val
        return
    data = {}
    if os.path.exists(fn):
        with np.load(fn, allow_pickle=False) as old:
            data = {k: old[k] for k in old.files}
    if gn + '/r' in data:
        raise ValueError('group %r exists in %s' % (gn, fn))
    data.update({gn + '/r': r, gn + '/v': v, gn + '/m': m, gn + '/name': names,
                 gn + '/attrs': np.array(json.dumps(attrs))})
    with open(fn, 'wb') as f:
        np.savez(f, **data)


def groups(fn):
    """Group names stored in ``fn``."""
    h5 = _h5py()
    if h5 is not None and not fn.endswith('.npz'):
        with h5.File(fn, 'r') as f:
            return list(f.keys())
    with np.load(fn, allow_pickle=False) as d:
        return sorted({k.rsplit('/', 1)[0] for k in d.files})


def load_simulation(universe_class, fn, gn, threads=None):
    """A universe of ``universe_class`` holding group ``gn`` of ``fn`` (not started); the
    stored, already scaled, values are added with ``scale_off`` (:186-214)."""
    h5 = _h5py()
    if h5 is not None and not fn.endswith('.npz'):
        with h5.File(fn, 'r') as f:
            g = f[gn]
            attrs = {a: g.attrs[a] for a in g.attrs.keys()}
            r, v, m, names = g['r'][()], g['v'][()], g['m'][()], g['name'][()]
    else:
        with np.load(fn, allow_pickle=False) as d:
            attrs = json.loads(str(d[gn + '/attrs']))
            r, v, m, names = d[gn + '/r'], d[gn + '/v'], d[gn + '/m'], d[gn + '/name']
    if isinstance(threads, int):
        attrs['threads'] = threads
    if isinstance(attrs.get('dtype'), bytes):
        attrs['dtype'] = attrs['dtype'].decode()
    u = universe_class(scale_off=True, **attrs)
    for k in range(r.shape[0]):
        u.add_object(scale_off=True, name=bytes(names[k]).decode('utf-8'),
                     r=[float(x) for x in r[k]], v=[float(x) for x in v[k]], m=float(m[k]))
    return u
