"""Worker JSON-lines harness, seeded scenarios and state files, on the CPU with a CPU test
kernel.  A worker log is accepted here by the reference's analyze rules
(src/gravitation/cli/analyze.py:47-106), restated in ``_analyze``."""

import io
import json
import math

import numpy as np
import pytest

from gravitation_amd.kernel import _base_ as B
from gravitation_amd.lib import simulation, worker
from oracle import oracle as O


class cpu_universe(B.universe_base):
    """FP64 oracle stage 1, base stage 2 (test kernel)."""

    def step_stage1(self):
        pos = np.array([pm._r for pm in self], dtype=np.float64)
        m = np.array([pm._m for pm in self])
        for pm, ai in zip(self, O.accel_f64(pos, m, self._G)):
            pm._a[:] = list(ai)


class failing_universe(cpu_universe):
    def step_stage1(self):
        if self._t > 0:
            raise RuntimeError('boom')
        super().step_stage1()


def _analyze(text):
    """The reference analyze's validity rules for one worker run."""
    lines = [json.loads(ln) for ln in text.split('\n') if ln.strip()]
    assert lines[0] == {'log': 'START'}
    assert not [d for d in lines if d['log'] == 'ERROR']
    inputs = [d for d in lines if d['log'] == 'INPUT']
    sizes = [d for d in lines if d['log'] == 'SIZE']
    assert len(inputs) == 1 and len(sizes) == 1
    counter = [d['counter'] for d in lines if d['log'] == 'STEP']
    assert counter and counter == list(range(counter[0], counter[0] + len(counter)))
    assert lines[-1] == {'log': 'EXIT', 'msg': 'OK'}
    best = [d['value'] for d in lines if d['log'] == 'BEST_TIME']
    runtimes = [d['runtime'] for d in lines if d['log'] == 'STEP']
    assert best[-1] == min(runtimes)
    return inputs[0], sizes[0]['value'], runtimes


def test_worker_log_is_accepted_and_stores_state(tmp_path):
    out = io.StringIO()
    fn = str(tmp_path / 'data.npz')
    rc = worker.run_worker('cpu', 'solarsystem', {}, fn, save_after_iteration=(0, 2),
                           min_iterations=3, min_total_runtime=0, universe_class=cpu_universe,
                           out=out)
    assert rc == 0
    inp, size, rts = _analyze(out.getvalue())
    assert size == 2 and len(rts) == 3
    assert inp['simulation']['scenario'] == 'solarsystem'
    gs = simulation.groups(fn)
    assert gs == ['kernel=cpu;len=2;step=0', 'kernel=cpu;len=2;step=2']
    u = simulation.load_simulation(cpu_universe, fn, gs[1])
    assert len(u) == 2 and u._t == pytest.approx(2 * 1e3) and u._dtype == 'float32'
    earth = list(u)[1]
    assert earth._name == 'earth'


def test_worker_reports_errors():
    out = io.StringIO()
    rc = worker.run_worker('cpu', 'solarsystem', {}, 'unused.npz', min_iterations=3,
                           min_total_runtime=0, universe_class=failing_universe, out=out)
    assert rc == 1
    lines = [json.loads(ln) for ln in out.getvalue().split('\n') if ln.strip()]
    assert lines[-2]['log'] == 'ERROR' and 'boom' in lines[-2]['msg']
    assert lines[-1] == {'log': 'EXIT', 'msg': 'BAD'}


def test_inventory_lists_hip1():
    assert 'hip1' in worker.inventory()


def test_galaxy_is_seeded_and_physical():
    a = simulation.create_simulation('galaxy', B.universe_base, {'stars_len': 200}, seed=7)
    b = simulation.create_simulation('galaxy', B.universe_base, {'stars_len': 200}, seed=7)
    c = simulation.create_simulation('galaxy', B.universe_base, {'stars_len': 200}, seed=8)
    ra = np.array([pm._r for pm in a])
    assert len(a) == 200 and np.array_equal(ra, np.array([pm._r for pm in b]))
    assert not np.array_equal(ra, np.array([pm._r for pm in c]))
    hole = list(a)[0]
    assert hole._name == 'back hole' and hole._m == pytest.approx(4e40 * 1e-30)
    # every star starts on a circular orbit about the hole: v^2 |r| = G M_hole
    for pm in list(a)[1:]:
        r = math.sqrt(sum(x * x for x in pm._r))
        v2 = sum(x * x for x in pm._v)
        assert v2 * r == pytest.approx(a._G * hole._m, rel=1e-9)
    assert a._meta.get('stars_len') == 200      # as in the reference, it lands in _meta


def test_plummer_and_cube_scenarios():
    u = simulation.create_simulation('plummer', B.universe_base, {'stars_len': 300}, seed=3)
    assert len(u) == 300 and u._G == 1.0 and u._meta['eps'] == 1e-3
    u = simulation.create_simulation('cube', B.universe_base, {'stars_len': 50}, seed=3)
    assert len(u) == 50 and u._G == pytest.approx(6.6743e-11)
    with pytest.raises(ValueError):
        simulation.create_simulation('nope', B.universe_base)


def test_store_load_roundtrip(tmp_path):
    u = simulation.create_simulation('galaxy', cpu_universe, {'stars_len': 20}, seed=1)
    u.step()
    fn = str(tmp_path / 's.npz')
    simulation.store_simulation(u, fn, 'kernel=cpu;len=20;step=1')
    with pytest.raises(ValueError):
        simulation.store_simulation(u, fn, 'kernel=cpu;len=20;step=1')
    w = simulation.load_simulation(cpu_universe, fn, 'kernel=cpu;len=20;step=1')
    assert w._G == u._G and w._T == u._T and w._t == u._t
    assert w._scale_m == u._scale_m and w._scale_r == u._scale_r
    for p, q in zip(u, w):
        assert p._name == q._name
        np.testing.assert_array_equal(np.float32(p._r), np.float32(q._r))
        assert np.float32(p._m) == np.float32(q._m)


class _FakeH5:
    """The part of h5py's API that store/load use, in memory (h5py is not installed here):
    File(fn, mode) as a context manager, create_group / keys / [name] on files, create_dataset,
    attrs and [name][()] on groups.  Datasets keep copies, as HDF5 does."""

    files = {}

    class _Group:
        def __init__(self):
            self.data, self.attrs = {}, {}

        def create_dataset(self, name, data):
            if name in self.data:
                raise ValueError('dataset %r exists' % name)
            self.data[name] = np.array(data, copy=True)

        def __getitem__(self, name):
            return self.data[name]

    class File:
        def __init__(self, fn, mode):
            if mode == 'r' and fn not in _FakeH5.files:
                raise OSError('no file %s' % fn)
            self.groups = _FakeH5.files.setdefault(fn, {})

        def __enter__(self):
            return self

        def __exit__(self, *exc):
            return False

        def create_group(self, name):
            if name in self.groups:
                raise ValueError('group %r exists' % name)
            g = _FakeH5._Group()
            self.groups[name] = g
            return g

        def keys(self):
            return self.groups.keys()

        def __getitem__(self, name):
            return self.groups[name]


def test_store_load_hdf5_branch(monkeypatch, tmp_path):
    # The HDF5 branch of store/load (taken when h5py imports) against an in-memory stand-in
    # of h5py: the same group names, datasets r/v/m/name and attributes as the .npz mirror,
    # and the same universe back.
    monkeypatch.setattr(simulation, '_h5py', lambda: _FakeH5)
    monkeypatch.setattr(_FakeH5, 'files', {})
    u = simulation.create_simulation('galaxy', cpu_universe, {'stars_len': 20}, seed=1)
    u.step()
    fn = str(tmp_path / 's.h5')
    gn = 'kernel=cpu;len=20;step=1'
    simulation.store_simulation(u, fn, gn)
    with pytest.raises(ValueError):
        simulation.store_simulation(u, fn, gn)
    simulation.store_simulation(u, fn, 'second')
    assert sorted(simulation.groups(fn)) == sorted([gn, 'second'])
    g = _FakeH5.files[fn][gn]
    assert set(g.data) == {'r', 'v', 'm', 'name'}
    assert g.data['r'].shape == (len(list(u)), 3) and g.data['name'].dtype.kind == 'S'
    assert g.attrs['G'] == u._G and g.attrs['t'] == u._t
    w = simulation.load_simulation(cpu_universe, fn, gn)
    npz = str(tmp_path / 's.npz')
    simulation.store_simulation(u, npz, gn)
    x = simulation.load_simulation(cpu_universe, npz, gn)
    for a in (w, x):
        assert a._G == u._G and a._T == u._T and a._t == u._t
        assert a._scale_m == u._scale_m and a._scale_r == u._scale_r
    for p, q, s in zip(u, w, x):
        assert p._name == q._name == s._name
        np.testing.assert_array_equal(np.float32(q._r), np.float32(s._r))
        np.testing.assert_array_equal(np.float32(q._v), np.float32(s._v))
        assert q._m == s._m
