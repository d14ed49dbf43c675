"""The drop-in kernel ``hip1`` through the reference's front-end contract on the GPU:
add_object / start / step / stop, per-step positions and velocities in the iterable,
per-step accelerations against the FP64 oracle, trajectories against the oracle."""

import numpy as np
import pytest
import torch

from gravitation_amd.kernel import hip1
from gravitation_amd.lib import scenarios as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _make(r, v, m, **kw):
    u = hip1.universe(**kw)
    for k in range(r.shape[0]):
        u.add_object(name='b%d' % k, r=list(r[k]), v=list(v[k]), m=float(m[k]))
    return u


def test_c1_per_step_accelerations(cuda):
    # 1024 bodies in a cube of side 1e3, G = 6.6743e-11, dt = 1, 10 steps, FP32; every step's
    # accelerations against the FP64 oracle on the identical FP32 positions.
    r, v, m = S.uniform_cube(1024, seed=0)
    u = _make(r, v, m, T=1.0, G=6.6743e-11, dtype='float32')
    u.start()
    m32 = np.float32(m).astype(np.float64)
    for step in range(10):
        pos = np.array([pm._r for pm in u], dtype=np.float64)   # FP32 state, host view
        u.step()
        a = u.acc.cpu().numpy().T.astype(np.float64)
        err = O.rel_err(a, O.accel_f64(pos, m32, 6.6743e-11))
        assert err < 1e-6, (step, err)
    assert u._t == 10.0
    u.stop()


def test_c1_trajectory_fp64_masters(cuda):
    r, v, m = S.uniform_cube(1024, seed=0)
    u = _make(r, v, m, T=1.0, G=6.6743e-11, dtype='float32', stage2_dtype='float64')
    u.start()
    for _ in range(10):
        u.step()
    rg = np.array([pm._r for pm in u])
    vg = np.array([pm._v for pm in u])
    rr, vr, _ = O.simulate(r, v, m, 6.6743e-11, 0.0, 1.0, 10, dtype=np.float32)
    assert np.abs(rg - rr).max() < 1e-9 * np.abs(rr).max()
    assert O.rel_err(vg, vr) < 1e-5
    u.stop()


def test_fp64_universe_trajectory(cuda):
    r, v, m = S.plummer(2048, seed=4)
    v = np.random.default_rng(0).normal(0, 0.1, v.shape)
    u = _make(r, v, m, T=1e-4, G=1.0, dtype='float64', eps=1e-3)
    u.start()
    for _ in range(20):
        u.step()
    rg = np.array([pm._r for pm in u])
    vg = np.array([pm._v for pm in u])
    rr, vr, _ = O.simulate(r, v, m, 1.0, 1e-3, 1e-4, 20, dtype=np.float64)
    assert O.rel_err(rg, rr) < 1e-13
    assert O.rel_err(vg, vr) < 1e-11
    u.stop()


def test_fp32_stage2_matches_reference_semantics(cuda):
    # FP32 state: one step equals oracle stage 1 (on the FP32 inputs) + FP32 stage 2.
    r, v, m = S.uniform_cube(256, seed=1)
    v = np.random.default_rng(2).normal(0, 1e-3, v.shape)
    u = _make(r, v, m, T=1.0, G=6.6743e-11, dtype='float32', sync_host=False)
    u.start()
    r32 = np.array([pm._r for pm in u])
    v32 = np.array([pm._v for pm in u])
    u.step()
    u.sync_to_host()
    a32 = u.acc.cpu().numpy().T
    rn, vn = O.stage2(r32, v32, a32, np.float32(1.0))
    assert np.array_equal(np.array([pm._r for pm in u]), rn)
    assert np.array_equal(np.array([pm._v for pm in u]), vn)
    assert all(np.all(pm._a == 0) for pm in u)
    u.stop()


def test_option_errors(cuda):
    r, v, m = S.uniform_cube(8, seed=1)
    for kw in ({'dtype': 'float16'}, {'accum': 'fast'}, {'stage2_dtype': 'int8'},
               {'dtype': 'float64', 'stage2_dtype': 'float32'}, {'stage1': 'fast'},
               {'stage1_cfg': (4, 4, 256)}, {'stage1': 'symmetric', 'accum': 'plain'},
               {'stage1': 'split', 'accum': 'plain'},
               {'stage1': 'split', 'dtype': 'float64', 'stage2_dtype': 'float64',
                'stage1_cfg': (1, 2, 8)},
               {'stage1': 'split', 'stage1_cfg': (4, 8, 4)},
               {'stage1': 'tune', 'accum': 'plain'}, {'stage1': 'tune', 'stage1_cfg': (1, 8, 4)}):
        u = _make(r, v, m, **kw)
        with pytest.raises(ValueError):
            u.start()


def test_worker_galaxy_with_hip1(cuda, tmp_path):
    # The reference's default workload through the JSON-lines worker with the hip1 kernel;
    # the log satisfies the reference analyze's rules and snapshots round-trip.
    import io
    import json

    from gravitation_amd.lib import simulation, worker
    out = io.StringIO()
    fn = str(tmp_path / 'g.npz')
    rc = worker.run_worker('hip1', 'galaxy', {'stars_len': 4096}, fn, save_after_iteration=(3,),
                           min_iterations=3, min_total_runtime=0, seed=1, out=out)
    lines = [json.loads(ln) for ln in out.getvalue().split('\n') if ln.strip()]
    assert rc == 0 and lines[-1] == {'log': 'EXIT', 'msg': 'OK'}
    assert [d['counter'] for d in lines if d['log'] == 'STEP'] == [1, 2, 3]
    assert [d['value'] for d in lines if d['log'] == 'SIZE'] == [4096]
    assert not [d for d in lines if d['log'] == 'ERROR']
    u = simulation.load_simulation(hip1.universe, fn, 'kernel=hip1;len=4096;step=3')
    assert len(u) == 4096 and u._t == 3 * 2.0e12
    u.start()
    u.step()
    assert np.all(np.isfinite(np.array([pm._r for pm in u])))
    u.stop()


def test_split_form_in_the_universe(cuda):
    # 5000 bodies: 'auto' takes the j-split form; its trajectory equals owner-computes' to
    # FP32 rounding, with the default and an explicit shape, in eager and graph stepping
    r, v, m = S.plummer(5000, seed=19)
    a = _make(r, v, m, T=1e-4, G=1.0, eps=1e-3, stage2_dtype='float64', stage1='owner')
    b = _make(r, v, m, T=1e-4, G=1.0, eps=1e-3, stage2_dtype='float64')
    c = _make(r, v, m, T=1e-4, G=1.0, eps=1e-3, stage2_dtype='float64', stage1='split',
              stage1_cfg=(2, 4, 3), graph=True)
    for u in (a, b, c):
        u.start()
    assert b._form == 'split' and c._form == 'split' and c._graphs is not None
    for _ in range(3):
        for u in (a, b, c):
            u.step()
    ra, rb, rc = (np.array([pm._r for pm in u]) for u in (a, b, c))
    assert O.rel_err(rb - r, ra - r) < 1e-5 and O.rel_err(rc - r, ra - r) < 1e-5
    for u in (a, b, c):
        u.stop()


def test_symmetric_form_in_the_universe(cuda):
    r, v, m = S.plummer(5000, seed=9)
    a = _make(r, v, m, T=1e-4, G=1.0, eps=1e-3, stage2_dtype='float64')
    b = _make(r, v, m, T=1e-4, G=1.0, eps=1e-3, stage2_dtype='float64', stage1='symmetric')
    a.start()
    b.start()
    for _ in range(3):
        a.step()
        b.step()
    ra = np.array([pm._r for pm in a])
    rb = np.array([pm._r for pm in b])
    assert O.rel_err(rb - r, ra - r) < 1e-5
    a.stop()
    b.stop()


def test_auto_form_by_universe_size(cuda):
    from gravitation_amd import stages
    for n, form in ((stages.SPLIT_FROM - 1, 'owner'), (stages.SPLIT_FROM, 'split'),
                    (stages.SYMMETRIC_FROM - 1, 'split'), (stages.SYMMETRIC_FROM, 'symmetric')):
        r, v, m = S.plummer(n, seed=3)
        u = hip1.universe(T=1e-4, G=1.0, eps=1e-3, stage2_dtype='float64')
        u.add_objects(r, v, m)
        u.start()
        assert u._form == form
        u.step()
        pos = np.array([pm._r for pm in u])
        assert np.all(np.isfinite(pos))
        u.stop()


def test_fp64_forms_in_the_universe(cuda):
    # dtype='float64' with 5000 bodies: 'auto' takes the FP64 j-split form; its trajectory,
    # the symmetric form's and an explicit split shape's equal the owner-computes one to FP64
    # rounding.
    r, v, m = S.plummer(5000, seed=10)
    v = np.random.default_rng(1).normal(0, 0.1, v.shape)
    kw = dict(T=1e-4, G=1.0, dtype='float64', eps=1e-3)
    a = _make(r, v, m, stage1='owner', **kw)
    us = [_make(r, v, m, **kw), _make(r, v, m, stage1='symmetric', **kw),
          _make(r, v, m, stage1='split', stage1_cfg=(2, 4, 7), **kw)]
    for u in [a] + us:
        u.start()
    assert [u._form for u in us] == ['split', 'symmetric', 'split']
    for _ in range(5):
        for u in [a] + us:
            u.step()
    ra = np.array([pm._r for pm in a])
    for u in us:
        assert O.rel_err(np.array([pm._r for pm in u]) - r, ra - r) < 1e-11, u._form
    for u in [a] + us:
        u.stop()


def test_auto_form_fp64_by_universe_size(cuda):
    from gravitation_amd import stages
    for n, form in ((stages.SPLIT_F64_FROM - 1, 'owner'), (stages.SPLIT_F64_FROM, 'split'),
                    (stages.SYMMETRIC_F64_FROM, 'symmetric')):
        r, v, m = S.plummer(n, seed=5)
        u = _make(r, v, m, T=1e-4, G=1.0, dtype='float64', eps=1e-3)
        u.start()
        assert u._form == form
        u.step()
        assert np.all(np.isfinite(np.array([pm._r for pm in u])))
        u.stop()


@pytest.mark.parametrize('dtype, stage2, n', [('float32', 'float32', 1000),
                                              ('float32', 'float64', 1000),
                                              ('float32', 'float64', 30000),
                                              ('float32', 'float64', 41000),
                                              ('float64', 'float64', 5000),
                                              ('float64', 'float64', 13000),
                                              ('float64', 'float64', 700)])
def test_graph_steps_equal_eager_steps(cuda, dtype, stage2, n):
    # The captured stages replay to exactly the eager results, step after step, with the host
    # views read (and the device synchronised) between replays; the sizes take every stage-1
    # form of both precisions (owner-computes, j-split, symmetric).
    r, v, m = S.plummer(n, seed=11)
    v = np.random.default_rng(3).normal(0, 0.1, v.shape)
    kw = dict(T=1e-4, G=1.0, eps=1e-3, dtype=dtype, stage2_dtype=stage2)
    g = _make(r, v, m, graph=True, **kw)
    e = _make(r, v, m, graph=False, **kw)
    g.start()
    e.start()
    assert g._graphs is not None and e._graphs is None
    for _ in range(6):
        g.step()
        e.step()
        assert np.array_equal(np.array([pm._r for pm in g]), np.array([pm._r for pm in e]))
        assert np.array_equal(np.array([pm._v for pm in g]), np.array([pm._v for pm in e]))
        assert torch.equal(g.acc, e.acc)
    g.stop()
    e.stop()


@pytest.mark.parametrize('dtype', ['float32', 'float64'])
def test_tuned_stage1_form_in_the_universe(cuda, dtype):
    # stage1='tune' times the forms on the universe's own state at start() and keeps the
    # fastest; the trajectory equals owner-computes' to the form's rounding
    r, v, m = S.plummer(6000, seed=23)
    kw = dict(T=1e-4, G=1.0, eps=1e-3, dtype=dtype, stage2_dtype='float64')
    a = _make(r, v, m, stage1='owner', **kw)
    b = _make(r, v, m, stage1='tune', **kw)
    a.start()
    b.start()
    assert b._form in ('owner', 'split', 'symmetric')
    for _ in range(3):
        a.step()
        b.step()
    ra = np.array([pm._r for pm in a])
    rb = np.array([pm._r for pm in b])
    assert O.rel_err(rb - r, ra - r) < (1e-11 if dtype == 'float64' else 1e-5)
    a.stop()
    b.stop()


@pytest.mark.parametrize('dim', [1, 2])
@pytest.mark.parametrize('dtype', ['float32', 'float64'])
def test_lower_dimensional_universes(cuda, dim, dtype):
    # the reference's dimension-generic kernels allow 1-D and 2-D universes; hip1 holds the
    # missing coordinates at 0 and matches the oracle integrating the same D-dimensional arrays
    rng = np.random.default_rng(dim)
    n = 300
    r = rng.normal(size=(n, dim))
    v = rng.normal(0, 0.1, (n, dim))
    m = rng.uniform(0.5, 1.5, n) / n
    u = _make(r, v, m, T=1e-3, G=1.0, eps=1e-2, dtype=dtype, stage2_dtype='float64')
    u.start()
    assert all(len(pm._r) == dim and len(pm._v) == dim and len(pm._a) == dim for pm in u)
    for _ in range(5):
        u.step()
    rg = np.array([pm._r for pm in u])
    vg = np.array([pm._v for pm in u])
    rr, vr, _ = O.simulate(r, v, m, 1.0, 1e-2, 1e-3, 5, dtype=np.float64)
    tol = 1e-11 if dtype == 'float64' else 1e-5
    assert O.rel_err(vg - v, vr - v) < tol
    assert O.rel_err(rg - r, rr - r) < tol
    u.stop()


GALAXY_FORM = {1024: 'owner', 4096: 'split', 16384: 'split'}


@pytest.mark.parametrize('stars', [1024, 4096, 16384])
@pytest.mark.parametrize('form', ['auto', 'owner', 'split', 'symmetric', 'float64'])
def test_galaxy_accelerations_vs_oracle(cuda, stars, form):
    # The reference's default workload (src/gravitation/lib/simulation.py:57-83, :99-184) at
    # its published sizes through hip1: scale_r 1e-10, scale_m 1e-30 (G unchanged), eps = 0
    # (self term masked), a central mass 2e10 times each star's.  Every step's accelerations
    # against the FP64 oracle on the positions and masses the step used: <= 1e-5 in FP32 (the
    # form 'auto' takes, and each form forced), <= 1e-12 with dtype='float64'.
    from conftest import oracle_rows
    from gravitation_amd.lib import simulation
    print('galaxy stars=%d form=%s' % (stars, form), flush=True)
    param = {'stars_len': stars}
    if form == 'float64':
        param['dtype'] = 'float64'
    else:
        param['stage1'] = form
    u = simulation.create_simulation('galaxy', hip1.universe, param, seed=stars)
    if form == 'auto':
        assert u._form == GALAXY_FORM[stars]
    assert u._eps == 0.0 and u._G == pytest.approx(6.6740831e-11, rel=1e-15)
    rows = None if stars <= 4096 else np.unique(np.concatenate(
        [[0, 1, 2, 63, 64, stars - 1], np.random.default_rng(stars).choice(stars, 2000, False)]))
    bar = 1e-12 if form == 'float64' else 1e-5
    worst = 0.0
    for step in range(3):
        p = u.posm.cpu().numpy().astype(np.float64)
        u.step()
        a = u.acc.cpu().numpy().T.astype(np.float64)
        ref = oracle_rows(p[:, :3], p[:, 3], u._G, 0.0, rows)
        err = O.rel_err(a if rows is None else a[rows], ref)
        worst = max(worst, err)
        assert err <= bar, (step, err)
    print('galaxy stars=%d form=%s (%s): worst rel err %.2e' % (stars, form, u._form, worst),
          flush=True)
    u.stop()


@pytest.mark.parametrize('graph', [False, True])
@pytest.mark.parametrize('dtype, stage2', [('float32', 'float32'), ('float32', 'float64'),
                                           ('float64', 'float64')])
def test_host_views_mirror_device_state(cuda, dtype, stage2, graph):
    # With sync_host the stage-2 kernel writes the new state straight into the mapped host
    # mirror: after every step the per-body views equal the device state bit for bit (r from
    # the positions, or from the FP64 masters in the mixed mode; v), eager and graph-replayed
    r, v, m = S.plummer(700, seed=9)
    u = _make(r, v, m, T=1e-3, G=1.0, eps=1e-3, dtype=dtype, stage2_dtype=stage2, graph=graph)
    u.start()
    for _ in range(4):
        u.step()
        dev_r = (u.r64.t() if stage2 == 'float64' and dtype == 'float32'
                 else u.posm[:, :3]).cpu().numpy()
        dev_v = u.vel.t().cpu().numpy()
        host_r = np.array([pm._r for pm in u])
        host_v = np.array([pm._v for pm in u])
        assert host_r.dtype == dev_r.dtype and host_v.dtype == dev_v.dtype
        assert np.array_equal(host_r, dev_r) and np.array_equal(host_v, dev_v)
    u.stop()
    # the views outlive stop() and keep the final state
    assert np.array_equal(np.array([pm._r for pm in u]), dev_r)
