"""The hip1 package against the reference's plugin contract, on the CPU.

The reference's loader lists src/gravitation/kernel/ entries not starting with '_'
(src/gravitation/lib/load.py:45-48), reads nine meta dunders from the package's __init__.py
as literals without importing it (load.py:77-99, :125-144) and instantiates the module-level
class ``universe`` (load.py:62-66)."""

import ast
import os

import pytest

from gravitation_amd.kernel import _base_
from gravitation_amd.kernel import hip1

HERE = os.path.dirname(os.path.abspath(__file__))
INIT = os.path.join(HERE, '..', 'gravitation_amd', 'kernel', 'hip1', '__init__.py')
DUNDERS = ('longname', 'version', 'description', 'requirements', 'externalrequirements',
           'interpreters', 'parallel', 'license', 'authors')


def test_meta_dunders_are_literals():
    tree = ast.parse(open(INIT).read())
    found = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1:
            name = getattr(node.targets[0], 'id', '')
            if name.startswith('__') and name.endswith('__'):
                found[name[2:-2]] = ast.literal_eval(node.value)
    for d in DUNDERS:
        assert d in found, d
    assert found['parallel'] is False
    assert isinstance(found['authors'], list) and isinstance(found['version'], str)


def test_universe_is_a_base_subclass_with_the_contract_hooks():
    u = hip1.universe
    assert issubclass(u, _base_.universe_base)
    for hook in ('start_kernel', 'step_stage1', 'step_stage2', 'stop_kernel'):
        assert hook in u.__dict__
    for fixed in ('__init__', 'step', 'start', 'stop', 'add_object', '__iter__', '__len__'):
        assert fixed not in u.__dict__


def test_start_without_a_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip('a HIP device is visible')
    u = hip1.universe(G=1.0)
    u.add_object(name='a', r=[0.0, 0.0, 0.0], v=[0.0, 0.0, 0.0], m=1.0)
    with pytest.raises(RuntimeError, match='HIP device'):
        u.start()


def test_auto_stage1_form_threshold():
    from gravitation_amd import stages
    assert stages.auto_form(stages.SPLIT_FROM - 1) == 'owner'
    assert stages.auto_form(stages.SPLIT_FROM) == 'split'
    assert stages.auto_form(stages.SYMMETRIC_FROM - 1) == 'split'
    assert stages.auto_form(stages.SYMMETRIC_FROM) == 'symmetric'
    assert stages.auto_form(1 << 18, 'plain') == 'owner'
    assert stages.auto_form(stages.SPLIT_FROM, 'plain') == 'owner'


def test_auto_form_needs_the_workspace_to_fit():
    # The symmetric form works in slabs of rows down to about 50 bytes per body, so it is
    # taken for any size whose smallest workspace fits half the free memory; its one-slab
    # workspace (about 116 GB at 2^22 bodies) is allocated only when it fits a quarter of the
    # free memory, else that quarter is.
    from gravitation_amd import stages
    free = 280 * 2 ** 30
    n = 1 << 22
    full, least = stages.sym_workspace_bytes(n), stages.sym_workspace_bytes(n, minimum=True)
    assert full > free // 4 and least < 0.25e9
    assert stages.auto_form(1 << 18, 'kahan', free) == 'symmetric'
    assert stages.auto_form(n, 'kahan', free) == 'symmetric'
    assert stages.auto_form(n, 'kahan', 2 * least - 1) == 'owner'
    assert stages.auto_form(n, 'kahan') == 'symmetric'
    assert stages.sym_workspace_size(n, free_bytes=free) == (free // 4) & ~255
    assert stages.sym_workspace_size(1 << 18, free_bytes=free) == stages.sym_workspace_bytes(
        1 << 18)
    assert stages.sym_workspace_size(n, free_bytes=least) == least


def test_auto_form_f64_threshold():
    from gravitation_amd import stages
    assert stages.auto_form(stages.SPLIT_F64_FROM - 1, dtype='float64') == 'owner'
    assert stages.auto_form(stages.SPLIT_F64_FROM, dtype='float64') == 'split'
    assert stages.auto_form(stages.SYMMETRIC_F64_FROM - 1, dtype='float64') == 'split'
    assert stages.auto_form(stages.SYMMETRIC_F64_FROM, 'plain', dtype='float64') == 'symmetric'
    free = 280 * 2 ** 30
    least = stages.sym_f64_workspace_bytes(1 << 22, minimum=True)
    assert stages.auto_form(1 << 22, 'kahan', free, 'float64') == 'symmetric'
    assert stages.auto_form(1 << 22, 'kahan', 2 * least - 1, 'float64') == 'owner'
    assert stages.sym_workspace_size(1 << 22, free_bytes=free, dtype='float64') == \
        (free // 4) & ~255


def test_bench_refuses_more_than_one_gpu():
    # one GPU per box: --gpus N > 1 exits with a message instead of measuring anything
    import subprocess
    import sys
    root = os.path.join(HERE, '..')
    p = subprocess.run([sys.executable, os.path.join(root, 'bench.py'), '--gpus', '2'],
                       cwd=root, capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert 'one GPU' in p.stderr
