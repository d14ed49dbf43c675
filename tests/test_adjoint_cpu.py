"""Host-side logic of the adjoint: checkpoint interval from the HBM budget."""

import numpy as np
import pytest

from gravitation_amd import adjoint


def test_checkpoint_interval_choice():
    # the whole trajectory fits: record it, no recompute
    assert adjoint.checkpoint_interval(1000, 200, 10 ** 12) == (200, True)
    # it does not: checkpoints every ceil(sqrt(3T)) steps
    c, store = adjoint.checkpoint_interval(1000, 200, 200 * 16 * 1000 - 1)
    assert not store and c == int(np.ceil(np.sqrt(600)))
    with pytest.raises(MemoryError):
        adjoint.checkpoint_interval(1000, 200, 1000)
    assert adjoint.checkpoint_interval(10, 0, 0) == (1, True)
    # two record buffers (the paired backward): C = ceil(sqrt(3T/2)), counted twice
    c, store = adjoint.checkpoint_interval(1000, 200, 200 * 16 * 1000 - 1, 2)
    assert not store and c == int(np.ceil(np.sqrt(300)))
    need = int(np.ceil(200 / c)) * 48 * 1000 + 2 * c * 16 * 1000
    assert adjoint.checkpoint_interval(1000, 200, need, 2) == (c, False)
    with pytest.raises(MemoryError):
        adjoint.checkpoint_interval(1000, 200, need - 1, 2)


def test_trajectory_form_is_checked_first():
    z = np.zeros((4, 3))
    with pytest.raises(ValueError, match='form'):
        adjoint.trajectory(z, z, np.ones(4), 3, form='split')
