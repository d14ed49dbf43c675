"""grav_stage1_f32_split_noalloc (partials in the library's static device array) against
grav_stage1_f32_split (partials in a caller workspace) through the C ABI: the same kernels,
so the same bits, on consecutive calls whose second input is smaller (the array keeps the
first call's partials, which the second must not see), in the default and explicit shapes
and with eps = 0 and > 0; requests beyond the array's size are rejected."""

import ctypes

import numpy as np
import pytest
import torch

from gravitation_amd import _native, stages

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20


def _posm(n, seed, dev):
    rng = np.random.default_rng(seed)
    p = torch.empty((n, 4), dtype=torch.float32)
    p[:, :3] = torch.from_numpy(rng.normal(size=(n, 3)))
    p[:, 3] = torch.from_numpy(rng.uniform(0.5, 1.5, n))
    return p.to(dev)


def _split(posm, G, eps, cfg):
    n = posm.shape[0]
    acc = torch.full((3, n), float('nan'), dtype=torch.float32, device=posm.device)
    ws = torch.empty((max(256, stages.split_workspace_bytes(n, cfg if cfg[2] else None)),),
                     dtype=torch.uint8, device=posm.device)
    _native.call('grav_stage1_f32_split', posm.data_ptr(), acc.data_ptr(), n, G, eps, *cfg,
                 ws.data_ptr(), ws.numel(), torch.cuda.current_stream().cuda_stream)
    return acc


def _noalloc(posm, G, eps, cfg):
    n = posm.shape[0]
    acc = torch.full((3, n), float('nan'), dtype=torch.float32, device=posm.device)
    _native.call('grav_stage1_f32_split_noalloc', posm.data_ptr(), acc.data_ptr(), n, G, eps,
                 *cfg, torch.cuda.current_stream().cuda_stream)
    return acc


@pytest.mark.parametrize('cfg', [(0, 0, 0), (2, 8, 16), (1, 4, 7)])
@pytest.mark.parametrize('eps', [0.0, 1e-3])
def test_noalloc_equals_workspace_form(cuda, cfg, eps):
    for n, seed in ((30000, 1), (2500, 2), (777, 3)):     # each call smaller than the last
        print('noalloc n=%d cfg=%s eps=%g' % (n, cfg, eps), flush=True)
        posm = _posm(n, seed, cuda)
        got = _noalloc(posm, 0.8, eps, cfg)
        want = _split(posm, 0.8, eps, cfg)
        torch.cuda.synchronize()
        assert torch.isfinite(got).all()
        assert torch.equal(got, want), (n, cfg, eps)


def test_noalloc_rejects_sizes_beyond_the_array(cuda):
    posm = _posm(8, 4, cuda)
    acc = torch.empty((3, 8), dtype=torch.float32, device=cuda)
    big = SCRATCH_BYTES // (32 * 64) + 1          # 64 j-ranges of this many bodies overflow
    with pytest.raises(_native.GravInvalidArgument, match='array'):
        _native.call('grav_stage1_f32_split_noalloc', posm.data_ptr(), acc.data_ptr(), big, 1.0,
                     0.0, 1, 8, 64, ctypes.c_void_p(None))
    # the largest size that fits still runs (pointers of a large enough input)
    n = SCRATCH_BYTES // (32 * 64)
    posm = _posm(n, 5, cuda)
    got = _noalloc(posm, 1.0, 1e-3, (1, 8, 64))
    torch.cuda.synchronize()
    assert torch.equal(got, _split(posm, 1.0, 1e-3, (1, 8, 64)))
