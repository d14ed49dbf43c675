"""Host side of dividing work among GPUs, on the CPU: whole universes of a batch bin-packed
into shares, and the arrays of one share."""

import itertools

import numpy as np
import pytest

from gravitation_amd import batched
from gravitation_amd.lib import scenarios as S


def _offsets(sizes):
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


@pytest.mark.parametrize('shares', [1, 3, 8, 64])
def test_shard_universes_partition_and_bound(shares):
    off = _offsets(S.batch_sizes(4096, 32, 2048, seed=0))
    bins = batched.shard_universes(off, shares)
    assert len(bins) == shares
    every = np.sort(np.concatenate(bins))
    assert np.array_equal(every, np.arange(4096))
    cost = np.diff(off) ** 2
    loads = np.array([cost[b].sum() for b in bins])
    for b in bins:
        assert np.all(np.diff(b) > 0)
    # list scheduling: the heaviest share is at most the mean plus one universe
    assert loads.max() <= cost.sum() / shares + cost.max()
    again = batched.shard_universes(off, shares)
    assert all(np.array_equal(a, b) for a, b in zip(bins, again))


def test_shard_universes_against_brute_force():
    # longest processing time first is within (4/3 - 1/(3P)) of the best partition (Graham);
    # checked against every assignment of a few universes
    rng = np.random.default_rng(5)
    for trial in range(30):
        u = int(rng.integers(2, 8))
        p = int(rng.integers(2, 4))
        sizes = rng.integers(1, 40, u)
        cost = sizes.astype(np.int64) ** 2
        bins = batched.shard_universes(_offsets(sizes), p)
        lpt = max(int(cost[b].sum()) for b in bins)
        best = min(max(sum(int(cost[i]) for i in range(u) if a[i] == k) for k in range(p))
                   for a in itertools.product(range(p), repeat=u))
        assert lpt * 3 * p <= best * (4 * p - 1), (sizes, p, lpt, best)


def test_shard_universes_more_shares_than_universes_and_errors():
    bins = batched.shard_universes(_offsets([5, 100, 7]), 5)
    assert [b.tolist() for b in bins] == [[1], [2], [0], [], []]
    with pytest.raises(ValueError):
        batched.shard_universes(_offsets([5, 7]), 0)
    with pytest.raises(ValueError):
        batched.shard_universes(np.array([1, 5]), 2)


def test_take_universes_layout():
    rng = np.random.default_rng(2)
    sizes = np.array([3, 1, 4, 1, 5])
    off = _offsets(sizes)
    r, v, m = rng.normal(size=(14, 3)), rng.normal(size=(14, 3)), rng.uniform(1, 2, 14)
    rs, vs, ms, os_ = batched.take_universes(r, v, m, off, [4, 0, 2])
    assert os_.tolist() == [0, 5, 8, 12]
    assert np.array_equal(rs[:5], r[9:14]) and np.array_equal(vs[5:8], v[0:3])
    assert np.array_equal(ms[8:12], m[4:8])
    # the shares of a partition hold every body once
    parts = [batched.take_universes(r, v, m, off, b) for b in batched.shard_universes(off, 2)]
    assert sum(p[3][-1] for p in parts) == 14
    assert np.allclose(np.sort(np.concatenate([p[2] for p in parts])), np.sort(m))
    empty = batched.take_universes(r, v, m, off, [])
    assert empty[0].shape == (0, 3) and empty[3].tolist() == [0]
    with pytest.raises(ValueError):
        batched.take_universes(r, v, m, off, [5])
