"""Host logic of the MPM scatter windows (no GPU): the cuts of the particle
order that P2G and the G2P adjoint run one workgroup per window over."""

import numpy as np
import pytest

from diffloco import mpm as dm


def _box(base):
    return int(np.prod(base.max(0) - base.min(0) + 3))


def _check_partition(win, P, pb):
    assert win.dtype == np.int32 and win.shape[1] == 2
    assert win[0, 0] == 0 and win[-1, 1] == P
    assert np.all(win[1:, 0] == win[:-1, 1])
    sizes = win[:, 1] - win[:, 0]
    assert sizes.min() >= 1 and sizes.max() <= pb


@pytest.mark.parametrize("limit", [27, 200, 512, 4096])
def test_windows_partition_bound_and_maximal(limit):
    """Two clusters far apart, block-sorted, then a shuffled tail: every
    window's stencil box stays within the limit, and every window is
    maximal (full, last, or the next particle would break the limit)."""
    rng = np.random.default_rng(3)
    n = 64
    a = rng.uniform([0.1, 0.1, 0.1], [0.2, 0.15, 0.2], (900, 3))
    b = rng.uniform([0.6, 0.1, 0.6], [0.7, 0.15, 0.7], (700, 3))
    pos = np.concatenate([a, b])
    pos = pos[dm.block_order(pos, n)]
    pos = np.concatenate([pos, rng.permutation(pos)[:300]])
    win = dm.scatter_windows(pos, n, limit)
    _check_partition(win, len(pos), dm.MPM_PB)
    base = np.floor(pos * n - 0.5).astype(np.int64)
    for s, e in win:
        assert _box(base[s:e]) <= limit
        if e - s < dm.MPM_PB and e < len(pos):
            assert _box(base[s:e + 1]) > limit


def test_windows_without_a_binding_limit_are_plain_chunks():
    rng = np.random.default_rng(4)
    pos = rng.uniform(0.2, 0.3, (1000, 3))
    win = dm.scatter_windows(pos, 32, 10 ** 9)
    P, pb = len(pos), dm.MPM_PB
    exp = np.array([(s, min(s + pb, P)) for s in range(0, P, pb)])
    assert np.array_equal(win, exp)


def test_windows_of_the_quadruped_stay_few():
    """The 64^3 quadruped (27,648 particles): a 512-node box limit costs a
    handful of extra workgroups over the 108 plain chunks."""
    pos, ids, n_act = dm.quadruped3d_particles(cells_per_cube=6, n_grid=64)
    pos = np.asarray(pos)[dm.block_order(np.asarray(pos), 64)]
    win = dm.scatter_windows(pos, 64, 512)
    _check_partition(win, len(pos), dm.MPM_PB)
    assert 108 <= len(win) <= 130
