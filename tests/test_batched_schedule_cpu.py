"""Host-side tile schedule of the batched path (no GPU needed)."""
import ctypes

import numpy as np
import pytest

from gputools_amd import _lib


def schedule(sizes, dims, tile):
    L = _lib.lib()
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    dims = np.ascontiguousarray(dims, dtype=np.int64)
    cnt = ctypes.c_int64()
    assert L.gt_dist_batched_schedule(sizes.ctypes.data, dims.ctypes.data, sizes.size, tile,
                                      None, 0, ctypes.byref(cnt)) == 0
    out = np.empty(max(cnt.value, 1), np.int64)
    assert L.gt_dist_batched_schedule(sizes.ctypes.data, dims.ctypes.data, sizes.size, tile,
                                      out.ctypes.data, out.size, ctypes.byref(cnt)) == 0
    return out[:cnt.value]


@pytest.mark.parametrize("tile,side", [(1, 32), (2, 64), (3, 128)])
def test_schedule_covers_every_lower_tile_once(tile, side):
    rng = np.random.default_rng(7)
    sizes = rng.integers(0, 700, 60)
    sizes[:3] = [0, 1, 2]
    dims = rng.choice([16, 32, 64, 128], 60)
    sched = schedule(sizes, dims, tile)
    s, I, J = sched >> 32, (sched >> 16) & 0xffff, sched & 0xffff
    want = set()
    for si, n in enumerate(sizes):
        if n < 2:
            continue
        T = -(-int(n) // side)
        want |= {(si, a, b) for a in range(T) for b in range(a + 1)}
    got = list(zip(s.tolist(), I.tolist(), J.tolist()))
    assert len(got) == len(set(got)) == len(want)
    assert set(got) == want
    # most expensive (more features, then larger sets) first
    d = dims[s]
    n = sizes[s]
    assert np.all(np.diff(d) <= 0)
    for dv in np.unique(d):
        assert np.all(np.diff(n[d == dv]) <= 0)


def test_schedule_rejects_bad_input():
    L = _lib.lib()
    cnt = ctypes.c_int64()
    sizes = np.array([5, 3], np.int64)
    dims = np.array([4, 0], np.int64)
    assert L.gt_dist_batched_schedule(sizes.ctypes.data, dims.ctypes.data, 2, 2, None, 0,
                                      ctypes.byref(cnt)) == _lib.GT_ERR_INVALID
    dims[1] = 4
    assert L.gt_dist_batched_schedule(sizes.ctypes.data, dims.ctypes.data, 2, 9, None, 0,
                                      ctypes.byref(cnt)) == _lib.GT_ERR_INVALID
    sizes[0] = -1
    assert L.gt_dist_batched_schedule(sizes.ctypes.data, dims.ctypes.data, 2, 2, None, 0,
                                      ctypes.byref(cnt)) == _lib.GT_ERR_INVALID
    assert L.gt_batched_tile_count() == 3


def test_batched_launch_validation_before_the_device():
    L = _lib.lib()
    d = ctypes.c_void_p(16)
    # unsupported method, unknown tile, null pointers: rejected without a launch
    assert L.gt_dist_batched(d, d, d, d, 1, d, d, 3, 2, d, 5, None) == _lib.GT_ERR_UNSUPPORTED
    assert L.gt_dist_batched(d, d, d, d, 1, d, d, 0, 7, d, 5, None) == _lib.GT_ERR_INVALID
    assert L.gt_dist_batched(None, d, d, d, 1, d, d, 0, 2, d, 5, None) == _lib.GT_ERR_INVALID
    assert L.gt_dist_batched(d, d, d, d, 1, d, d, 0, 2, d, 0, None) == _lib.GT_OK
