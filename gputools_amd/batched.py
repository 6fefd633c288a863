"""Many independent point sets in one persistent launch, optionally replayed
from a captured HIP graph.

    b = Batch(list_of_arrays, method="euclidean")
    b.run(tile=2)                 # one launch for every set
    d3 = b.result(3)              # packed 'dist' vector of set 3 (device view)
    g = b.capture(tile=2)         # torch.cuda.CUDAGraph around the same launch
    g.replay()

The tile schedule (every lower-triangular tile of every set, most expensive
first) is built once on the host by gt_dist_batched_schedule and kept on the
device; the launch itself allocates and synchronises nothing.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .dist import _method_id, _stream_ptr

BATCHED_METHODS = ("euclidean", "manhattan", "maximum")


def tile_count() -> int:
    return int(lib().gt_batched_tile_count())


def tile_side(tile: int) -> int:
    return {1: 32, 2: 64, 3: 128}[tile]


class Batch:
    """Point sets laid out back to back in one device buffer."""

    def __init__(self, sets=None, *, method="euclidean", x=None, sizes=None, dims=None,
                 device="cuda"):
        if method not in BATCHED_METHODS:
            raise ValueError("batched path supports %s" % (BATCHED_METHODS,))
        self.method = method
        self.device = torch.device(device)
        if sets is not None:
            arrs = [np.ascontiguousarray(np.asarray(s, dtype=np.float32)) for s in sets]
            if any(a.ndim != 2 for a in arrs):
                raise ValueError("every set must be a 2-D array")
            self.sizes = np.array([a.shape[0] for a in arrs], dtype=np.int64)
            self.dims = np.array([a.shape[1] for a in arrs], dtype=np.int64)
            flat = np.concatenate([a.ravel() for a in arrs]) if arrs else np.zeros(0, np.float32)
            self.x = torch.from_numpy(flat).to(self.device)
        else:
            self.sizes = np.asarray(sizes, dtype=np.int64)
            self.dims = np.asarray(dims, dtype=np.int64)
            if x.dtype != torch.float32 or not x.is_cuda:
                raise TypeError("x must be a float32 device tensor")
            self.x = x.reshape(-1)
        elems = self.sizes * self.dims
        self.x_off = np.concatenate([[0], np.cumsum(elems)[:-1]]).astype(np.int64)
        if self.x.numel() < int(elems.sum()):
            raise ValueError("x holds fewer values than the sets describe")
        pairs = self.sizes * (self.sizes - 1) // 2
        self.out_off = np.concatenate([[0], np.cumsum(pairs)[:-1]]).astype(np.int64)
        self.pairs = int(pairs.sum())
        self.out = torch.empty(max(self.pairs, 1), dtype=torch.float32, device=self.device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)  # noqa: E731
        self.d_sizes, self.d_dims = t(self.sizes), t(self.dims)
        self.d_xoff, self.d_outoff = t(self.x_off), t(self.out_off)
        self._schedules = {}

    @property
    def nsets(self) -> int:
        return int(self.sizes.shape[0])

    def schedule(self, tile: int):
        if tile not in self._schedules:
            cnt = ctypes.c_int64(0)
            sz = np.ascontiguousarray(self.sizes)
            dm = np.ascontiguousarray(self.dims)
            check(lib().gt_dist_batched_schedule(sz.ctypes.data, dm.ctypes.data, self.nsets, tile,
                                                 None, 0, ctypes.byref(cnt)))
            host = np.empty(max(cnt.value, 1), dtype=np.int64)
            check(lib().gt_dist_batched_schedule(sz.ctypes.data, dm.ctypes.data, self.nsets, tile,
                                                 host.ctypes.data, host.shape[0], ctypes.byref(cnt)))
            self._schedules[tile] = (torch.from_numpy(host).to(self.device), cnt.value)
        return self._schedules[tile]

    def run(self, tile: int = 2, stream=None) -> None:
        sched, count = self.schedule(tile)
        check(lib().gt_dist_batched(self.x.data_ptr(), self.d_xoff.data_ptr(),
                                    self.d_sizes.data_ptr(), self.d_dims.data_ptr(), self.nsets,
                                    self.out.data_ptr(), self.d_outoff.data_ptr(),
                                    _method_id(self.method), tile, sched.data_ptr(), count,
                                    _stream_ptr(stream)))

    def result(self, s: int):
        o = int(self.out_off[s])
        n = int(self.sizes[s])
        return self.out[o:o + n * (n - 1) // 2]

    def set_points(self, s: int):
        o = int(self.x_off[s])
        return self.x[o:o + int(self.sizes[s] * self.dims[s])].view(int(self.sizes[s]),
                                                                      int(self.dims[s]))

    def capture(self, tile: int = 2):
        """Capture one batched launch in a HIP graph (torch.cuda.CUDAGraph)."""
        self.schedule(tile)            # host work and uploads stay outside the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            self.run(tile, stream=s)   # warm-up on the capture stream
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            self.run(tile, stream=s)
        torch.cuda.synchronize()
        return g

    def run_per_set(self) -> None:
        """Reference strategy: one gt_dist launch per set (no batching)."""
        m = _method_id(self.method)
        sp = _stream_ptr()
        for s in range(self.nsets):
            n, d = int(self.sizes[s]), int(self.dims[s])
            if n < 2:
                continue
            check(lib().gt_dist(self.x.data_ptr() + 4 * int(self.x_off[s]), n, d, d, m, 2.0,
                                _lib.PREC_FP32, _lib.LAYOUT_PACKED,
                                self.out.data_ptr() + 4 * int(self.out_off[s]), n, sp))


def dist_batched(sets, method="euclidean", tile=None):
    """Packed distances of every set (list of device views into one buffer)."""
    b = Batch(sets, method=method)
    if tile is None:
        from .autotune import tune_batched
        tile = tune_batched(b)
    b.run(tile)
    return [b.result(s) for s in range(b.nsets)]
