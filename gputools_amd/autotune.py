"""Tile-shape autotuner for the batched persistent kernel.

For a workload (method, feature counts, distribution of set sizes) every tile
shape the library offers is timed with HIP events on the actual data and the
fastest is kept in a JSON cache keyed by GPU architecture and a workload
signature, so later calls with a similar workload skip the search.

Cache file: $GPUTOOLS_AMD_TUNE_CACHE, else ~/.cache/gputools_amd/autotune.json.
"""
from __future__ import annotations

import json
import os

import numpy as np
import torch

from .batched import Batch, tile_count


def cache_path() -> str:
    p = os.environ.get("GPUTOOLS_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "gputools_amd", "autotune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache: dict) -> None:
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass   # a read-only home only disables caching


def signature(batch: Batch) -> str:
    """Workload key: arch, method, feature counts, set-size quartiles (log2
    buckets) and the number of sets (log2 bucket)."""
    arch = torch.cuda.get_device_properties(batch.device).gcnArchName.split(":")[0]
    live = batch.sizes[batch.sizes >= 2]
    if live.size == 0:
        return "%s|%s|empty" % (arch, batch.method)
    q = np.percentile(live, [25, 50, 75]).astype(np.int64)
    bucket = lambda v: int(np.ceil(np.log2(max(int(v), 1))))  # noqa: E731
    dims = ",".join(str(int(d)) for d in sorted(set(batch.dims[batch.sizes >= 2].tolist())))
    return "%s|%s|d=%s|q=%s|sets=%d" % (arch, batch.method, dims,
                                        "/".join(str(bucket(v)) for v in q),
                                        bucket(live.size))


def time_tile(batch: Batch, tile: int, reps: int = 5) -> float:
    batch.run(tile)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        batch.run(tile)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_batched(batch: Batch, reps: int = 5, use_cache: bool = True, verbose: bool = False):
    """Return the fastest tile shape id for this batch (and cache it)."""
    key = signature(batch)
    cache = _load() if use_cache else {}
    if key in cache:
        return int(cache[key]["tile"])
    times = {t: time_tile(batch, t, reps) for t in range(1, tile_count() + 1)}
    best = min(times, key=times.get)
    if verbose:
        print("autotune %s: %s -> tile %d" % (key, times, best))
    if use_cache:
        cache = _load()
        cache[key] = {"tile": best, "ms": {str(t): round(v, 4) for t, v in times.items()}}
        _store(cache)
    return best
