"""Host-side mirror of gputools' clustering API over the HIP C-ABI.

``gpuHclust(distances, method)`` (R/gpuHclust.R:40-91) and
``gpuDistClust(points, distmethod, clustmethod)`` (R/gpuHclust.R:93-146) keep
the R functions' argument handling: ``pmatch`` of the method names with their
error messages, the checks on the dissimilarities, ``as.single(...)``
conversion, and an ``hclust``-like result (merge, height, order, labels,
method, call, dist.method).  As in the reference, the clustering method
string reaches the C layer unchanged, where names other than the exact
getClusterEnum spellings (src/rinterface.cpp:158-170) select single linkage:
R's "flexible group" therefore clusters with single linkage, as gputools
does.  ``gpuDistClust`` also keeps the reference's handling of "euclidian":
the alias is assigned to an unused variable (R/gpuHclust.R:96-98), so that
spelling is rejected as an invalid distance method.

The device API (:func:`hclust`) runs the clustering stage on a full n x n
distance matrix that already lives in HBM (torch tensor), on torch's current
stream.
"""
from __future__ import annotations

import ctypes
import warnings
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import check, lib
from .dist import Dist, METHODS, _as_points, pmatch

# the clustering methods gpuHclust / gpuDistClust accept (R/gpuHclust.R:43-44)
R_CLUST_METHODS = ("ward", "single", "complete", "average", "mcquitty", "median", "centroid",
                   "flexible", "flexible group", "wpgma")


class HClust:
    """An R ``hclust`` object: merge ((n-1) x 2, 1-based, negative entries
    are singletons), height, order (1-based), labels, method, call and
    dist.method."""

    def __init__(self, merge, height, order, labels=None, method="complete", call=None,
                 dist_method=None):
        self.merge = merge
        self.height = height
        self.order = order
        self.labels = labels
        self.method = method
        self.call = call
        self.dist_method = dist_method

    def __len__(self) -> int:
        return self.order.shape[0]

    def __repr__(self):
        return "HClust(n=%d, method=%r, dist.method=%r)" % (len(self), self.method, self.dist_method)


def match_clust_method(method: str) -> str:
    idx = pmatch(method, R_CLUST_METHODS)
    if idx is None:
        raise ValueError("invalid clustering method")
    return R_CLUST_METHODS[idx]


def _c_strings(*values: str):
    return [(ctypes.c_char_p * 1)(v.encode()) for v in values]


def _tree(merge_flat: np.ndarray, order: np.ndarray, val: np.ndarray, n: int, **kw) -> HClust:
    merge = merge_flat.reshape(2, n - 1).T.copy()   # matrix(a$merge, n-1, 2)
    return HClust(merge, val.astype(np.float64), order, **kw)


def gpuHclust(distances, method: str = "complete") -> HClust:
    """Hierarchical clustering of a ``dist`` object (R/gpuHclust.R:40-91).

    ``distances`` is a :class:`~gputools_amd.dist.Dist` (e.g. from
    :func:`~gputools_amd.dist.gpuDist`).  ``method`` is one of "ward",
    "single", "complete", "average", "mcquitty", "median", "centroid",
    "flexible", "flexible group", "wpgma" (unique prefixes accepted).
    """
    call = "gpuHclust(distances = distances, method = %r)" % (method,)
    method = match_clust_method(method)
    if not isinstance(distances, Dist):
        raise ValueError("invalid dissimilarities")
    n = distances.size
    if n < 2:
        raise ValueError("must have n >= 2 objects to cluster")
    need = n * (n - 1) // 2
    if len(distances) != need:
        if len(distances) < need:
            raise ValueError("dissimilarities of improper length")
        warnings.warn("dissimilarities of improper length")
    d = np.ascontiguousarray(distances.as_matrix(), dtype=np.float32)   # as.single(as.matrix(.))
    merge = np.zeros(2 * (n - 1), np.int32)
    order = np.zeros(n, np.int32)
    val = np.zeros(n - 1, np.float32)
    cn = ctypes.c_int(n)
    fp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    (cm,) = _c_strings(method)
    lib().Rhcluster(d.ctypes.data_as(fp), ctypes.byref(cn), merge.ctypes.data_as(ip),
                    order.ctypes.data_as(ip), val.ctypes.data_as(fp), cm)
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
    return _tree(merge, order, val, n, labels=distances.attrs.get("Labels"), method=method,
                 call=call, dist_method=distances.method)


def gpuDistClust(points, distmethod: str = "euclidean", clustmethod: str = "complete") -> HClust:
    """Distances and clustering in one call (R/gpuHclust.R:93-146): the
    distance matrix stays in HBM between the two stages (p = 2 for
    minkowski, as Rdistclust fixes it)."""
    call = "gpuDistClust(points = points, distmethod = %r, clustmethod = %r)" % (distmethod,
                                                                                 clustmethod)
    idx = pmatch(distmethod, METHODS)
    if idx is None:
        raise ValueError("invalid distance method")
    distmethod = METHODS[idx]
    clustmethod = match_clust_method(clustmethod)
    x, labels = _as_points(points)
    n, k = x.shape
    if n < 1:
        raise ValueError("need at least one point")
    merge = np.zeros(max(2 * (n - 1), 1), np.int32)
    order = np.zeros(n, np.int32)
    val = np.zeros(max(n - 1, 1), np.float32)
    cn, ck = ctypes.c_int(n), ctypes.c_int(k)
    fp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    dm, cm = _c_strings(distmethod, clustmethod)
    lib().Rdistclust(dm, cm, x.ctypes.data_as(fp), ctypes.byref(cn), ctypes.byref(ck),
                     merge.ctypes.data_as(ip), order.ctypes.data_as(ip), val.ctypes.data_as(fp))
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
    return _tree(merge[:2 * (n - 1)], order, val[:n - 1], n, labels=labels, method=clustmethod,
                 call=call, dist_method=distmethod)


def hclust_method_id(method) -> int:
    """getClusterEnum's mapping of a C-level method name (or an id)."""
    if isinstance(method, str):
        return int(lib().gt_hclust_method_from_name(method.encode()))
    return int(method)


HC_STATES = {"auto": 0, "lds": 1, "global": 2}


def hclust(d, method="complete", lam: float = 0.5, beta: float = 0.5, stream=None,
           state: str = "auto"):
    """Clustering stage on a full n x n FP32 device matrix ``d`` (torch
    tensor on the GPU, consumed: it is overwritten).  Returns device int32
    tensors ``sub``, ``sup`` and float32 ``height`` of length n - 1: at step s
    the clusters represented by sub[s] < sup[s] merge at height[s].
    ``state`` places the kernel's per-row state: "auto" (LDS up to 9216
    points, global memory above), "lds" or "global"; every placement gives
    the same result."""
    import torch
    if not isinstance(d, torch.Tensor) or not d.is_cuda or d.dtype != torch.float32:
        raise TypeError("d must be a float32 torch tensor on the GPU")
    if d.dim() != 2 or d.shape[0] != d.shape[1] or d.stride(1) != 1:
        raise ValueError("d must be a square matrix with unit column stride")
    n = d.shape[0]
    m = max(n - 1, 0)
    sub = torch.empty(m, dtype=torch.int32, device=d.device)
    sup = torch.empty(m, dtype=torch.int32, device=d.device)
    height = torch.empty(m, dtype=torch.float32, device=d.device)
    b = ctypes.c_size_t()
    check(lib().gt_hclust_workspace_size(n, ctypes.byref(b)))
    ws = torch.empty(max(b.value, 1), dtype=torch.uint8, device=d.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().gt_hclust_state(d.data_ptr(), n, d.stride(0), hclust_method_id(method),
                                float(lam), float(beta), sub.data_ptr() if m else None,
                                sup.data_ptr() if m else None,
                                height.data_ptr() if m else None, ws.data_ptr(), b.value,
                                HC_STATES[state], s.cuda_stream or None))
    return sub, sup, height


def hclust_format(sub: Sequence[int], sup: Sequence[int]):
    """R's merge matrix ((n-1) x 2) and order (n) from a merge sequence."""
    sub = np.ascontiguousarray(sub, dtype=np.int32)
    sup = np.ascontiguousarray(sup, dtype=np.int32)
    ln = sub.shape[0]
    merge = np.zeros(max(2 * ln, 1), np.int32)
    order = np.zeros(ln + 1, np.int32)
    check(lib().gt_hclust_format(ln, sub.ctypes.data if ln else None,
                                 sup.ctypes.data if ln else None, merge.ctypes.data,
                                 order.ctypes.data))
    return merge[:2 * ln].reshape(2, ln).T.copy(), order
