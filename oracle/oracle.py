"""ctypes wrapper of the CPU oracle (oracle/dist_oracle.c, oracle/hclust_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of ``bench.py``.  The product package never imports it.

The oracle restates R ``stats::dist`` in FP64 for the six methods of gputools'
``gpuDist`` (R/gpuHclust.R:1-39, kernels inst/cuda/distance.cu:40-493); see the
header of dist_oracle.c for the per-method citations and where R and gputools
differ.  It also restates the clustering stage of gpuDistClust / gpuHclust
(src/hcluster.cpp:40-201, inst/cuda/hcluster.cu, src/rinterface.cpp:241-292)
in the reference's own FP32 arithmetic; see hclust_oracle.c.  And it restates
gpuCor's Pearson correlation (src/correlation.cpp:51-160,
inst/cuda/correlation.cu) in FP64 with the t-scores in FP32, and its Kendall
tau (inst/cuda/kendall.cu); see cor_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

METHODS = ("euclidean", "maximum", "manhattan", "canberra", "binary", "minkowski")
# hc_method order (src/hcluster.h:5-16)
HC_METHODS = ("single", "complete", "wpgma", "average", "median", "centroid", "flexible_group",
              "flexible", "ward", "mcquitty")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64


def build() -> None:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_packed_index.argtypes = [_i64, _i64, _i64]
        L.oracle_packed_index.restype = _i64
        L.oracle_dist_packed.argtypes = [_f32p, _i64, _i64, _i64, ctypes.c_int,
                                         ctypes.c_double, _f64p, ctypes.c_int]
        L.oracle_dist_packed.restype = ctypes.c_int
        L.oracle_dist_pairs.argtypes = [_f32p, _i64, _i64, _i64p, _i64p, _i64,
                                        ctypes.c_int, ctypes.c_double, _f64p, ctypes.c_int]
        L.oracle_dist_pairs.restype = ctypes.c_int
        L.oracle_dist_cross.argtypes = [_f32p, _i64, _f32p, _i64, _i64, ctypes.c_int,
                                        ctypes.c_double, _f64p, ctypes.c_int]
        L.oracle_dist_cross.restype = ctypes.c_int
        L.oracle_compare_packed.argtypes = [_f32p, _i64, _i64, _i64, ctypes.c_int,
                                            ctypes.c_double, _f32p, ctypes.c_double,
                                            ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                            ctypes.c_int]
        L.oracle_compare_packed.restype = ctypes.c_double
        L.oracle_dist_backward.argtypes = [_f32p, _i64, _i64, ctypes.c_int, ctypes.c_double,
                                           _f64p, _f64p, ctypes.c_int]
        L.oracle_dist_backward.restype = ctypes.c_int
        L.oracle_mds_stress.argtypes = [_f32p, _i64, _i64, _f64p, ctypes.c_void_p,
                                        ctypes.c_int]
        L.oracle_mds_stress.restype = ctypes.c_double
        _i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.oracle_hclust.argtypes = [_f32p, _i64, _i64, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, _i32p, _i32p, _f32p]
        L.oracle_hclust.restype = ctypes.c_int
        L.oracle_hclust_format.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p]
        L.oracle_hclust_format.restype = ctypes.c_int
        L.oracle_pmcc.argtypes = [ctypes.c_int, _f32p, _i64, _f32p, _i64, _i64, _f64p, _f64p,
                                  ctypes.c_int]
        L.oracle_pmcc.restype = ctypes.c_int
        L.oracle_signif.argtypes = [_f32p, _f32p, _i64, ctypes.c_int, _f32p]
        L.oracle_signif.restype = None
        L.oracle_kendall.argtypes = [_f32p, _i64, _f32p, _i64, _i64, _f64p, ctypes.c_void_p,
                                     ctypes.c_int]
        L.oracle_kendall.restype = ctypes.c_int
        _lib = L
    return _lib


def method_id(method) -> int:
    if isinstance(method, str):
        return METHODS.index(method)
    return int(method)


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def dist(x, method="euclidean", p=2.0, nthreads=0) -> np.ndarray:
    """Packed FP64 distances of the rows of x, in R 'dist' order."""
    x = _f32(x)
    if x.ndim != 2:
        raise ValueError("x must be 2-D")
    n, k = x.shape
    out = np.empty(n * (n - 1) // 2, dtype=np.float64)
    rc = lib().oracle_dist_packed(x, n, k, k, method_id(method), float(p), out, nthreads)
    if rc != 0:
        raise ValueError("oracle_dist_packed rejected its arguments")
    return out


def dist_pairs(x, ii, jj, method="euclidean", p=2.0, nthreads=0) -> np.ndarray:
    """FP64 distance of the listed row pairs (ii[q], jj[q])."""
    x = _f32(x)
    n, k = x.shape
    ii = np.ascontiguousarray(ii, dtype=np.int64)
    jj = np.ascontiguousarray(jj, dtype=np.int64)
    out = np.empty(ii.shape[0], dtype=np.float64)
    rc = lib().oracle_dist_pairs(x, k, k, ii, jj, ii.shape[0], method_id(method), float(p),
                                 out, nthreads)
    if rc != 0:
        raise ValueError("oracle_dist_pairs rejected its arguments")
    return out


def dist_cross(a, b, method="euclidean", p=2.0, nthreads=0) -> np.ndarray:
    """FP64 cross distances, nb x na row-major (out[y, x] = d(a[x], b[y]))."""
    a, b = _f32(a), _f32(b)
    na, k = a.shape
    nb = b.shape[0]
    out = np.empty(na * nb, dtype=np.float64)
    rc = lib().oracle_dist_cross(a, na, b, nb, k, method_id(method), float(p), out, nthreads)
    if rc != 0:
        raise ValueError("oracle_dist_cross rejected its arguments")
    return out.reshape(nb, na)


def compare_packed(x, got, method="euclidean", p=2.0, abs_floor=1e-30, nthreads=0):
    """Max relative error of an FP32 packed result vs the oracle (streamed)."""
    x = _f32(x)
    n, k = x.shape
    got = np.ascontiguousarray(got, dtype=np.float32)
    wi, wj = _i64(-1), _i64(-1)
    err = lib().oracle_compare_packed(x, n, k, k, method_id(method), float(p), got,
                                      float(abs_floor), ctypes.byref(wi), ctypes.byref(wj),
                                      nthreads)
    return err, (wi.value, wj.value)


def dist_backward(x, g_packed, method="euclidean", p=2.0, nthreads=0) -> np.ndarray:
    x = _f32(x)
    n, k = x.shape
    g = np.ascontiguousarray(g_packed, dtype=np.float64)
    dx = np.empty((n, k), dtype=np.float64)
    rc = lib().oracle_dist_backward(x, n, k, method_id(method), float(p), g, dx, nthreads)
    if rc != 0:
        raise ValueError("oracle_dist_backward rejected its arguments")
    return dx


def mds_stress(y, delta, with_grad=True, nthreads=0):
    y = _f32(y)
    n, q = y.shape
    delta = np.ascontiguousarray(delta, dtype=np.float64)
    dy = np.empty((n, q), dtype=np.float64) if with_grad else None
    s = lib().oracle_mds_stress(y, n, q, delta,
                                dy.ctypes.data if dy is not None else None, nthreads)
    return (s, dy) if with_grad else s


def packed_index(n, i, j) -> int:
    return int(lib().oracle_packed_index(n, i, j))


def hclust(d_full, method="complete", lam=0.5, beta=0.5):
    """The reference's merge loop on a full n x n FP32 distance matrix (a copy
    is consumed).  Returns (sub, sup, height): at step s clusters represented
    by sub[s] < sup[s] merge at height[s] (0-based point indices)."""
    d = np.array(d_full, dtype=np.float32, order="C", copy=True)
    n = d.shape[0]
    if d.ndim != 2 or d.shape[1] != n:
        raise ValueError("d_full must be square")
    m = HC_METHODS.index(method) if isinstance(method, str) else int(method)
    sub = np.zeros(max(n - 1, 0), np.int32)
    sup = np.zeros(max(n - 1, 0), np.int32)
    h = np.zeros(max(n - 1, 0), np.float32)
    if lib().oracle_hclust(d, n, n, m, lam, beta, sub, sup, h) != 0:
        raise ValueError("oracle_hclust rejected its arguments")
    return sub, sup, h


def hclust_format(sub, sup):
    """R's hclust merge matrix ((n-1) x 2, 1-based, negative = singleton) and
    order (1-based), as formatClustering / getPrintOrder build them."""
    sub = np.ascontiguousarray(sub, dtype=np.int32)
    sup = np.ascontiguousarray(sup, dtype=np.int32)
    ln = sub.shape[0]
    merge = np.zeros(2 * ln, np.int32)
    order = np.zeros(ln + 1, np.int32)
    if lib().oracle_hclust_format(ln, sub, sup, merge, order) != 0:
        raise ValueError("oracle_hclust_format rejected its arguments")
    return merge.reshape(2, ln).T.copy(), order


USES = ("everything", "pairwise.complete.obs")


def pmcc(a, b=None, use="everything", nthreads=0):
    """Pearson correlation of every variable (row) of ``a`` with every row of
    ``b`` (default ``a``), as gputools' pmcc computes it, in FP64.  ``a`` is
    na x dim (row x = variable x: a column of R's matrix).  Returns
    (pairs, corr), both na x nb FP64."""
    a = _f32(a)
    b = a if b is None else _f32(b)
    if a.ndim != 2 or b.ndim != 2 or a.shape[1] != b.shape[1]:
        raise ValueError("a and b must be 2-D with the same number of samples")
    u = USES.index(use) if isinstance(use, str) else int(use)
    na, dim = a.shape
    nb = b.shape[0]
    pairs = np.empty(na * nb, dtype=np.float64)
    corr = np.empty(na * nb, dtype=np.float64)
    if lib().oracle_pmcc(u, a, na, b, nb, dim, pairs, corr, nthreads) != 0:
        raise ValueError("oracle_pmcc rejected its arguments")
    return pairs.reshape(na, nb), corr.reshape(na, nb)


def signif(pairs, corr, cap=True) -> np.ndarray:
    """FP32 t-scores: hostSignif (cap, what gpuCor returns) or gpuSignif (no
    cap, what gpuTtest returns)."""
    p = _f32(pairs).ravel()
    c = _f32(corr).ravel()
    if p.shape != c.shape:
        raise ValueError("pairs and corr differ in size")
    ts = np.empty_like(c)
    lib().oracle_signif(p, c, c.shape[0], 1 if cap else 0, ts)
    return ts.reshape(np.shape(corr))


def kendall(a, b=None, nthreads=0):
    """Kendall's tau as gputools' gpuKendall defines it, for every row
    (variable) of ``a`` against every row of ``b`` (default ``a``).  Returns
    (tau, count): nb x na arrays (entry (y, x)), tau FP64, count the exact
    number of concordant sample pairs (uint64)."""
    a = _f32(a)
    b = a if b is None else _f32(b)
    if a.ndim != 2 or b.ndim != 2 or a.shape[1] != b.shape[1]:
        raise ValueError("a and b must be 2-D with the same number of samples")
    na, n = a.shape
    nb = b.shape[0]
    tau = np.empty(na * nb, dtype=np.float64)
    count = np.empty(na * nb, dtype=np.uint64)
    if lib().oracle_kendall(a, na, b, nb, n, tau, count.ctypes.data, nthreads) != 0:
        raise ValueError("oracle_kendall rejected its arguments")
    return tau.reshape(nb, na), count.reshape(nb, na)
