"""Host-side mirror of gputools' distance API over the HIP C-ABI.

``gpuDist(points, method, p)`` keeps the argument handling of the R function
(R/gpuHclust.R:1-39): the "euclidian" alias, ``pmatch`` method matching with
its error messages, ``as.single(t(points))`` conversion to row-major FP32, the
refusal of NA/NaN/Inf input that ``.C`` applies, and a ``dist``-like result
(packed lower triangle in R's order, values widened to double as ``.C`` does,
attributes Size, Labels, Diag, Upper, method, call and -- minkowski only -- p).

Unlike the reference, the packed triangle is produced on the device and
copied back directly (n(n-1)/2 values instead of the n x n matrix and
``as.dist``), which also lifts the n <= 46340 limit of ``single(n * n)``
(R/gpuHclust.R:25).

The device API (:func:`dist`, :func:`dist_cross`) works on torch tensors that
already live in HBM and runs on torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib

METHODS = _lib.METHOD_NAMES


def pmatch(x: str, table: Sequence[str]) -> Optional[int]:
    """R's ``pmatch(x, table)`` for one string: exact match first, then a
    unique partial (prefix) match; the empty string and ambiguous prefixes
    match nothing (None stands for NA)."""
    if not isinstance(x, str) or x == "":
        return None
    for idx, name in enumerate(table):
        if x == name:
            return idx
    hits = [idx for idx, name in enumerate(table) if name.startswith(x)]
    return hits[0] if len(hits) == 1 else None


def match_method(method: str) -> str:
    """Method resolution of gpuDist (R/gpuHclust.R:3-16)."""
    if pmatch(method, ["euclidian"]) is not None:
        method = "euclidean"
    idx = pmatch(method, METHODS)
    if idx is None:
        raise ValueError("invalid distance method")
    return METHODS[idx]


class Dist:
    """An R ``dist`` object: the strict lower triangle in column order."""

    def __init__(self, values: np.ndarray, size: int, labels=None, method: str = "euclidean",
                 call: Optional[str] = None, p: Optional[float] = None):
        self.values = values
        self.attrs = {"Size": int(size), "Labels": labels, "Diag": False, "Upper": False,
                      "method": method, "call": call}
        if p is not None:
            self.attrs["p"] = p

    @property
    def size(self) -> int:
        return self.attrs["Size"]

    @property
    def method(self) -> str:
        return self.attrs["method"]

    def __len__(self) -> int:
        return self.values.shape[0]

    def index(self, i: int, j: int) -> int:
        """Position of pair (i, j), i != j, 0-based."""
        n = self.size
        if i == j or not (0 <= i < n and 0 <= j < n):
            raise IndexError("pair (%d, %d) out of range" % (i, j))
        if i < j:
            i, j = j, i
        return j * n - j * (j + 1) // 2 + (i - j - 1)

    def __getitem__(self, ij):
        i, j = ij
        return 0.0 if i == j else float(self.values[self.index(i, j)])

    def as_matrix(self) -> np.ndarray:
        """``as.matrix(d)``: the symmetric n x n matrix with a zero diagonal."""
        n = self.size
        m = np.zeros((n, n), dtype=self.values.dtype)
        jj, ii = np.triu_indices(n, 1)   # column j < row i, in dist order
        m[ii, jj] = self.values
        m[jj, ii] = self.values
        return m

    def __repr__(self):
        return "Dist(Size=%d, method=%r, len=%d)" % (self.size, self.method, len(self))


def _as_points(points) -> tuple:
    """``as.matrix`` + ``as.single(t(points))``: row-major FP32, plus labels."""
    labels = None
    try:
        import torch
        if isinstance(points, torch.Tensor):
            points = points.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    if hasattr(points, "index") and hasattr(points, "to_numpy"):  # pandas DataFrame
        labels = [str(v) for v in points.index]
        points = points.to_numpy()
    a = np.asarray(points)
    if a.ndim == 1:
        a = a.reshape(-1, 1)     # as.matrix of a vector: one column
    if a.ndim != 2:
        raise ValueError("points must be a matrix")
    if a.dtype == bool:
        a = a.astype(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        x = np.ascontiguousarray(a, dtype=np.float32)
    if not np.all(np.isfinite(x)):
        raise ValueError("NA/NaN/Inf in foreign function call (arg 1)")
    return x, labels


def gpuDist(points, method: str = "euclidean", p: float = 2.0, precision: str = "fp32") -> Dist:
    """Distances between the rows of ``points`` (R/gpuHclust.R:1-39).

    ``method`` is one of "euclidean", "maximum", "manhattan", "canberra",
    "binary", "minkowski" (unique prefixes accepted).  ``p`` is the Minkowski
    power.  ``precision`` (not in the reference, whose gpuDist has one FP32
    path) selects "fp32" or, for euclidean, the split-precision Gram form
    "bf16x3" / "fp16x3" (max relative error 1e-5).  Returns a :class:`Dist`
    whose values are the FP32 GPU results widened to double.
    """
    call = "gpuDist(points = points, method = %r, p = %r)" % (method, p)
    if precision not in _lib.PRECISIONS:
        raise ValueError("precision must be one of %s" % ", ".join(_lib.PRECISIONS))
    method = match_method(method)
    x, labels = _as_points(points)
    n, k = x.shape
    p32 = float(np.float32(p))
    out = np.empty(max(n * (n - 1) // 2, 0), dtype=np.float32)
    m = METHODS.index(method)
    check(lib().gt_dist_host(x.ctypes.data, n, k, m, p32, _lib.PRECISIONS[precision],
                             _lib.LAYOUT_PACKED, out.ctypes.data if out.size else None))
    return Dist(out.astype(np.float64), n, labels=labels, method=method, call=call,
                p=p if method == "minkowski" else None)


def Rdistances(points: np.ndarray, method: str, p: float) -> np.ndarray:
    """Call the .C export exactly as R does (R/gpuHclust.R:21-27): host
    pointers to copies of every argument, a caller-allocated n*n output."""
    x = np.ascontiguousarray(points, dtype=np.float32)
    n, k = x.shape
    d = np.zeros(n * n, dtype=np.float32)
    cn, ck = ctypes.c_int(n), ctypes.c_int(k)
    cm = (ctypes.c_char_p * 1)(method.encode())
    cp = ctypes.c_float(p)
    fp = ctypes.POINTER(ctypes.c_float)
    lib().Rdistances(x.ctypes.data_as(fp), ctypes.byref(cn), ctypes.byref(ck),
                     d.ctypes.data_as(fp), cm, ctypes.byref(cp))
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
    return d.reshape(n, n)


def Rdistances_packed(points: np.ndarray, method: str, p: float) -> np.ndarray:
    """The packed .C export, called as an R shim would: host pointers, a
    caller-allocated n(n-1)/2 output in R 'dist' order."""
    x = np.ascontiguousarray(points, dtype=np.float32)
    n, k = x.shape
    d = np.zeros(max(n * (n - 1) // 2, 1), dtype=np.float32)
    cn, ck = ctypes.c_int(n), ctypes.c_int(k)
    cm = (ctypes.c_char_p * 1)(method.encode())
    cp = ctypes.c_float(p)
    fp = ctypes.POINTER(ctypes.c_float)
    lib().Rdistances_packed(x.ctypes.data_as(fp), ctypes.byref(cn), ctypes.byref(ck),
                            d.ctypes.data_as(fp), cm, ctypes.byref(cp))
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
    return d[:n * (n - 1) // 2]


def Rdistances_packed_prec(points: np.ndarray, method: str, p: float, precision: str) -> np.ndarray:
    """The packed .C export with a precision string, called as an R shim
    would (every argument a pointer, strings as char**)."""
    x = np.ascontiguousarray(points, dtype=np.float32)
    n, k = x.shape
    d = np.zeros(max(n * (n - 1) // 2, 1), dtype=np.float32)
    cn, ck = ctypes.c_int(n), ctypes.c_int(k)
    cm = (ctypes.c_char_p * 1)(method.encode())
    cprec = (ctypes.c_char_p * 1)(precision.encode())
    cp = ctypes.c_float(p)
    fp = ctypes.POINTER(ctypes.c_float)
    lib().Rdistances_packed_prec(x.ctypes.data_as(fp), ctypes.byref(cn), ctypes.byref(ck),
                                 d.ctypes.data_as(fp), cm, ctypes.byref(cp), cprec)
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
    return d[:n * (n - 1) // 2]


# --------------------------------------------------------------------------
# Device API (torch tensors in HBM)
# --------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _stream_ptr(stream=None) -> Optional[int]:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream or None


def _method_id(method) -> int:
    if isinstance(method, int):
        return method
    return METHODS.index(match_method(method))


def _check_points(x, name="x"):
    torch = _torch()
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError("%s must be a torch tensor on the GPU" % name)
    if x.dtype != torch.float32 or x.dim() != 2:
        raise TypeError("%s must be a 2-D float32 tensor" % name)
    if x.stride(1) != 1:
        raise ValueError("%s must have unit stride along features" % name)


def workspace_size(n: int, k: int, method="euclidean", precision="fp32") -> int:
    b = ctypes.c_size_t(0)
    check(lib().gt_dist_workspace_size(n, k, _method_id(method), _lib.PRECISIONS[precision],
                                       ctypes.byref(b)))
    return b.value


def dist(x, method="euclidean", p: float = 2.0, precision: str = "fp32",
         layout: str = "packed", out=None, stream=None):
    """All-pairs distances of the rows of a device tensor ``x`` (n x k FP32).

    layout "packed": n(n-1)/2 values in R 'dist' order; "full": n x n
    symmetric with zero diagonal.  precision "fp32" (any method) or
    "bf16x3" / "fp16x3" (euclidean, split-precision MFMA Gram form).
    """
    torch = _torch()
    _check_points(x)
    n, k = x.shape
    m = _method_id(method)
    prec = _lib.PRECISIONS[precision]
    lay = {"packed": _lib.LAYOUT_PACKED, "full": _lib.LAYOUT_FULL}[layout]
    if out is None:
        if lay == _lib.LAYOUT_PACKED:
            out = torch.empty(max(n * (n - 1) // 2, 0), dtype=torch.float32, device=x.device)
        else:
            out = torch.empty((n, n), dtype=torch.float32, device=x.device)
    ldd = out.stride(0) if lay == _lib.LAYOUT_FULL else n
    ws_bytes = workspace_size(n, k, m, precision)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=x.device) if ws_bytes else None
    check(lib().gt_dist_ws(x.data_ptr() if x.numel() else None, n, k, x.stride(0) if n else k,
                           m, float(p), prec, lay, out.data_ptr() if out.numel() else None, ldd,
                           ws.data_ptr() if ws is not None else None, ws_bytes,
                           _stream_ptr(stream)))
    return out


def dist_cross(a, b, method="euclidean", p: float = 2.0, out=None, stream=None):
    """Cross distances: out[y, x] = d(a[x], b[y]) (nb x na)."""
    torch = _torch()
    _check_points(a, "a")
    _check_points(b, "b")
    if a.shape[1] != b.shape[1]:
        raise ValueError("a and b must have the same number of features")
    na, k = a.shape
    nb = b.shape[0]
    if out is None:
        out = torch.empty((nb, na), dtype=torch.float32, device=a.device)
    check(lib().gt_dist_cross(a.data_ptr(), na, a.stride(0), b.data_ptr(), nb, b.stride(0), k,
                              _method_id(method), float(p), out.data_ptr(), out.stride(0),
                              _stream_ptr(stream)))
    return out


def dist_device_full(points, method="euclidean", p: float = 2.0):
    """distanceLeaveOnGpu (src/distance.cpp:318-344): host points in, a
    device-resident pitched n x n matrix out.  Returns (device pointer,
    pitch in bytes); free it with :func:`free_device`."""
    x, _ = _as_points(points)
    n, k = x.shape
    ptr = ctypes.c_void_p()
    pitch = ctypes.c_size_t()
    check(lib().gt_dist_device_full(x.ctypes.data, n, k, _method_id(method), float(p),
                                    ctypes.byref(ptr), ctypes.byref(pitch)))
    return ptr.value, pitch.value


def free_device(ptr) -> None:
    check(lib().gt_free(ptr))
