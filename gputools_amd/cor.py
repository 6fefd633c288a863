"""Host-side mirror of gputools' correlation API over the HIP C-ABI.

``gpuCor(x, y, use, method)`` (R/gpuCor.R:1-66) and ``gpuTtest(goodPairs,
coeffs)`` (R/gpuCor.R:68-77) keep the R functions' argument handling:
``as.matrix`` (a vector is one variable), the columns of x and y are the
variables and the rows the samples, ``as.single`` conversion, and the result
list (``coefficients``, ``ts`` and ``pairs`` as nx x ny matrices, values
widened to double as ``.C`` returns them).  The reference matches ``method``
and ``use`` with ``pmatch(..., -1)``, whose third argument is ``nomatch``: an
unknown or ambiguous name yields -1, so every failed match stops with
"ambiguous correlation method" (R/gpuCor.R:15-31); that is kept.  Pearson
is called with NAOK = TRUE (NaN is data), Kendall without it (R refuses
NA/NaN/Inf before the call).  Unlike the
reference, y must have as many rows as x (the reference reads past the end
of y otherwise).

The device API (:func:`pmcc`, :func:`kendall`, :func:`ttest`) works on torch tensors that
already live in HBM (variables as rows) and runs on torch's current stream.
"""
from __future__ import annotations

import ctypes
import warnings
import numpy as np

from . import _lib
from ._lib import check, lib
from .dist import pmatch

COR_METHODS = ("pearson", "kendall")
USES = ("everything", "pairwise.complete.obs")


def _match(x, table) -> str:
    idx = pmatch(x, table)
    if idx is None:
        raise ValueError("ambiguous correlation method")
    return table[idx]


def _as_matrix(x) -> np.ndarray:
    """R's as.matrix for numeric input: a vector becomes one column."""
    a = np.asarray(x)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.ndim != 2:
        raise ValueError("x must be a vector or a matrix")
    return a


def _variables(m: np.ndarray) -> np.ndarray:
    """as.single of an R matrix: column-major FP32, i.e. one variable per row."""
    return np.ascontiguousarray(m.T, dtype=np.float32)


def gpuCor(x, y=None, use: str = "everything", method: str = "pearson") -> dict:
    """Correlation of every column of ``x`` with every column of ``y``
    (default ``x``), R/gpuCor.R:1-66.

    method "pearson": returns ``{"coefficients", "ts", "pairs"}``, nx x ny
    float64 matrices (correlation, t-score, samples used).  use
    "pairwise.complete.obs" skips, per pair of columns, the rows where either
    value is NaN.  method "kendall": ``{"coefficients", "pairs"}``.
    """
    xm = _as_matrix(x)
    nx, size = xm.shape[1], xm.shape[0]
    ym = xm if y is None else _as_matrix(y)
    ny = ym.shape[1]
    if ym.shape[0] != size:
        raise ValueError("x and y must have the same number of rows")
    method = _match(method, COR_METHODS)
    use = _match(use, USES)
    a = _variables(xm)
    b = a if y is None else _variables(ym)
    n = nx * ny
    if method == "pearson":
        pairs = np.zeros(max(n, 1), np.float32)
        corr = np.zeros(max(n, 1), np.float32)
        ts = np.zeros(max(n, 1), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        cu = ctypes.c_int(USES.index(use))
        cnx, cny, csz = ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(size)
        lib().rpmcc(ctypes.byref(cu), a.ctypes.data_as(fp), ctypes.byref(cnx),
                    b.ctypes.data_as(fp), ctypes.byref(cny), ctypes.byref(csz),
                    pairs.ctypes.data_as(fp), corr.ctypes.data_as(fp), ts.ctypes.data_as(fp))
        err = _lib.last_error()
        if err:
            raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
        # t(matrix(answer$corr, ny, nx)): entry (x, y) at x * ny + y
        shape = (nx, ny)
        return {"coefficients": corr[:n].reshape(shape).astype(np.float64),
                "ts": ts[:n].reshape(shape).astype(np.float64),
                "pairs": pairs[:n].reshape(shape).astype(np.float64)}
    # kendall (R/gpuCor.R:46-59)
    if use != "everything":
        warnings.warn("NA handling for Kendall's is not yet supported. Defaulting to using "
                      "everything. Sorry for any inconvenience.")
    # .C("RgpuKendall", ...) runs without NAOK = TRUE, so R itself refuses
    # NA/NaN/Inf in x (argument 1) and y (argument 3)
    for arr, arg in ((a, 1), (b, 3)):
        if not np.all(np.isfinite(arr)):
            raise ValueError("NA/NaN/Inf in foreign function call (arg %d)" % arg)
    result = np.zeros(max(n, 1), np.float64)
    fp = ctypes.POINTER(ctypes.c_float)
    cnx, cny, csz = ctypes.c_int(nx), ctypes.c_int(ny), ctypes.c_int(size)
    lib().RgpuKendall(a.ctypes.data_as(fp), ctypes.byref(cnx), b.ctypes.data_as(fp),
                      ctypes.byref(cny), ctypes.byref(csz),
                      result.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
    # matrix(a$result, nx, ny): column-major, entry (x, y) at y * nx + x
    return {"coefficients": result[:n].reshape(ny, nx).T.copy(),
            "pairs": np.full((nx, ny), float(size))}


def gpuTtest(goodPairs, coeffs) -> np.ndarray:
    """t-scores coeffs * sqrt((goodPairs - 2) / (1 - coeffs^2)) on the GPU
    (R/gpuCor.R:68-77; no cap at 0.999, unlike gpuCor's ts)."""
    p = np.ascontiguousarray(np.asarray(goodPairs, dtype=np.float32).ravel())
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float32).ravel())
    n = p.shape[0]
    if c.shape[0] < n:
        raise ValueError("coeffs is shorter than goodPairs")
    out = np.zeros(max(n, 1), np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    cn = ctypes.c_int(n)
    lib().rtestT(p.ctypes.data_as(fp), c.ctypes.data_as(fp), ctypes.byref(cn),
                 out.ctypes.data_as(fp))
    err = _lib.last_error()
    if err:
        raise _lib.GputoolsError(_lib.GT_ERR_HIP, err)
    return out[:n].astype(np.float64)


# ------------------------------- device API --------------------------------

def _use_id(use) -> int:
    if isinstance(use, str):
        return USES.index(use)
    return int(use)


def pmcc(a, b=None, use="everything", precision: str = "bf16x3", stream=None, out=None):
    """Pearson correlation of the rows of ``a`` (na x dim float32 CUDA
    tensor, one variable per row) with the rows of ``b`` (default ``a``:
    symmetric result).  Returns ``(pairs, corr, ts)``, na x nb float32
    tensors (or fills ``out``, a tuple of three such tensors)."""
    import torch
    for t, name in ((a, "a"), (b, "b")):
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
            raise TypeError("%s must be a float32 torch tensor on the GPU" % name)
        if t.dim() != 2 or t.stride(1) != 1:
            raise ValueError("%s must be 2-D with unit column stride" % name)
    if b is not None and b.shape[1] != a.shape[1]:
        raise ValueError("a and b must have the same number of samples")
    u = _use_id(use)
    prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    na, dim = a.shape
    nb = na if b is None else b.shape[0]
    if out is None:
        out = tuple(torch.empty((na, nb), dtype=torch.float32, device=a.device) for _ in range(3))
    pairs, corr, ts = out
    for t in out:
        if (t.shape != (na, nb) or t.dtype != torch.float32 or not t.is_cuda
                or t.stride() != pairs.stride() or t.stride(1) != 1):
            raise ValueError("out must be three float32 na x nb tensors with equal strides")
    nbytes = ctypes.c_size_t()
    check(lib().gt_pmcc_workspace_size(u, prec, na, nb, dim, 1 if b is None else 0,
                                       ctypes.byref(nbytes)))
    ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=a.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().gt_pmcc(u, prec, a.data_ptr(), na, a.stride(0),
                        None if b is None else b.data_ptr(), nb,
                        a.stride(0) if b is None else b.stride(0), dim, pairs.data_ptr(),
                        corr.data_ptr(), ts.data_ptr(), pairs.stride(0), ws.data_ptr(),
                        nbytes.value, s.cuda_stream or None))
    return pairs, corr, ts


def kendall(a, b=None, stream=None, out=None, workspace=None):
    """Kendall's tau (the reference's definition: sample pairs tied in both
    variables count as concordant) of the rows of ``a`` (na x n float32 CUDA
    tensor, one variable per row) with the rows of ``b`` (default ``a``).
    Returns an nb x na float64 tensor, entry (y, x) = tau(a[x], b[y]) (the
    reference's layout), or fills ``out``.  ``workspace``: an optional uint8
    CUDA tensor to use instead of a fresh one (any contents)."""
    import torch
    for t, name in ((a, "a"), (b, "b")):
        if t is None:
            continue
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
            raise TypeError("%s must be a float32 torch tensor on the GPU" % name)
        if t.dim() != 2 or t.stride(1) != 1:
            raise ValueError("%s must be 2-D with unit column stride" % name)
    if b is not None and b.shape[1] != a.shape[1]:
        raise ValueError("a and b must have the same number of samples")
    na, n = a.shape
    nb = na if b is None else b.shape[0]
    if out is None:
        out = torch.empty((nb, na), dtype=torch.float64, device=a.device)
    if out.shape != (nb, na) or out.dtype != torch.float64 or out.stride(1) != 1:
        raise ValueError("out must be an nb x na float64 tensor with unit column stride")
    nbytes = ctypes.c_size_t()
    check(lib().gt_kendall_workspace_size(na, nb, n, 1 if b is None else 0, ctypes.byref(nbytes)))
    if workspace is None:
        ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=a.device)
    else:
        if workspace.dtype != torch.uint8 or not workspace.is_cuda or workspace.numel() < nbytes.value:
            raise ValueError("workspace must be a uint8 CUDA tensor of at least %d bytes"
                             % nbytes.value)
        ws = workspace
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().gt_kendall(a.data_ptr(), na, a.stride(0), None if b is None else b.data_ptr(), nb,
                           a.stride(0) if b is None else b.stride(0), n, out.data_ptr(),
                           out.stride(0), ws.data_ptr(), ws.numel(), s.cuda_stream or None))
    return out


def ttest(pairs, corr, stream=None):
    """gpuSignif on device tensors of equal size: t-scores without the cap."""
    import torch
    if pairs.shape != corr.shape or not pairs.is_contiguous() or not corr.is_contiguous():
        raise ValueError("pairs and corr must be contiguous tensors of the same shape")
    ts = torch.empty_like(corr)
    s = stream if stream is not None else torch.cuda.current_stream()
    check(lib().gt_ttest(pairs.data_ptr(), corr.data_ptr(), corr.numel(), ts.data_ptr(),
                         s.cuda_stream or None))
    return ts
