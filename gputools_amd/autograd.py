"""torch.autograd wrappers of the distance kernels (ctypes into the C-ABI).

``pdist(x, method, p)`` returns the packed distance vector (R 'dist' order)
and differentiates through ``gt_dist_backward`` (euclidean, minkowski p >= 1),
which recomputes distances tile by tile instead of storing them.

``mds_stress(y, delta)`` is the metric-MDS objective
S(Y) = sum_{i<j} (||y_i - y_j|| - delta_ij)^2 with delta packed; forward and
gradient come from one fused kernel (``gt_mds_stress``), S is reduced in FP64.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check, lib
from .dist import _check_points, _method_id, _stream_ptr, dist


class PairwiseDistance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, method="euclidean", p=2.0, precision="fp32"):
        m = _method_id(method)
        if m not in (_lib.EUCLIDEAN, _lib.MINKOWSKI):
            raise ValueError("pdist is differentiable for euclidean and minkowski only")
        x = x.contiguous()
        ctx.save_for_backward(x)
        ctx.method, ctx.p = m, float(p)
        return dist(x, m, p, precision=precision)

    @staticmethod
    def backward(ctx, grad):
        (x,) = ctx.saved_tensors
        n, k = x.shape
        grad = grad.contiguous()
        dx = torch.empty_like(x)
        check(lib().gt_dist_backward(x.data_ptr(), n, k, x.stride(0), ctx.method, ctx.p,
                                     grad.data_ptr() if grad.numel() else None, dx.data_ptr(),
                                     dx.stride(0), _stream_ptr()))
        return dx, None, None, None


def pdist(x, method="euclidean", p=2.0, precision="fp32"):
    """Differentiable packed pairwise distances of the rows of x (n x k FP32)."""
    _check_points(x)
    return PairwiseDistance.apply(x, method, p, precision)


def dist_backward(x, grad_packed, method="euclidean", p=2.0):
    """dL/dX for a packed upstream gradient, without autograd."""
    _check_points(x)
    n, k = x.shape
    dx = torch.empty_like(x)
    check(lib().gt_dist_backward(x.data_ptr(), n, k, x.stride(0), _method_id(method), float(p),
                                 grad_packed.data_ptr() if grad_packed.numel() else None,
                                 dx.data_ptr(), dx.stride(0), _stream_ptr()))
    return dx


def _mds(y, delta, want_grad):
    _check_points(y, "y")
    n, q = y.shape
    if delta.numel() != n * (n - 1) // 2 or delta.dtype != torch.float32 or not delta.is_cuda:
        raise ValueError("delta must be the packed FP32 distance vector of n points on the GPU")
    delta = delta.contiguous()
    stress = torch.empty((), dtype=torch.float64, device=y.device)
    dy = torch.empty_like(y) if want_grad else None
    check(lib().gt_mds_stress(y.data_ptr(), n, q, y.stride(0),
                              delta.data_ptr() if delta.numel() else None,
                              dy.data_ptr() if dy is not None else None,
                              dy.stride(0) if dy is not None else q, stress.data_ptr(),
                              _stream_ptr()))
    return stress, dy


class MDSStress(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, delta):
        y = y.contiguous()
        stress, dy = _mds(y, delta, want_grad=y.requires_grad)
        ctx.save_for_backward(dy if dy is not None else torch.empty(0, device=y.device))
        return stress

    @staticmethod
    def backward(ctx, grad):
        (dy,) = ctx.saved_tensors
        return (grad.to(dy.dtype) * dy), None


def mds_stress(y, delta):
    """Metric-MDS stress (FP64 scalar tensor), differentiable w.r.t. y."""
    return MDSStress.apply(y, delta)
