"""GPU parity of the backward kernels (dL/dX) and the fused MDS stress.

The FP64 oracle (oracle_dist_backward / oracle_mds_stress, itself pinned
against central differences in test_oracle.py) is the reference.  The GPU
computes in FP32 and sums n terms per output, so gradients are compared by
relative Frobenius error (<= 1e-5) plus an elementwise check scaled by the
largest gradient entry (<= 1e-4); the stress by relative error <= 1e-6.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def assert_grad_close(got, want, frob=1e-5, elem=1e-4):
    got = got.astype(np.float64)
    scale = np.max(np.abs(want))
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel <= frob, "relative Frobenius error %.3g" % rel
    assert np.max(np.abs(got - want)) <= elem * scale


@pytest.mark.parametrize("method,p", [("euclidean", 2.0), ("minkowski", 3.0), ("minkowski", 1.5),
                                      ("minkowski", 2.0), ("minkowski", 4.5)])
@pytest.mark.parametrize("n,k", [(300, 37), (129, 5), (256, 128), (70, 200)])
def test_backward_vs_oracle(method, p, n, k):
    rng = np.random.default_rng(n + k)
    x = rng.standard_normal((n, k)).astype(np.float32)
    g = rng.standard_normal(n * (n - 1) // 2).astype(np.float32)
    got = gt.dist_backward(dev(x), dev(g), method, p).cpu().numpy()
    want = orc.dist_backward(x, g.astype(np.float64), method, p)
    assert_grad_close(got, want)


def test_backward_duplicate_rows_are_zero_subgradient():
    rng = np.random.default_rng(40)
    x = rng.standard_normal((50, 8)).astype(np.float32)
    x[10] = x[3]
    g = np.ones(50 * 49 // 2, np.float32)
    got = gt.dist_backward(dev(x), dev(g), "euclidean").cpu().numpy()
    want = orc.dist_backward(x, g.astype(np.float64), "euclidean")
    assert np.all(np.isfinite(got))
    assert_grad_close(got, want)


def test_autograd_function_matches_backward():
    rng = np.random.default_rng(41)
    x = dev(rng.standard_normal((200, 30)))
    w = dev(rng.standard_normal(200 * 199 // 2))
    xr = x.clone().requires_grad_(True)
    loss = (gt.pdist(xr, "euclidean") * w).sum()
    loss.backward()
    direct = gt.dist_backward(x, w, "euclidean")
    np.testing.assert_array_equal(xr.grad.cpu().numpy(), direct.cpu().numpy())
    # forward value of pdist equals gt.dist
    np.testing.assert_array_equal(gt.pdist(x).detach().cpu().numpy(), gt.dist(x).cpu().numpy())


def test_gradcheck_n256_against_fp64_finite_differences():
    """n = 256: the FP32 kernel gradient against central differences of the
    FP64 oracle loss L(X) = sum_q w_q d_q(X), on a random subset of entries."""
    rng = np.random.default_rng(42)
    n, k = 256, 6
    x = rng.standard_normal((n, k)).astype(np.float32)
    w = rng.standard_normal(n * (n - 1) // 2)
    got = gt.dist_backward(dev(x), dev(w), "euclidean").cpu().numpy()
    h = 1e-3
    for idx in [tuple(v) for v in rng.integers(0, [n, k], size=(20, 2))]:
        xp = x.copy(); xm = x.copy()
        xp[idx] = np.float32(x[idx] + h)
        xm[idx] = np.float32(x[idx] - h)
        fd = (np.dot(w, orc.dist(xp)) - np.dot(w, orc.dist(xm))) / (float(xp[idx]) - float(xm[idx]))
        assert abs(got[idx] - fd) <= 2e-3 * max(1.0, abs(fd)), (idx, got[idx], fd)


@pytest.mark.parametrize("q", [2, 32])
def test_mds_stress_n256_vs_oracle_and_autograd(q):
    rng = np.random.default_rng(43 + q)
    n = 256
    xs = rng.standard_normal((n, 16)).astype(np.float32)
    delta = orc.dist(xs).astype(np.float32)
    y = rng.standard_normal((n, q)).astype(np.float32)
    yt = dev(y).requires_grad_(True)
    s = gt.mds_stress(yt, dev(delta))
    s.backward()
    s_ref, dy_ref = orc.mds_stress(y, delta.astype(np.float64))
    assert s.dtype == torch.float64
    assert abs(s.item() - s_ref) <= 1e-6 * s_ref
    assert_grad_close(yt.grad.cpu().numpy(), dy_ref)


@pytest.mark.parametrize("q", [2, 32])
def test_mds_stress_16384(q):
    """The differentiation workload: delta from 16384 x 128 N(0,1) points,
    embedding Y 16384 x q, forward stress plus dY with tile recompute."""
    n = 16384
    print("case mds 16384 q=%d" % q, flush=True)
    x = wl.normal(n, 128, seed=4)
    delta = gt.dist(dev(x))
    y = wl.normal(n, q, seed=44)
    yt = dev(y).requires_grad_(True)
    s = gt.mds_stress(yt, delta)
    s.backward()
    torch.cuda.synchronize()
    s_ref, dy_ref = orc.mds_stress(y, delta.cpu().numpy().astype(np.float64))
    assert abs(s.item() - s_ref) <= 1e-6 * s_ref
    assert_grad_close(yt.grad.cpu().numpy(), dy_ref)


def test_backward_rejects_other_methods():
    x = dev(np.ones((4, 3)))
    g = dev(np.ones(6))
    with pytest.raises(gt.GputoolsError, match="UNSUPPORTED"):
        gt.dist_backward(x, g, "manhattan")
    with pytest.raises(gt.GputoolsError, match="INVALID"):
        gt.dist_backward(x, g, "minkowski", 0.5)
