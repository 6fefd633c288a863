"""Size-independent properties of the distances on the GPU (all methods, the
split-precision forms included).

* Scaling by a power of two is exact in FP32 arithmetic barring overflow and
  underflow: every difference, product, sum, centre (on its power-of-two
  grid), split term and correctly rounded root scales exactly, so
  dist(2^s x) = 2^s dist(x) bit for bit for euclidean (every precision),
  maximum and manhattan, and canberra and binary do not change at all.
  Minkowski's root goes through exp2/log2 and is checked to 1e-6.
* Permuting the points permutes the distances: each pair's arithmetic does
  not depend on where its rows sit (bit for bit for the FP32 kernels; the
  split forms' column centres are sums in row order, so those are checked
  to the 1e-5 bar).  Canberra chooses its per-chunk form (reciprocal with
  fma, with a zero mask, or the IEEE divide) from every value its tile
  stages, so a pair's rounding can depend on the rows beside it: within
  1.5 ulp per term, checked to 1e-6.
* The triangle inequality for the metrics (p >= 1) on random triples, with
  FP32 rounding slack.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

N, K = 700, 100
CASES = [("euclidean", "fp32"), ("euclidean", "bf16x3"), ("euclidean", "fp16x3"),
         ("maximum", "fp32"), ("manhattan", "fp32"), ("canberra", "fp32"),
         ("binary", "fp32"), ("minkowski", "fp32")]


def full(x, method, precision):
    p = 3.0 if method == "minkowski" else 2.0
    d = gt.dist(torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda(), method, p,
                precision=precision, layout="full")
    torch.cuda.synchronize()
    return d.cpu().numpy()


def inputs(method, seed):
    x = wl.method_input(method, N, K, seed=seed)
    if method in ("euclidean", "maximum", "manhattan", "minkowski"):
        x = x + np.float32(0.5)          # a mean: the centring and the norms matter
    return x


@pytest.mark.parametrize("method,precision", CASES)
@pytest.mark.parametrize("s", [3, -5])
def test_power_of_two_scaling(method, precision, s):
    x = inputs(method, 61)
    print("case scaling %s %s 2^%d" % (method, precision, s), flush=True)
    d0 = full(x, method, precision)
    d1 = full(x * np.float32(2.0 ** s), method, precision)
    if method in ("canberra", "binary"):
        np.testing.assert_array_equal(d1, d0)
    elif method == "minkowski":
        np.testing.assert_allclose(d1, d0 * np.float32(2.0 ** s), rtol=1e-6, atol=0)
    else:
        np.testing.assert_array_equal(d1, d0 * np.float32(2.0 ** s))


@pytest.mark.parametrize("method,precision", CASES)
def test_permutation_equivariance(method, precision):
    x = inputs(method, 62)
    perm = np.random.default_rng(63).permutation(N)
    print("case permutation %s %s" % (method, precision), flush=True)
    d0 = full(x, method, precision)
    d1 = full(x[perm], method, precision)
    want = d0[np.ix_(perm, perm)]
    if method == "canberra":
        np.testing.assert_allclose(d1, want, rtol=1e-6, atol=0)
    elif precision == "fp32":
        np.testing.assert_array_equal(d1, want)
    else:
        np.testing.assert_allclose(d1, want, rtol=1e-5, atol=0)


@pytest.mark.parametrize("method,precision", [c for c in CASES
                                              if c[0] not in ("canberra", "binary")])
def test_triangle_inequality(method, precision):
    x = inputs(method, 64)
    d = full(x, method, precision).astype(np.float64)
    rng = np.random.default_rng(65)
    i, j, k = (rng.integers(0, N, 200000) for _ in range(3))
    slack = 4e-6 * (d[i, j] + d[j, k]) + 1e-30
    assert np.all(d[i, k] <= d[i, j] + d[j, k] + slack)


@pytest.mark.parametrize("method,precision", CASES)
def test_known_answers(method, precision):
    """Unit vectors e_i, e_j (i != j) of k = 64 features: euclidean sqrt 2,
    maximum 1, manhattan 2, minkowski 2^(1/p), binary 1, canberra k (R drops
    the 0/0 terms and rescales by k / 2); a repeated point is at 0; x and
    x + c 1 (c = 0.5, all features) are |c| sqrt k apart in euclidean, |c| in
    maximum and |c| k in manhattan."""
    k = 64
    x = np.vstack([np.eye(k, dtype=np.float32), np.eye(k, dtype=np.float32)[:1]])   # row 64 = e_0
    d = full(x, method, precision)
    want = {"euclidean": np.sqrt(np.float32(2.0)), "maximum": 1.0, "manhattan": 2.0,
            "binary": 1.0, "canberra": float(k), "minkowski": 2.0 ** (1.0 / 3.0)}[method]
    off = ~np.eye(k + 1, dtype=bool)
    off[0, k] = off[k, 0] = False
    vals = d[off]
    if method == "minkowski":
        np.testing.assert_allclose(vals, want, rtol=1e-6)
    else:
        assert np.all(vals == np.float32(want)), (method, np.unique(vals)[:5])
    assert d[0, k] == 0.0 and d[k, 0] == 0.0
    if method in ("euclidean", "maximum", "manhattan"):
        y = wl.normal(2, k, seed=66)
        y[1] = y[0] + np.float32(0.5)
        d2 = full(y, method, precision)[1, 0]
        expect = {"euclidean": 0.5 * np.sqrt(k), "maximum": 0.5, "manhattan": 0.5 * k}[method]
        assert abs(float(d2) - expect) <= 1e-6 * expect, (d2, expect)
