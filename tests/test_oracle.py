"""Pins the FP64 CPU oracle (oracle/dist_oracle.c) before it is trusted.

Checks, each chosen so that a plausible slip in the oracle fails one of them:
  * golden fixtures from an independent FP64 implementation (scipy pdist);
  * closed-form known answers (unit vectors, constant shifts, R's canberra
    0/0-omission-with-rescale and a = -b -> Inf rules, binary ratios);
  * R's 'dist' packed order against numpy's lower-triangle enumeration;
  * the backward and MDS-stress gradients against central differences.
"""
import math

import numpy as np
import pytest

from conftest import golden_cases, load_golden
from oracle import oracle as orc

METHODS = orc.METHODS


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_golden(name):
    g = load_golden(name)
    got = orc.dist(g["x"], g["method"], g["p"])
    assert got.shape == g["expected"].shape
    if g["method"] in ("binary", "maximum"):
        np.testing.assert_array_equal(got, g["expected"])
    else:
        np.testing.assert_allclose(got, g["expected"], rtol=1e-12, atol=1e-13)


def test_packed_order_is_r_dist_order():
    # R: as.dist(m) keeps m[lower.tri(m)], i.e. column-major lower triangle.
    n = 9
    m = np.arange(n * n, dtype=np.float64).reshape(n, n, order="F")  # m[i, j]
    ii, jj = np.tril_indices(n, -1)
    order = np.lexsort((ii, jj))          # column-major: by j, then i
    want = m[ii[order], jj[order]]
    got = np.empty_like(want)
    for i in range(n):
        for j in range(i):
            got[orc.packed_index(n, i, j)] = m[i, j]
    np.testing.assert_array_equal(got, want)
    assert orc.packed_index(n, 2, 5) == orc.packed_index(n, 5, 2)


@pytest.mark.parametrize("k", [1, 3, 64, 513])
def test_unit_vector_known_answers(k):
    if k < 2:
        pytest.skip("needs two coordinates")
    x = np.zeros((2, k), np.float32)
    x[0, 0] = 1.0
    x[1, k - 1] = 1.0
    assert orc.dist(x, "euclidean")[0] == pytest.approx(math.sqrt(2), rel=1e-15)
    assert orc.dist(x, "maximum")[0] == 1.0
    assert orc.dist(x, "manhattan")[0] == 2.0
    # R omits the k-2 zero/zero terms and rescales the sum of 2 by k/2.
    assert orc.dist(x, "canberra")[0] == pytest.approx(k, rel=1e-15)
    assert orc.dist(x, "binary")[0] == 1.0
    for p in (1.0, 3.0, 0.5, 7.25):
        assert orc.dist(x, "minkowski", p)[0] == pytest.approx(2 ** (1 / p), rel=1e-14)


@pytest.mark.parametrize("c", [0.5, -3.0, 1e-3])
def test_constant_shift_known_answers(c):
    rng = np.random.default_rng(1)
    k = 40
    # dyadic grid: a + c is exact in FP32 for c = 0.5, so the closed form holds
    a = (np.round(rng.standard_normal(k) * 1024) / 1024).astype(np.float32)
    b = (a.astype(np.float64) + c).astype(np.float32)
    shift = b.astype(np.float64) - a.astype(np.float64)  # exact per-coordinate shift
    x = np.stack([a, b])
    assert orc.dist(x, "euclidean")[0] == pytest.approx(math.sqrt(np.sum(shift ** 2)), rel=1e-14)
    assert orc.dist(x, "maximum")[0] == np.max(np.abs(shift))
    assert orc.dist(x, "manhattan")[0] == pytest.approx(np.sum(np.abs(shift)), rel=1e-14)
    if c == 0.5:  # exact shift: closed form |c| k^(1/p)
        assert orc.dist(x, "euclidean")[0] == pytest.approx(abs(c) * math.sqrt(k), rel=1e-14)
        assert orc.dist(x, "minkowski", 3.0)[0] == pytest.approx(abs(c) * k ** (1 / 3), rel=1e-14)


def test_canberra_r_edge_rules():
    # 0/0 terms omitted and the sum rescaled by k/count: 0.5 * 3/2
    x = np.array([[1, 0, 2], [3, 0, 2]], np.float32)
    assert orc.dist(x, "canberra")[0] == pytest.approx(0.75, rel=1e-15)
    # a = -b != 0 gives x/0 = Inf
    x = np.array([[1, 0, 0, 2], [3, 0, 0, -2]], np.float32)
    assert orc.dist(x, "canberra")[0] == math.inf
    # every term 0/0: NA
    x = np.zeros((2, 4), np.float32)
    assert math.isnan(orc.dist(x, "canberra")[0])
    # mixed signs, no zero terms: sum |a-b|/|a+b| (not scipy's |a|+|b|)
    x = np.array([[1.0, -2.0], [3.0, 1.0]], np.float32)
    assert orc.dist(x, "canberra")[0] == pytest.approx(2 / 4 + 3 / 1, rel=1e-15)


def test_binary_ratio_and_zero_rows():
    x = np.array([[1, 0, 0, 5, 0, 0],
                  [2, 7, 0, 0, 0, 0],
                  [0, 0, 0, 0, 0, 0],
                  [0, 0, 0, 0, 0, 0]], np.float32)
    d = orc.dist(x, "binary")
    # (0,1): on in either = {0,1,3}, on in one = {1,3}: 2/3
    assert d[orc.packed_index(4, 1, 0)] == 2 / 3
    # (0,2): {0,3} both in one: 1
    assert d[orc.packed_index(4, 2, 0)] == 1.0
    # (2,3): no bit on anywhere: 0 (R returns 0, not NA)
    assert d[orc.packed_index(4, 3, 2)] == 0.0


def test_zero_features_give_na():
    x = np.zeros((3, 0), np.float32)
    for m in METHODS:
        assert np.all(np.isnan(orc.dist(x, m, 2.0)))


def test_invalid_minkowski_p_rejected():
    x = np.ones((3, 2), np.float32)
    for p in (0.0, -1.0, math.inf):
        with pytest.raises(ValueError):
            orc.dist(x, "minkowski", p)


def test_pairs_and_cross_agree_with_packed():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((23, 11)).astype(np.float32)
    for m in METHODS:
        full = orc.dist(x, m, 3.0)
        ii = rng.integers(0, 23, 50)
        jj = (ii + rng.integers(1, 23, 50)) % 23
        got = orc.dist_pairs(x, ii, jj, m, 3.0)
        want = np.array([full[orc.packed_index(23, i, j)] for i, j in zip(ii, jj)])
        np.testing.assert_array_equal(got, want)
        cross = orc.dist_cross(x, x, m, 3.0)
        for i, j in zip(ii, jj):
            assert cross[j, i] == full[orc.packed_index(23, i, j)]
        assert np.all(np.diag(cross) == 0.0)


def test_compare_packed_finds_planted_error():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((40, 9)).astype(np.float32)
    ref = orc.dist(x, "euclidean").astype(np.float32)
    err, _ = orc.compare_packed(x, ref, "euclidean")
    assert err < 1e-7
    bad = ref.copy()
    q = orc.packed_index(40, 31, 7)
    bad[q] *= 1.001
    err, (wi, wj) = orc.compare_packed(x, bad, "euclidean")
    assert err == pytest.approx(1e-3, rel=1e-3)
    assert (wi, wj) == (31, 7)


def _fd_grad(loss, x, h=1e-3):
    """Central differences w.r.t. an FP32 matrix, dividing by the perturbation
    that FP32 actually represents."""
    g = np.zeros(x.shape, np.float64)
    for idx in np.ndindex(*x.shape):
        xp = x.copy(); xm = x.copy()
        xp[idx] = np.float32(x[idx] + h)
        xm[idx] = np.float32(x[idx] - h)
        g[idx] = (loss(xp) - loss(xm)) / (float(xp[idx]) - float(xm[idx]))
    return g


@pytest.mark.parametrize("method,p", [("euclidean", 2.0), ("minkowski", 3.0),
                                      ("minkowski", 1.5)])
def test_backward_matches_central_differences(method, p):
    rng = np.random.default_rng(5)
    n, k = 7, 4
    x = rng.standard_normal((n, k)).astype(np.float32)
    w = rng.standard_normal(n * (n - 1) // 2)

    def loss(xx):
        return float(np.dot(w, orc.dist(xx, method, p)))

    got = orc.dist_backward(x, w, method, p)
    want = _fd_grad(loss, x)
    np.testing.assert_allclose(got, want, rtol=2e-4, atol=2e-4)


def test_mds_stress_value_and_gradient():
    rng = np.random.default_rng(6)
    n, q = 9, 2
    y = rng.standard_normal((n, q)).astype(np.float32)
    # delta equal to the embedding's own distances: stress 0, gradient 0
    delta = orc.dist(y, "euclidean")
    s, dy = orc.mds_stress(y, delta)
    assert s == pytest.approx(0.0, abs=1e-24)
    assert np.max(np.abs(dy)) < 1e-12
    delta = orc.dist(rng.standard_normal((n, 5)).astype(np.float32), "euclidean")
    s, dy = orc.mds_stress(y, delta)
    want = np.sum((orc.dist(y, "euclidean") - delta) ** 2)
    assert s == pytest.approx(want, rel=1e-13)
    fd = _fd_grad(lambda yy: orc.mds_stress(yy, delta, with_grad=False), y)
    np.testing.assert_allclose(dy, fd, rtol=2e-4, atol=2e-4)
