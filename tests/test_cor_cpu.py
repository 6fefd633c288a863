"""CPU checks of the correlation path: the oracle against independent
references, and the host-side argument handling (no GPU compute here).

Pins of the oracle (oracle/cor_oracle.c, pmcc in FP64 and the t-scores in FP32):
  * numpy.corrcoef for use = "everything";
  * pandas DataFrame.corr (pairwise-complete observations, its own Cython
    implementation) for use = "pairwise.complete.obs", and the sample counts
    against a product of presence masks;
  * textbook cases in closed form (a perfect line, orthogonal vectors, a
    five-point example with r = sqrt(0.6)) and the reference's special values
    (no samples: 0 / (0 - 1) = -0; one sample or a constant variable: NaN);
  * the t-score against t = r sqrt((n - 2) / (1 - r^2)) in FP64, the cap of
    hostSignif, and the FP32 rounding sequence of the host code (cor^2 rounded
    before the subtraction), which a fused multiply-add would change.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

import gputools_amd as gt  # noqa: E402
from gputools_amd import _lib  # noqa: E402
from gputools_amd.cor import COR_METHODS, USES, _match  # noqa: E402


def _vars(rng, nvar, size, shift=0.0):
    return (rng.standard_normal((nvar, size)) + shift).astype(np.float32)


@pytest.mark.parametrize("na,nb,size", [(5, 10, 30), (17, 3, 200), (1, 1, 2), (40, 40, 7)])
def test_oracle_everything_matches_numpy_corrcoef(na, nb, size):
    rng = np.random.default_rng(na * 1000 + size)
    a = _vars(rng, na, size, shift=3.0)
    b = _vars(rng, nb, size, shift=-1.5)
    pairs, corr = orc.pmcc(a, b, "everything")
    ref = np.corrcoef(a.astype(np.float64), b.astype(np.float64))[:na, na:]
    np.testing.assert_allclose(corr, ref, rtol=0, atol=1e-12)
    assert np.all(pairs == size)


def test_oracle_pairwise_matches_pandas():
    pd = pytest.importorskip("pandas")
    rng = np.random.default_rng(3)
    a = _vars(rng, 12, 60, 2.0)
    b = _vars(rng, 9, 60, -4.0)
    a[rng.random(a.shape) < 0.2] = np.nan
    b[rng.random(b.shape) < 0.3] = np.nan
    pairs, corr = orc.pmcc(a, b, "pairwise.complete.obs")
    frame = pd.DataFrame(np.concatenate([a, b]).T.astype(np.float64))
    ref = frame.corr(method="pearson", min_periods=1).values[:12, 12:]
    np.testing.assert_allclose(corr, ref, rtol=0, atol=1e-12)
    pa, pb = ~np.isnan(a), ~np.isnan(b)
    np.testing.assert_array_equal(pairs, pa.astype(np.float64) @ pb.T.astype(np.float64))


def test_oracle_closed_forms():
    x = np.array([[1, 2, 3, 4, 5]], np.float32)
    y = np.array([[2, 4, 5, 4, 5]], np.float32)
    assert abs(orc.pmcc(x, y)[1][0, 0] - np.sqrt(0.6)) < 1e-15
    line = np.array([[-3 * v + 7 for v in range(1, 9)]], np.float32)
    assert abs(orc.pmcc(np.arange(1, 9, dtype=np.float32)[None], line)[1][0, 0] + 1) < 1e-15
    u = np.array([[1, -1, 1, -1]], np.float32)
    v = np.array([[1, 1, -1, -1]], np.float32)
    assert orc.pmcc(u, v)[1][0, 0] == 0.0
    # pairwise: rows 1, 4, 5 are complete and there y = 2 x
    xp = np.array([[1, 2, np.nan, 4, 5]], np.float32)
    yp = np.array([[2, np.nan, 6, 8, 10]], np.float32)
    pairs, corr = orc.pmcc(xp, yp, "pairwise.complete.obs")
    assert pairs[0, 0] == 3 and abs(corr[0, 0] - 1) < 1e-15


def test_oracle_special_values_follow_the_reference_formulas():
    nan = np.nan
    a = np.array([[nan, 1, 2, 3], [1, 1, 1, 1], [5, nan, nan, nan]], np.float32)
    b = np.array([[1, nan, nan, nan], [4, 3, 2, 1]], np.float32)
    pairs, corr = orc.pmcc(a, b, "pairwise.complete.obs")
    # no complete sample: sum 0 over n - 1 = -1
    assert pairs[0, 0] == 0 and corr[0, 0] == 0 and np.signbit(corr[0, 0])
    # one complete sample: zero deviation over zero -> NaN
    assert pairs[1, 0] == 1 and np.isnan(corr[1, 0])
    # a constant variable: standard deviation 0 -> NaN
    assert pairs[1, 1] == 4 and np.isnan(corr[1, 1])
    # "everything": a NaN anywhere in a variable makes its correlations NaN
    _, corr_e = orc.pmcc(a, b, "everything")
    assert np.isnan(corr_e[0]).all() and np.isnan(corr_e[:, 0]).all()
    # no samples at all
    pairs0, corr0 = orc.pmcc(np.zeros((2, 0), np.float32), np.zeros((3, 0), np.float32))
    assert (pairs0 == 0).all() and (corr0 == 0).all() and np.signbit(corr0).all()


def test_signif_closed_form_and_cap():
    rng = np.random.default_rng(4)
    r = rng.uniform(-0.998, 0.998, 2000).astype(np.float32)
    n = rng.integers(3, 5000, 2000).astype(np.float32)
    t = orc.signif(n, r)
    r64, n64 = r.astype(np.float64), n.astype(np.float64)
    want = r64 * np.sqrt((n64 - 2) / (1 - r64 * r64))
    # FP32 rounding, amplified by the cancellation in 1 - r^2
    tol = 4e-7 * (1 + r64 * r64 / (1 - r64 * r64))
    assert np.all(np.abs(t - want) <= tol * np.abs(want))
    # hostSignif's cap, and its absence in gpuSignif
    assert orc.signif(np.float32([10, 10]), np.float32([0.999, 1.5])).tolist() == [10000, 10000]
    nocap = orc.signif(np.float32([10]), np.float32([0.9995]), cap=False)[0]
    assert abs(nocap - 0.9995 * np.sqrt(8 / (1 - 0.9995 ** 2))) < 1e-2
    # r = -1 divides by zero; fewer than two samples gives a negative radicand
    assert orc.signif(np.float32([10]), np.float32([-1.0]))[0] == -np.inf
    assert np.isnan(orc.signif(np.float32([1]), np.float32([0.5]))[0])
    assert np.isnan(orc.signif(np.float32([0]), np.float32([-0.0]))[0])


def test_signif_rounds_like_the_host_code():
    """Each FP32 operation rounded separately: 1 - r*r with r*r rounded first
    (an fma would round once) -- checked on values where the two differ."""
    rng = np.random.default_rng(5)
    r = rng.uniform(-0.99, 0.99, 200000).astype(np.float32)
    sep = np.float32(1) - r * r
    fused = (1.0 - r.astype(np.float64) * r.astype(np.float64)).astype(np.float32)
    differ = sep != fused
    assert differ.sum() > 1000
    n = np.full(r.shape, 50, np.float32)
    t = orc.signif(n, r)
    want = r * np.sqrt((n - np.float32(2)) / sep)
    np.testing.assert_array_equal(t.view(np.uint32), want.view(np.uint32))


def test_method_and_use_matching_follow_pmatch_nomatch():
    assert _match("p", COR_METHODS) == "pearson"
    assert _match("kendall", COR_METHODS) == "kendall"
    assert _match("pairwise", USES) == "pairwise.complete.obs"
    assert _match("every", USES) == "everything"
    x = np.zeros((4, 2))
    # pmatch(..., -1): unknown and ambiguous names both give -1
    for kw in ({"method": "spearman"}, {"method": ""}, {"use": "complete.obs"}, {"use": "all"}):
        with pytest.raises(ValueError, match="ambiguous correlation method"):
            gt.gpuCor(x, **kw)


def test_gpucor_rejects_unequal_rows():
    with pytest.raises(ValueError, match="same number of rows"):
        gt.gpuCor(np.zeros((4, 2)), np.zeros((5, 2)))


def test_capi_rejects_bad_pmcc_arguments_before_touching_the_device():
    L = _lib.lib()
    buf = (ctypes.c_float * 16)()
    p = ctypes.addressof(buf)
    # unknown use
    assert L.gt_pmcc(2, 0, p, 2, 2, p, 2, 2, 2, p, p, p, 2, None, 0, None) == _lib.GT_ERR_INVALID
    # pairwise complete is FP32 only
    assert L.gt_pmcc(1, 1, p, 2, 2, p, 2, 2, 2, p, p, p, 2, None, 0, None) == _lib.GT_ERR_UNSUPPORTED
    assert "FP32" in _lib.last_error()
    # ldo below nb, leading dimension below dim, null outputs, huge dim
    assert L.gt_pmcc(0, 0, p, 2, 2, p, 3, 2, 2, p, p, p, 2, None, 0, None) == _lib.GT_ERR_INVALID
    assert L.gt_pmcc(0, 0, p, 2, 1, p, 2, 2, 2, p, p, p, 2, None, 0, None) == _lib.GT_ERR_INVALID
    assert L.gt_pmcc(0, 0, p, 2, 2, p, 2, 2, 2, None, p, p, 2, None, 0, None) == _lib.GT_ERR_INVALID
    assert L.gt_pmcc(0, 0, p, 2, 1 << 24, p, 2, 1 << 24, 1 << 24, p, p, p, 2, None, 0,
                     None) == _lib.GT_ERR_INVALID
    # the split forms need a workspace
    assert L.gt_pmcc(0, 1, p, 2, 2, p, 2, 2, 2, p, p, p, 2, None, 0, None) == _lib.GT_ERR_WORKSPACE
    # rpmcc's own check (src/rinterface.cpp:47-48)
    w = ctypes.c_int(3)
    one = ctypes.c_int(1)
    fp = ctypes.cast(p, ctypes.POINTER(ctypes.c_float))
    L.rpmcc(ctypes.byref(w), fp, ctypes.byref(one), fp, ctypes.byref(one), ctypes.byref(one), fp,
            fp, fp)
    assert "unknown use method" in _lib.last_error()


def test_pmcc_workspace_sizes():
    L = _lib.lib()
    b = ctypes.c_size_t()
    assert L.gt_pmcc_workspace_size(1, 0, 100, 50, 30, 0, ctypes.byref(b)) == 0 and b.value == 0
    assert L.gt_pmcc_workspace_size(0, 0, 100, 50, 30, 0, ctypes.byref(b)) == 0
    fp32 = b.value
    assert fp32 >= 150 * 30 * 4
    assert L.gt_pmcc_workspace_size(0, 0, 100, 100, 30, 1, ctypes.byref(b)) == 0
    assert 100 * 30 * 4 <= b.value < fp32
    assert L.gt_pmcc_workspace_size(0, 1, 100, 50, 30, 0, ctypes.byref(b)) == 0
    # three BF16 terms of the samples padded to 32, plus one scale per variable
    assert b.value >= 3 * 150 * 32 * 2 + 150 * 4
    assert L.gt_pmcc_workspace_size(0, 1, 0, 50, 30, 0, ctypes.byref(b)) == 0 and b.value == 0


# ------------------------------- Kendall -----------------------------------
# Pins of oracle_kendall (gpuKendall's definition: pairs tied in both
# variables count as concordant, NaN comparisons are false):
#   * scipy.stats.kendalltau on tie-free data, where that definition and
#     scipy's tau-b coincide with (C - D) / (n (n - 1) / 2);
#   * closed forms: identical / reversed orderings (+1 / -1), a constant
#     variable against distinct values (no concordant pair: -1) and against
#     another constant (every pair tied in both: +1), n < 2 (0 / 0);
#   * brute-force enumeration of the three tests on tiny inputs with ties and
#     NaN, written independently of the oracle's loop order.

@pytest.mark.parametrize("na,nb,n", [(5, 10, 30), (3, 2, 200), (1, 1, 2)])
def test_kendall_oracle_matches_scipy_without_ties(na, nb, n):
    stats = pytest.importorskip("scipy.stats")
    rng = np.random.default_rng(na + nb + n)
    a = rng.standard_normal((na, n)).astype(np.float32)
    b = rng.standard_normal((nb, n)).astype(np.float32)
    tau, count = orc.kendall(a, b)
    for y in range(nb):
        for x in range(na):
            want = stats.kendalltau(a[x], b[y]).statistic
            assert abs(tau[y, x] - want) < 1e-12


def test_kendall_oracle_closed_forms():
    n = 40
    up = np.arange(n, dtype=np.float32)[None]
    tau, count = orc.kendall(up, np.concatenate([up, -up, np.zeros_like(up)]))
    assert tau[0, 0] == 1.0 and tau[1, 0] == -1.0 and tau[2, 0] == -1.0
    assert count[0, 0] == n * (n - 1) // 2 and count[2, 0] == 0
    const = np.full((1, n), 3.0, np.float32)
    assert orc.kendall(const, const)[0][0, 0] == 1.0
    # x = (1, 2, 3), y = (1, 3, 2): pairs (1,2) and (1,3) concordant, (2,3) not
    t, c = orc.kendall(np.float32([[1, 2, 3]]), np.float32([[1, 3, 2]]))
    assert c[0, 0] == 2 and abs(t[0, 0] - 1 / 3) < 1e-15
    for n0 in (0, 1):
        t0, _ = orc.kendall(np.zeros((2, n0), np.float32), np.zeros((1, n0), np.float32))
        assert np.isnan(t0).all()


def test_kendall_oracle_brute_force_with_ties_and_nan():
    import itertools
    rng = np.random.default_rng(9)
    for _ in range(30):
        n = int(rng.integers(2, 12))
        a = rng.integers(0, 3, n).astype(np.float32)
        b = rng.integers(0, 3, n).astype(np.float32)
        a[rng.random(n) < 0.15] = np.nan
        c = 0
        for i, j in itertools.combinations(range(n), 2):
            sa, sb = a[j] - a[i], b[j] - b[i]   # NaN when either value is NaN
            if (sa > 0 and sb > 0) or (sa < 0 and sb < 0) or (sa == 0 and sb == 0):
                c += 1
        tau, count = orc.kendall(a[None], b[None])
        assert count[0, 0] == c
        k = n * (n - 1) / 2
        assert abs(tau[0, 0] - (2 * c - k) / k) < 1e-12


def test_kendall_oracle_rounds_large_counts_in_fp32():
    """Above 2^24 concordant pairs the FP32 steps round (n = 6000: 17,997,000
    sample pairs; identical variables are concordant on all of them)."""
    n = 6000
    x = np.arange(n, dtype=np.float32)[None]
    tau, count = orc.kendall(x, x)
    assert count[0, 0] == n * (n - 1) // 2 and tau[0, 0] == 1.0
    y = x.copy()
    y[0, :1002] = y[0, :1002][::-1]   # 501,501 discordant pairs: an odd count above 2^24
    tau, count = orc.kendall(x, y)
    c = int(count[0, 0])
    cf = np.float32(c)
    denom = np.float32(np.float32(n) * np.float32(n - 1)) / np.float32(2)
    numer = cf - (denom - cf)
    assert tau[0, 0] == float(numer) / float(denom)
    assert c > 2 ** 24 and c % 2 == 1 and int(cf) != c   # the count itself rounds


def test_gpucor_kendall_warns_on_pairwise_use_and_refuses_na():
    x = np.zeros((4, 2))
    x[1, 1] = np.nan
    with pytest.warns(UserWarning, match="NA handling for Kendall"):
        # .C without NAOK: R refuses NA/NaN/Inf before the library is called
        with pytest.raises(ValueError, match="NA/NaN/Inf in foreign function call"):
            gt.gpuCor(x, method="kendall", use="pairwise.complete.obs")
    with pytest.raises(ValueError, match=r"\(arg 3\)"):
        gt.gpuCor(np.zeros((4, 2)), np.full((4, 1), np.inf), method="kendall")


def test_capi_rejects_bad_kendall_arguments():
    L = _lib.lib()
    buf = (ctypes.c_float * 16)()
    p = ctypes.addressof(buf)
    assert L.gt_kendall(p, 2, 1, p, 2, 2, 2, p, 2, None, 0, None) == _lib.GT_ERR_INVALID
    assert L.gt_kendall(p, 2, 2, p, 2, 2, 2, p, 1, None, 0, None) == _lib.GT_ERR_INVALID
    assert L.gt_kendall(p, 2, 2, p, 2, 2, 2, None, 2, None, 0, None) == _lib.GT_ERR_INVALID
    assert L.gt_kendall(p, 2, 2, p, 2, 2, (1 << 22) + 1, p, 2, None, 0, None) == _lib.GT_ERR_INVALID
    assert L.gt_kendall(p, 2, 2, p, 2, 2, 2, p, 2, None, 0, None) == _lib.GT_ERR_WORKSPACE
    b = ctypes.c_size_t()
    assert L.gt_kendall_workspace_size(4, 4, 1, 0, ctypes.byref(b)) == 0 and b.value == 0
    assert L.gt_kendall_workspace_size(4, 4, 100, 1, ctypes.byref(b)) == 0
    # 3 planes x 4 variables x ceil-padded words, plus 4 x 4 64-bit counts
    assert b.value >= 3 * 4 * (100 * 99 // 2) // 32 * 4 + 4 * 4 * 8
