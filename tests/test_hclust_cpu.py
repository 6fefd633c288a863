"""CPU checks of the clustering stage: the oracle against independent
references, and the library's host-side pieces (formatting, method names).

Pins of the oracle (oracle/hclust_oracle.c, the reference's merge loop):
  * scipy.cluster.hierarchy.linkage fixtures (tests/golden/hclust, written by
    make_golden_hclust.py) for single, complete, average, mcquitty / wpgma
    (scipy 'weighted'), median and centroid;
  * the second merge of three points in closed form from Lance & Williams'
    parameter table (alpha_i, alpha_j, beta, gamma), which also covers
    flexible (beta = 0.5), flexible group and ward, whose weights only
    coincide with the textbook ones for equal cluster sizes.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as orc

import gputools_amd as gt  # noqa: E402
from gputools_amd.hclust import R_CLUST_METHODS, hclust_method_id, match_clust_method  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "hclust", "*.npz")))
SCIPY_PINNED = ("single", "complete", "average", "mcquitty", "wpgma", "median", "centroid")


@pytest.mark.parametrize("method", SCIPY_PINNED)
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_matches_scipy_linkage(path, method):
    with np.load(path, allow_pickle=False) as z:
        d = z["d"]
        merge_want, h_want = z["merge_" + method], z["height_" + method]
    sub, sup, h = orc.hclust(d, method)
    merge, order = orc.hclust_format(sub, sup)
    np.testing.assert_array_equal(merge, merge_want)
    # FP32 Lance-Williams arithmetic against scipy's FP64
    np.testing.assert_allclose(h, h_want, rtol=2e-6, atol=0)
    assert sorted(order) == list(range(1, d.shape[0] + 1))


def _three_points():
    # d01 < d02 < d12: points 0 and 1 merge first, then {0,1} with 2
    d01, d02, d12 = np.float32(1.0), np.float32(2.5), np.float32(3.25)
    d = np.array([[0, d01, d02], [d01, 0, d12], [d02, d12, 0]], np.float32)
    return d, float(d01), float(d02), float(d12)


# Lance & Williams: d(k, i+j) = a_i d(k,i) + a_j d(k,j) + b d(i,j) + g |d(k,i) - d(k,j)|,
# here with n_i = n_j = n_k = 1 and, for the flexible methods, beta = 0.5 (alpha = 0.25)
@pytest.mark.parametrize("method,expect", [
    ("single", lambda a, b, c: min(b, c)),
    ("complete", lambda a, b, c: max(b, c)),
    ("average", lambda a, b, c: 0.5 * b + 0.5 * c),
    ("mcquitty", lambda a, b, c: 0.5 * b + 0.5 * c),
    ("wpgma", lambda a, b, c: 0.5 * b + 0.5 * c),
    ("median", lambda a, b, c: 0.5 * b + 0.5 * c - 0.25 * a),
    ("centroid", lambda a, b, c: 0.5 * b + 0.5 * c - 0.25 * a),
    ("flexible", lambda a, b, c: 0.25 * b + 0.25 * c + 0.5 * a),
    ("flexible_group", lambda a, b, c: 0.25 * b + 0.25 * c + 0.5 * a),
    ("ward", lambda a, b, c: (2 * b + 2 * c - a) / 3),
])
def test_oracle_three_points_closed_form(method, expect):
    d, d01, d02, d12 = _three_points()
    sub, sup, h = orc.hclust(d, method)
    assert list(sub) == [0, 1] and list(sup) == [1, 2]
    assert h[0] == np.float32(d01)
    assert abs(float(h[1]) - expect(d01, d02, d12)) <= 1e-6 * abs(expect(d01, d02, d12))


def _four_points():
    # d01 = 1 merges first (sub 0, sup 1), then {0, 1} (now at index 1, size
    # 2) with point 2, then that cluster of 3 with point 3: the second update
    # combines clusters of unequal size
    d = np.array([[0, 1, 4, 10],
                  [1, 0, 5, 12],
                  [4, 5, 0, 20],
                  [10, 12, 20, 0]], np.float32)
    return d


# Heights worked by hand from the update kernels' text (inst/cuda/hcluster.cu,
# ward at :380-388): when the cluster at index sub (np points) is absorbed
# into sup (nq points), nr = np + nq, for a cluster k of nk points with
# top = d(k, sub), bot = d(k, sup), val = the merge height:
#   ward            (bot (np + nk) + top (nq + nk) - val nk) / (nr + nk)
#   flexible        (bot (1 - lam) + top (1 - lam)) / 2 + beta val
#   flexible group  (bot (1 - lam) np + top (1 - lam) nq) / nr + beta val
# i.e. the surviving row is weighted by the absorbed cluster's size.  On the
# four points (round 2 absorbs the size-2 cluster at index 1 into point 2):
#   ward:  round 1 d(2,{01}) = (5*2 + 4*2 - 1)/3 = 17/3, d(3,{01}) = 43/3;
#          round 3 height (20*3 + 43/3*2 - 17/3)/4 = 83/4 = 20.75
#          (textbook weights would give (43/3*3 + 20*2 - 17/3)/4 = 19.33)
#   lam = 0.25, beta = -0.25:
#   flexible:       3.125, 8.0; then (20*.75 + 8*.75)/2 - 3.125/4 = 9.71875
#   flexible group: 3.125, 8.0; then (20*.75*2 + 8*.75*1)/3 - 0.78125 = 11.21875
#          (textbook weights would give 8.21875)
@pytest.mark.parametrize("method,lam,beta,heights", [
    ("ward", 0.5, 0.5, (1.0, 17.0 / 3.0, 20.75)),
    ("flexible", 0.25, -0.25, (1.0, 3.125, 9.71875)),
    ("flexible_group", 0.25, -0.25, (1.0, 3.125, 11.21875)),
    # R's lambda = beta = 0.5: round 1 (5+4)/4 + 0.5 = 2.75, (12+10)/4 + 0.5 = 6;
    # then flexible (20+6)/4 + 1.375 = 7.875, group (20*2 + 6)/6 + 1.375
    ("flexible", 0.5, 0.5, (1.0, 2.75, 7.875)),
    ("flexible_group", 0.5, 0.5, (1.0, 2.75, 46.0 / 6.0 + 1.375)),
])
def test_oracle_unequal_cluster_sizes_hand_worked(method, lam, beta, heights):
    sub, sup, h = orc.hclust(_four_points(), method, lam, beta)
    assert list(sub) == [0, 1, 2] and list(sup) == [1, 2, 3]
    np.testing.assert_allclose(h.astype(np.float64), heights, rtol=1e-6, atol=0)


def _random_sequence(rng, n):
    active, sub, sup = list(range(n)), [], []
    for _ in range(n - 1):
        a, b = sorted(rng.choice(active, 2, replace=False))
        sub.append(int(a))
        sup.append(int(b))
        active.remove(a)
    return sub, sup


def test_format_matches_reference_formatting():
    """gt_hclust_format (O(n)) against the reference's O(n^2) rewrite loop
    and recursive order (restated in the oracle) on random merge sequences."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        n = int(rng.integers(2, 80))
        sub, sup = _random_sequence(rng, n)
        m1, o1 = gt.hclust_format(sub, sup)
        m2, o2 = orc.hclust_format(sub, sup)
        np.testing.assert_array_equal(m1, m2)
        np.testing.assert_array_equal(o1, o2)


def _leaves(merge, row):
    out = []
    for e in merge[row]:
        out.extend([-e] if e < 0 else _leaves(merge, e - 1))
    return out


def test_order_has_no_crossings():
    """R's order: a permutation in which every cluster's leaves are contiguous."""
    rng = np.random.default_rng(8)
    for _ in range(50):
        n = int(rng.integers(2, 60))
        merge, order = gt.hclust_format(*_random_sequence(rng, n))
        assert sorted(order) == list(range(1, n + 1))
        pos = {int(v): i for i, v in enumerate(order)}
        for row in range(n - 1):
            p = sorted(pos[v] for v in _leaves(merge, row))
            assert p[-1] - p[0] == len(p) - 1


def test_format_single_point_and_invalid():
    merge, order = gt.hclust_format([], [])
    assert merge.shape == (0, 2) and list(order) == [1]
    with pytest.raises(gt.GputoolsError, match="INVALID"):
        gt.hclust_format([0], [0])


def test_method_names_follow_getclusterenum():
    ids = [hclust_method_id(m) for m in ("single", "complete", "wpgma", "average", "median",
                                         "centroid", "flexible_group", "flexible", "ward",
                                         "mcquitty")]
    assert ids == list(range(10))
    # gpuHclust passes R's spelling "flexible group": no match, single linkage
    assert hclust_method_id("flexible group") == 0
    assert hclust_method_id("") == 0


def test_r_method_matching():
    assert match_clust_method("comp") == "complete"
    assert match_clust_method("flexible") == "flexible"      # exact beats the longer name
    assert match_clust_method("flexible g") == "flexible group"
    for bad in ("flex", "m", "", "kmeans"):
        with pytest.raises(ValueError, match="invalid clustering method"):
            match_clust_method(bad)
    assert len(R_CLUST_METHODS) == 10


def test_gpudistclust_rejects_euclidian_like_the_reference():
    # R/gpuHclust.R:96-108 assigns the alias to an unused variable
    with pytest.raises(ValueError, match="invalid distance method"):
        gt.gpuDistClust(np.zeros((3, 2)), distmethod="euclidian")


def test_gpuhclust_argument_checks():
    d = gt.Dist(np.zeros(1), 2)
    with pytest.raises(ValueError, match="invalid clustering method"):
        gt.gpuHclust(d, "nope")
    with pytest.raises(ValueError, match="invalid dissimilarities"):
        gt.gpuHclust(np.zeros((2, 2)))
    with pytest.raises(ValueError, match="n >= 2"):
        gt.gpuHclust(gt.Dist(np.zeros(0), 1))
    with pytest.raises(ValueError, match="improper length"):
        gt.gpuHclust(gt.Dist(np.zeros(2), 3))


def test_oracle_single_linkage_heights_are_mst_weights():
    """Single linkage merges at the minimum-spanning-tree edge weights
    (Kruskal over the pairs, union-find), equal values included."""
    rng = np.random.default_rng(31)
    for n, levels in ((40, None), (120, 7)):
        a = rng.random((n, n), dtype=np.float32)
        if levels:
            a = np.floor(a * levels).astype(np.float32)   # many equal distances
        d = (a + a.T) * np.float32(0.5)
        iu = np.triu_indices(n, 1)
        parent = list(range(n))

        def root(x):
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x
        mst = []
        for v, i, j in sorted(zip(d[iu], iu[0], iu[1])):
            ri, rj = root(i), root(j)
            if ri != rj:
                parent[ri] = rj
                mst.append(v)
        _, _, h = orc.hclust(d.copy(), "single")
        np.testing.assert_array_equal(np.sort(h), np.array(mst, np.float32))
