"""GPU parity of the clustering stage (gt_hclust, Rhcluster, Rdistclust).

The oracle restates the reference's merge loop in its own arithmetic
(oracle/hclust_oracle.c, pinned in test_hclust_cpu.py), so the comparison is
exact: the same merge sequence, bit-identical FP32 heights, and the same R
merge matrix and order.  Integer-valued points produce many equal distances
and exercise the reference's tie order.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402

pytestmark = pytest.mark.gpu

METHODS = orc.HC_METHODS   # single .. mcquitty, in hc_method order


def full_matrix(x):
    """FP32 full euclidean matrix from the library's own distance kernel."""
    xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()
    return gt.dist(xt, "euclidean", layout="full").cpu().numpy()


def run_gpu(d, method, state="auto", lam=0.5, beta=0.5):
    dt = torch.from_numpy(np.ascontiguousarray(d, dtype=np.float32)).cuda()
    sub, sup, h = gt.hclust(dt, method, lam, beta, state=state)
    torch.cuda.synchronize()
    return sub.cpu().numpy(), sup.cpu().numpy(), h.cpu().numpy()


def assert_same(d, method, state="auto"):
    sub_g, sup_g, h_g = run_gpu(d, method, state)
    sub_o, sup_o, h_o = orc.hclust(d, method)
    np.testing.assert_array_equal(sub_g, sub_o)
    np.testing.assert_array_equal(sup_g, sup_o)
    np.testing.assert_array_equal(h_g.view(np.uint32), h_o.view(np.uint32))


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n", [2, 3, 5, 64, 65, 300])
def test_hclust_matches_oracle_random(method, n):
    rng = np.random.default_rng(1000 + n)
    d = full_matrix(rng.standard_normal((n, 7)))
    print("case hclust random n=%d %s" % (n, method), flush=True)
    assert_same(d, method)


@pytest.mark.parametrize("method", METHODS)
def test_hclust_matches_oracle_with_ties(method):
    """Small-integer points: many equal distances, the reference's tie order
    (value, then index mod 32, then index) decides."""
    rng = np.random.default_rng(77)
    d = full_matrix(rng.integers(0, 3, size=(200, 3)))
    print("case hclust ties n=200 %s" % method, flush=True)
    assert_same(d, method)


@pytest.mark.parametrize("method", ["single", "average", "ward"])
def test_hclust_matches_oracle_n2048(method):
    rng = np.random.default_rng(5)
    d = full_matrix(rng.standard_normal((2048, 16)))
    print("case hclust n=2048 %s" % method, flush=True)
    assert_same(d, method)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n", [3, 65, 300, 2048])
def test_hclust_global_state_matches_oracle(method, n):
    """The variant that keeps the per-row state in global memory (the one
    every n above 9216 runs) against the reference-order oracle, at sizes the
    oracle reaches."""
    rng = np.random.default_rng(2000 + n)
    d = full_matrix(rng.standard_normal((n, 11)))
    print("case hclust global state n=%d %s" % (n, method), flush=True)
    assert_same(d, method, state="global")


@pytest.mark.parametrize("method", METHODS)
def test_hclust_global_state_with_ties(method):
    rng = np.random.default_rng(78)
    d = full_matrix(rng.integers(0, 3, size=(300, 3)))
    print("case hclust global state ties n=300 %s" % method, flush=True)
    assert_same(d, method, state="global")


def test_hclust_global_state_clustered_data():
    rng = np.random.default_rng(9)
    c = rng.standard_normal((8, 4)).astype(np.float32) * 10
    x = c[np.arange(1500) % 8] + rng.standard_normal((1500, 4)).astype(np.float32)
    d = full_matrix(x)
    for method in ("single", "complete", "average", "ward", "centroid", "median"):
        print("case hclust global state mixture n=1500 %s" % method, flush=True)
        assert_same(d, method, state="global")


@pytest.mark.parametrize("state", ["lds", "global"])
@pytest.mark.parametrize("method,lam,beta", [("flexible", 0.25, -0.25),
                                             ("flexible_group", 0.25, -0.25),
                                             ("flexible", 0.7, 0.1)])
def test_hclust_flexible_parameters(method, lam, beta, state):
    """lambda and beta other than R's 0.5 / 0.5 reach the kernel unchanged."""
    rng = np.random.default_rng(31)
    d = full_matrix(rng.standard_normal((400, 5)))
    sub_g, sup_g, h_g = run_gpu(d, method, state, lam, beta)
    sub_o, sup_o, h_o = orc.hclust(d, method, lam, beta)
    np.testing.assert_array_equal(sub_g, sub_o)
    np.testing.assert_array_equal(sup_g, sup_o)
    np.testing.assert_array_equal(h_g.view(np.uint32), h_o.view(np.uint32))


def test_hclust_state_argument_checks():
    d = torch.zeros((9300, 9300), dtype=torch.float32, device="cuda")
    with pytest.raises(gt.GputoolsError, match="LDS placement"):
        gt.hclust(d, "single", state="lds")


def test_hclust_clustered_data_many_rescans():
    """A mixture: merges inside clusters first, then between; single linkage
    chains through large clusters (many cached minima point at them)."""
    rng = np.random.default_rng(6)
    c = rng.standard_normal((8, 4)).astype(np.float32) * 10
    x = c[np.arange(1500) % 8] + rng.standard_normal((1500, 4)).astype(np.float32)
    d = full_matrix(x)
    for method in ("single", "complete", "average"):
        print("case hclust mixture n=1500 %s" % method, flush=True)
        assert_same(d, method)


@pytest.mark.parametrize("clust", ["complete", "average", "ward", "flexible group", "single"])
def test_gpudistclust_matches_oracle(clust):
    rng = np.random.default_rng(11)
    x = rng.standard_normal((150, 9)).astype(np.float32)
    tree = gt.gpuDistClust(x, "euclidean", clust)
    # the C layer maps R's "flexible group" to single linkage, as gputools does
    cmethod = "single" if clust == "flexible group" else clust
    sub, sup, h = orc.hclust(full_matrix(x), cmethod)
    merge, order = orc.hclust_format(sub, sup)
    np.testing.assert_array_equal(tree.merge, merge)
    np.testing.assert_array_equal(tree.order, order)
    np.testing.assert_array_equal(tree.height, h.astype(np.float64))
    assert tree.method == clust and tree.dist_method == "euclidean"


def test_gpuhclust_on_gpudist_and_other_metrics():
    rng = np.random.default_rng(12)
    x = rng.standard_normal((120, 6)).astype(np.float32)
    for dmethod in ("manhattan", "maximum"):
        dd = gt.gpuDist(x, dmethod)
        tree = gt.gpuHclust(dd, "average")
        d = dd.as_matrix().astype(np.float32)
        sub, sup, h = orc.hclust(d, "average")
        merge, order = orc.hclust_format(sub, sup)
        np.testing.assert_array_equal(tree.merge, merge)
        np.testing.assert_array_equal(tree.order, order)
        np.testing.assert_array_equal(tree.height.astype(np.float32), h)
        assert tree.dist_method == dmethod


def test_gpudistclust_single_point_and_pair():
    t1 = gt.gpuDistClust(np.ones((1, 3)))
    assert t1.merge.shape == (0, 2) and list(t1.order) == [1]
    t2 = gt.gpuDistClust(np.array([[0.0, 0.0], [3.0, 4.0]]))
    assert t2.merge.tolist() == [[-1, -2]] and t2.height.tolist() == [5.0]


def mst_weights(d):
    """Sorted edge weights of a minimum spanning tree of the complete graph
    on the symmetric matrix d (Prim, O(n^2)).  Single linkage merges at
    exactly these heights, whatever the order of equal values."""
    n = d.shape[0]
    in_tree = np.zeros(n, bool)
    in_tree[0] = True
    best = d[0].copy()
    w = np.empty(n - 1, np.float32)
    for t in range(n - 1):
        cand = np.where(in_tree, np.float32(np.inf), best)
        j = int(np.argmin(cand))
        w[t] = cand[j]
        in_tree[j] = True
        np.minimum(best, d[j], out=best)
    return np.sort(w)


@pytest.mark.parametrize("n", [9216, 9300])
def test_single_linkage_heights_are_the_mst_weights(n):
    """Sizes the reference-order oracle cannot reach: the per-row state is
    held in LDS up to n = 9216 and in global memory above; both must give
    single linkage's closed form (the sorted minimum-spanning-tree weights)."""
    rng = np.random.default_rng(n)
    a = rng.random((n, n), dtype=np.float32)
    d = (a + a.T) * np.float32(0.5)
    del a
    print("case hclust single linkage n=%d" % n, flush=True)
    want = mst_weights(d)
    sub, sup, h = run_gpu(d, "single")
    assert (sub < sup).all()
    np.testing.assert_array_equal(np.sort(h), want)
    assert (np.diff(h) >= 0).all()          # single linkage heights never decrease


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("state", ["lds", "global"])
def test_hclust_power_of_two_scaling(method, state):
    """Every Lance-Williams update is linear in the distances with weights
    that do not depend on them, so clustering 8 D merges the same clusters
    at exactly 8 times the heights (FP32, bit for bit)."""
    d = full_matrix(np.random.default_rng(68).standard_normal((300, 40)).astype(np.float32))
    sub0, sup0, h0 = run_gpu(d, method, state)
    sub1, sup1, h1 = run_gpu(d * np.float32(8.0), method, state)
    np.testing.assert_array_equal(sub1, sub0)
    np.testing.assert_array_equal(sup1, sup0)
    np.testing.assert_array_equal(h1, h0 * np.float32(8.0))
