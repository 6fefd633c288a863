"""GPU parity of the split-precision MFMA Gram path (euclidean).

Bar: max relative error <= 1e-5 against the FP64 oracle for both splits
(bf16x3: three BF16 terms, six products; fp16x3: two FP16 terms, three
products); bit-exact on small-integer inputs, where every product and partial
sum is exact.  Inputs keep their squares in the normal FP32 range.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

RTOL = 1e-5
SPLITS = ["bf16x3", "fp16x3"]


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def run(x, precision, layout="packed"):
    out = gt.dist(dev(x), "euclidean", precision=precision, layout=layout)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def max_rel(got, want):
    got = got.astype(np.float64)
    zero = want == 0
    assert np.all(got[zero] == 0.0), "pairs at distance 0 must be exactly 0"
    w = want[~zero]
    return float(np.max(np.abs(got[~zero] - w) / w)) if w.size else 0.0


@pytest.mark.parametrize("precision", SPLITS)
@pytest.mark.parametrize("name,n,k", [("uniform", 1000, 64), ("normal", 2000, 256),
                                      ("mixture", 4096, 512), ("normal", 131, 70),
                                      ("normal", 257, 3), ("mixture", 700, 33),
                                      ("normal", 2, 1), ("uniform", 129, 4)])
def test_gram_vs_oracle(precision, name, n, k):
    x = getattr(wl, name)(n, k)
    print("case gram", precision, name, n, k, flush=True)
    got = run(x, precision)
    err = max_rel(got, orc.dist(x, "euclidean"))
    assert err <= RTOL, err


@pytest.mark.parametrize("precision", SPLITS)
def test_gram_small_integers_bit_exact(precision):
    rng = np.random.default_rng(31)
    x = rng.integers(-127, 128, size=(777, 512)).astype(np.float32)
    got = run(x, precision)
    np.testing.assert_array_equal(got, orc.dist(x, "euclidean").astype(np.float32))


@pytest.mark.parametrize("precision", SPLITS)
def test_gram_all_pairs_near_coincident(precision):
    """Points within 1e-3 of one centre at 1e3: without centring every pair
    would cancel; the column centres bring the rows to their spread, so the
    Gram form itself must be accurate here."""
    rng = np.random.default_rng(32)
    c = rng.standard_normal(96).astype(np.float32) * 1000
    x = c + rng.standard_normal((300, 96)).astype(np.float32) * np.float32(1e-3)
    x[7] = x[3]                   # an exact duplicate: distance 0
    got = run(x, precision)
    want = orc.dist(x, "euclidean")
    assert got[orc.packed_index(300, 7, 3)] == 0.0
    assert max_rel(got, want) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
def test_gram_two_far_clusters_many_fixup_passes(precision):
    """Two tight clusters 2e3 apart: centring cannot help the pairs inside a
    cluster (they cancel against the clusters' offset), so half of every
    tile's pairs are recomputed exactly, over many list passes."""
    rng = np.random.default_rng(36)
    c = rng.standard_normal(96).astype(np.float32) * 1000
    side = np.where(np.arange(600) % 2 == 0, 1.0, -1.0).astype(np.float32)[:, None]
    x = side * c + rng.standard_normal((600, 96)).astype(np.float32) * np.float32(1e-3)
    x[9] = x[5]                   # an exact duplicate: distance 0
    got = run(x, precision)
    want = orc.dist(x, "euclidean")
    assert got[orc.packed_index(600, 9, 5)] == 0.0
    assert max_rel(got, want) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
@pytest.mark.parametrize("name,n,k,shift", [("uniform", 4096, 512, 0.0),
                                            ("normal", 2000, 256, 100.0),
                                            ("normal", 1500, 300, 1e4),
                                            ("uniform", 3000, 64, 0.0)])
def test_gram_non_zero_mean(precision, name, n, k, shift):
    """Data far from the origin: the Gram form is taken about the column
    centres, so accuracy (and the share of exactly recomputed pairs) does
    not depend on the offset."""
    x = getattr(wl, name)(n, k, seed=37) + np.float32(shift)
    print("case gram offset", precision, name, n, k, shift, flush=True)
    got = run(x, precision)
    assert max_rel(got, orc.dist(x, "euclidean")) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
def test_gram_per_column_offsets(precision):
    rng = np.random.default_rng(38)
    x = (rng.standard_normal((1200, 200)) + rng.standard_normal(200) * 1e3).astype(np.float32)
    x[:, 17] = 5.0                # a constant column
    got = run(x, precision)
    assert max_rel(got, orc.dist(x, "euclidean")) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
@pytest.mark.parametrize("n,k", [(300, 4096), (257, 8192), (130, 16384), (64, 65536),
                                 (200, 3001)])
def test_gram_large_k(precision, n, k):
    """Many features: the accumulation term of the threshold grows like
    sqrt(k); past the point where nearly every pair would be recomputed the
    request runs on the FP32 difference kernel.  Every pair is checked."""
    x = wl.normal(n, k, seed=39) + np.float32(0.5)
    print("case gram large k", precision, n, k, flush=True)
    got = run(x, precision)
    assert max_rel(got, orc.dist(x, "euclidean")) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
def test_gram_large_k_mixture(precision):
    x = wl.mixture(512, 8192, seed=40)
    got = run(x, precision)
    assert max_rel(got, orc.dist(x, "euclidean")) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
@pytest.mark.parametrize("scale", [1e-15, 1e-3, 1e3, 1e15])
def test_gram_scale_range(precision, scale):
    x = wl.normal(300, 40, seed=33) * np.float32(scale)
    got = run(x, precision)
    assert max_rel(got, orc.dist(x, "euclidean")) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
def test_gram_full_layout_and_padding(precision):
    n, k = 333, 70
    x = wl.mixture(n, k, seed=34)
    xp = torch.zeros((n, k + 5), dtype=torch.float32, device="cuda")
    xp[:, :k] = dev(x)
    outp = torch.full((n, n + 3), -1.0, dtype=torch.float32, device="cuda")
    gt.dist(xp[:, :k], "euclidean", precision=precision, layout="full", out=outp[:, :n])
    torch.cuda.synchronize()
    full = outp.cpu().numpy()
    assert np.all(full[:, n:] == -1.0)
    m = full[:, :n]
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0.0)
    ii, jj = np.tril_indices(n, -1)
    order = np.lexsort((ii, jj))
    assert max_rel(m[ii[order], jj[order]], orc.dist(x, "euclidean")) <= RTOL


@pytest.mark.parametrize("precision", SPLITS)
def test_gram_mixture_65536x512_sampled(precision):
    """The benchmark workload at full size: random pairs and same-centre pairs."""
    n, k = 65536, 512
    x = wl.mixture(n, k, seed=3)
    print("case gram mixture 65536x512", precision, flush=True)
    got = gt.dist(dev(x), "euclidean", precision=precision)
    torch.cuda.synchronize()
    rng = np.random.default_rng(1)
    ii = rng.integers(0, n, 200000)
    jj = rng.integers(0, n, 200000)
    same = rng.integers(0, n - 64 * 8, 50000)
    ii = np.concatenate([ii, same + 64 * rng.integers(1, 8, same.size)])
    jj = np.concatenate([jj, same])
    keep = ii != jj
    ii, jj = ii[keep], jj[keep]
    hi, lo = np.maximum(ii, jj), np.minimum(ii, jj)
    idx = lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)
    vals = got[torch.from_numpy(idx).cuda()].cpu().numpy()
    want = orc.dist_pairs(x, ii, jj, "euclidean")
    assert max_rel(vals, want) <= RTOL


def test_gram_deterministic():
    x = dev(wl.mixture(3000, 128, seed=35))
    a = gt.dist(x, "euclidean", precision="bf16x3").cpu().numpy()
    b = gt.dist(x, "euclidean", precision="bf16x3").cpu().numpy()
    assert np.array_equal(a, b)


def test_gram_rejects_other_methods():
    x = dev(wl.normal(10, 4))
    with pytest.raises(gt.GputoolsError, match="UNSUPPORTED"):
        gt.dist(x, "manhattan", precision="bf16x3")
