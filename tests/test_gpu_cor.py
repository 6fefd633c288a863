"""GPU parity of gpuCor's Pearson path (rpmcc, gt_pmcc) and gpuTtest (rtestT).

Correlations are compared with the FP64 oracle (oracle/cor_oracle.c, pinned in
test_cor_cpu.py) within an absolute tolerance of 1e-5 (|r| <= 1; FP32
arithmetic leaves about 1e-6); sample counts exactly; t-scores bit-exactly
against hostSignif / gpuSignif applied by the oracle to the GPU's own
correlations and counts; special values (NaN, -0) exactly.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-5
PRECISIONS = ("bf16x3", "fp16x3", "fp32")


def assert_bits_equal(got, want):
    """Bit-identical FP32 values (any NaN matches any NaN: payloads differ
    between the host and the GPU)."""
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    nan = np.isnan(want)
    np.testing.assert_array_equal(np.isnan(got), nan)
    np.testing.assert_array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def _data(rng, size, nvar, shift=0.0, scale=1.0):
    """R layout: samples x variables."""
    return (rng.standard_normal((size, nvar)) * scale + shift).astype(np.float32)


def _check(res, x, y, use):
    a = np.ascontiguousarray(x.T)
    b = a if y is None else np.ascontiguousarray(y.T)
    pairs, corr = orc.pmcc(a, b, use, nthreads=8)
    got = res["coefficients"]
    np.testing.assert_array_equal(res["pairs"], pairs)
    fin = np.isfinite(corr)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    err = np.abs(got[fin] - corr[fin]).max() if fin.any() else 0.0
    print("max abs error %.3g" % err, flush=True)
    assert err <= TOL, err
    ts = orc.signif(res["pairs"].astype(np.float32), got.astype(np.float32))
    assert_bits_equal(res["ts"], ts)
    return err


@pytest.mark.parametrize("size,nx,ny", [(30, 5, 10), (200, 257, 3), (77, 300, 513), (1000, 64, 64),
                                        (2, 3, 4), (1, 2, 2)])
def test_gpucor_everything_matches_oracle(size, nx, ny):
    rng = np.random.default_rng(size + nx + ny)
    x = _data(rng, size, nx, shift=5.0, scale=2.0)
    y = _data(rng, size, ny, shift=-3.0)
    print("case gpuCor everything size=%d nx=%d ny=%d" % (size, nx, ny), flush=True)
    _check(gt.gpuCor(x, y), x, y, "everything")


@pytest.mark.parametrize("size,nx", [(30, 5), (500, 300), (64, 513)])
def test_gpucor_self_is_symmetric(size, nx):
    rng = np.random.default_rng(10 + nx)
    x = _data(rng, size, nx, shift=1.0)
    print("case gpuCor self size=%d nx=%d" % (size, nx), flush=True)
    res = gt.gpuCor(x)
    _check(res, x, None, "everything")
    c = res["coefficients"]
    np.testing.assert_array_equal(c, c.T)
    np.testing.assert_array_equal(res["ts"], res["ts"].T)
    # the diagonal is 1 within rounding, so every t-score there is capped
    assert np.abs(np.diag(c) - 1).max() < 1e-6
    assert (np.diag(res["ts"]) == 10000).all()


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("self_case", [False, True])
def test_device_pmcc_all_precisions(precision, self_case):
    rng = np.random.default_rng(21)
    a = _data(rng, 700, 333, shift=2.0).T.copy()    # variables as rows
    b = _data(rng, 700, 270, shift=-1.0).T.copy()
    da = torch.from_numpy(a).cuda()
    db = None if self_case else torch.from_numpy(b).cuda()
    print("case pmcc %s self=%s" % (precision, self_case), flush=True)
    pairs, corr, ts = gt.pmcc(da, db, precision=precision)
    torch.cuda.synchronize()
    ref_p, ref_c = orc.pmcc(a, a if self_case else b, nthreads=8)
    got = corr.cpu().numpy()
    print("max abs error %.3g" % np.abs(got - ref_c).max(), flush=True)
    assert np.abs(got - ref_c).max() <= TOL
    assert (pairs.cpu().numpy() == 700).all()
    want_ts = orc.signif(pairs.cpu().numpy(), got)
    assert_bits_equal(ts.cpu().numpy(), want_ts)
    if self_case:
        np.testing.assert_array_equal(got, got.T)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_device_pmcc_self_whole_offdiagonal_tiles(precision):
    """600 variables: tile (1, 0) of the symmetric schedule is a whole
    256 x 256 tile, whose mirrored half goes out through LDS as rows."""
    rng = np.random.default_rng(23)
    a = _data(rng, 300, 600, shift=0.5).T.copy()
    print("case pmcc self 600 vars %s" % precision, flush=True)
    pairs, corr, ts = gt.pmcc(torch.from_numpy(a).cuda(), precision=precision)
    torch.cuda.synchronize()
    ref_p, ref_c = orc.pmcc(a, a, nthreads=8)
    got = corr.cpu().numpy()
    print("max abs error %.3g" % np.abs(got - ref_c).max(), flush=True)
    assert np.abs(got - ref_c).max() <= TOL
    np.testing.assert_array_equal(got, got.T)
    t = ts.cpu().numpy()
    np.testing.assert_array_equal(t, t.T)
    assert_bits_equal(t, orc.signif(pairs.cpu().numpy(), got))
    assert (pairs.cpu().numpy() == 300).all()


def test_device_pmcc_strided_output_and_input():
    rng = np.random.default_rng(22)
    base = torch.from_numpy(_data(rng, 300, 150).T.copy()).cuda()
    a = base[:, :257]                      # row stride 300, 257 samples
    b = base[:100, 3:260]
    out = tuple(torch.full((150, 104), 7.0, device="cuda")[:, :100] for _ in range(3))
    print("case pmcc strided", flush=True)
    gt.pmcc(a, b, out=out)
    torch.cuda.synchronize()
    _, ref = orc.pmcc(a.cpu().numpy(), b.cpu().numpy())
    assert np.abs(out[1].cpu().numpy() - ref).max() <= TOL
    assert (out[0].cpu().numpy() == 257).all()


def test_gpucor_everything_nan_and_constant_variables():
    rng = np.random.default_rng(23)
    x = _data(rng, 40, 6)
    x[5, 1] = np.nan          # a NaN makes every correlation of variable 1 NaN
    x[:, 3] = 2.5             # a constant variable: zero deviation -> NaN
    print("case gpuCor everything nan/constant", flush=True)
    res = gt.gpuCor(x)
    c = res["coefficients"]
    assert np.isnan(c[1]).all() and np.isnan(c[:, 1]).all()
    assert np.isnan(c[3]).all() and np.isnan(c[:, 3]).all()
    _check(res, x, None, "everything")


@pytest.mark.parametrize("size,nx,ny,miss", [(30, 5, 10, 0.1), (120, 70, 130, 0.3),
                                             (257, 65, 64, 0.05), (9, 20, 20, 0.5)])
def test_gpucor_pairwise_matches_oracle(size, nx, ny, miss):
    rng = np.random.default_rng(size * 7 + nx)
    x = _data(rng, size, nx, shift=4.0)
    y = _data(rng, size, ny, shift=-2.0, scale=3.0)
    x[rng.random(x.shape) < miss] = np.nan
    y[rng.random(y.shape) < miss] = np.nan
    print("case gpuCor pairwise size=%d nx=%d ny=%d miss=%g" % (size, nx, ny, miss), flush=True)
    _check(gt.gpuCor(x, y, use="pairwise.complete.obs"), x, y, "pairwise.complete.obs")


def test_gpucor_pairwise_self_symmetric():
    rng = np.random.default_rng(31)
    x = _data(rng, 150, 200, shift=1.0)
    x[rng.random(x.shape) < 0.2] = np.nan
    print("case gpuCor pairwise self", flush=True)
    res = gt.gpuCor(x, use="pairwise.complete.obs")
    _check(res, x, None, "pairwise.complete.obs")
    np.testing.assert_array_equal(res["coefficients"], res["coefficients"].T)
    np.testing.assert_array_equal(res["pairs"], res["pairs"].T)


def test_gpucor_pairwise_special_cases():
    nan, inf = np.nan, np.inf
    # columns: x0 and y0 share no complete row; x1/y0 share one; x2 is
    # constant on its complete rows (small integers: sums are exact); x3 has
    # an Inf only where y1 is missing, which must not leak into (x3, y1)
    x = np.array([[nan, 1, 3, 1], [nan, nan, 3, 2], [4, nan, 3, inf], [5, nan, 3, 4]], np.float32)
    y = np.array([[1, 4], [2, 3], [nan, nan], [nan, 1]], np.float32)
    print("case gpuCor pairwise special", flush=True)
    res = gt.gpuCor(x, y, use="pairwise.complete.obs")
    c, p = res["coefficients"], res["pairs"]
    assert p[0, 0] == 0 and c[0, 0] == 0 and np.signbit(c[0, 0])
    assert p[1, 0] == 1 and np.isnan(c[1, 0])
    assert np.isnan(c[2, 0]) and np.isnan(c[2, 1])
    assert p[3, 1] == 3 and np.isfinite(c[3, 1])
    _check(res, x, y, "pairwise.complete.obs")


@pytest.mark.parametrize("self_case", [False, True])
def test_gpucor_pairwise_offsets_and_drift(self_case):
    """The one-pass form shifts each variable by its mean over the first 32
    samples.  Variables here sit far from zero (N(1000, 1)).  Some drift by
    500 after the first 32 samples: the shift is off by most of the drift,
    and the moments cancel by about 3 bits, within the one pass.  In others
    the only present value among the first 32 is an outlier of 1e4, so the
    shift is off by 9000 standard deviations: the moments cancel completely
    and those tiles take the two passes.  Some have every one of the first
    32 samples missing (shift 0).  Several 64 x 64 tiles, so one-pass and
    two-pass tiles run side by side."""
    rng = np.random.default_rng(91 + self_case)
    size, nx, ny = 300, 150, 140
    x = _data(rng, size, nx, shift=1000.0)
    y = _data(rng, size, ny, shift=-250.0, scale=0.5)
    x[32:, 20:30] += 500.0                  # drift in the first tile of x
    y[32:, 5:9] -= 800.0                    # drift in the first tile of y
    x[:32, 70:74] = np.nan                  # an outlier as the shift: the
    x[3, 70:74] = 1.0e4                     # second tile of x is ill-conditioned
    x[:32, 130:134] = np.nan                # first chunk all missing (shift 0):
    y[:32, 130] = np.nan                    # ill-conditioned as well
    x[rng.random(x.shape) < 0.1] = np.nan
    y[rng.random(y.shape) < 0.1] = np.nan
    print("case gpuCor pairwise offsets/drift self=%s" % self_case, flush=True)
    if self_case:
        res = gt.gpuCor(x, use="pairwise.complete.obs")
        _check(res, x, None, "pairwise.complete.obs")
        np.testing.assert_array_equal(res["coefficients"], res["coefficients"].T)
    else:
        _check(gt.gpuCor(x, y, use="pairwise.complete.obs"), x, y, "pairwise.complete.obs")


def test_gpucor_zero_samples():
    print("case gpuCor zero samples", flush=True)
    res = gt.gpuCor(np.zeros((0, 3), np.float32), np.zeros((0, 2), np.float32))
    assert (res["pairs"] == 0).all()
    assert (res["coefficients"] == 0).all() and np.signbit(res["coefficients"]).all()
    assert np.isnan(res["ts"]).all()


def test_gpucor_larger_self_sampled_rows():
    """4096 samples x 1536 variables (6 x 6 tiles, symmetric schedule): the
    oracle checks the first 96 variables against all."""
    rng = np.random.default_rng(41)
    x = _data(rng, 4096, 1536, shift=10.0, scale=0.5)
    print("case gpuCor self 4096x1536", flush=True)
    res = gt.gpuCor(x)
    a = np.ascontiguousarray(x.T)
    pairs, corr = orc.pmcc(a[:96], a, nthreads=16)
    print("max abs error %.3g" % np.abs(res["coefficients"][:96] - corr).max(), flush=True)
    assert np.abs(res["coefficients"][:96] - corr).max() <= TOL
    np.testing.assert_array_equal(res["coefficients"], res["coefficients"].T)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
@pytest.mark.parametrize("self_case,na,nb,dim", [(True, 5788, None, 300),
                                                 (False, 4302, 4089, 200)])
def test_device_pmcc_larger_ragged(precision, self_case, na, nb, dim):
    """More tiles than CUs with a short last round (276 = 256 + 20 tiles
    symmetric, 17 x 16 = 272 = 256 + 16 cross), ragged edges: the bottom two
    tile rows (the last tiles in schedule order) against every column, and
    the first 64 rows."""
    rng = np.random.default_rng(24)
    a = _data(rng, dim, na, shift=0.25).T.copy()
    b = None if self_case else _data(rng, dim, nb, shift=-0.5).T.copy()
    print("case pmcc larger ragged %s self=%s %d %s %d" % (precision, self_case, na, nb, dim), flush=True)
    pairs, corr, ts = gt.pmcc(torch.from_numpy(a).cuda(),
                              None if self_case else torch.from_numpy(b).cuda(), precision=precision)
    torch.cuda.synchronize()
    got = corr.cpu().numpy()
    other = a if self_case else b
    x0 = (na - 1) // 256 * 256 - 256           # the last two tile rows
    for rows in (np.arange(x0, na), np.arange(0, 64)):
        _, ref = orc.pmcc(a[rows], other, nthreads=16)
        err = np.abs(got[rows] - ref).max()
        print("rows %d..%d max abs error %.3g" % (rows[0], rows[-1], err), flush=True)
        assert err <= TOL
        assert_bits_equal(ts.cpu().numpy()[rows], orc.signif(pairs.cpu().numpy()[rows], got[rows]))
    assert (pairs.cpu().numpy() == dim).all()
    if self_case:
        np.testing.assert_array_equal(got, got.T)


def test_gputtest_matches_gpusignif():
    rng = np.random.default_rng(51)
    r = np.concatenate([rng.uniform(-1, 1, 5000), [0.9995, 1.0, -1.0, 0.0, -0.0, 2.0, np.nan]])
    n = np.concatenate([rng.integers(0, 1000, 5000), [10, 10, 10, 1, 0, 5, 7]]).astype(np.float64)
    print("case gpuTtest", flush=True)
    got = gt.gpuTtest(n, r).astype(np.float32)
    want = orc.signif(n.astype(np.float32), r.astype(np.float32), cap=False)
    assert_bits_equal(got, want)
    dt = gt.ttest(torch.from_numpy(n.astype(np.float32)).cuda(),
                  torch.from_numpy(r.astype(np.float32)).cuda())
    assert_bits_equal(dt.cpu().numpy(), want)
