"""GPU parity of the FP32 distance path (all six methods) against the oracle.

Tolerances (FP32 result vs FP64 oracle):
  maximum, binary                      bit-exact vs the oracle rounded to FP32
  euclidean, manhattan, canberra,
  minkowski                            max relative error <= 1e-5
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import golden_cases, load_golden
from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

RTOL = 1e-5
EXACT = ("maximum", "binary")


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def run(x, method, p=2.0, layout="packed", precision="fp32"):
    out = gt.dist(dev(x), method, p, precision=precision, layout=layout)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def assert_parity(got, want, method, rtol=RTOL):
    want32 = want.astype(np.float32)
    assert got.shape == want.shape
    if method in EXACT:
        np.testing.assert_array_equal(got, want32)
        return
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    inf = np.isinf(want)
    assert np.array_equal(got[inf], want32[inf])
    ok = ~(nan | inf)
    w = want[ok]
    g = got[ok].astype(np.float64)
    den = np.where(np.abs(w) > 0, np.abs(w), 1.0)
    err = np.max(np.abs(g - w) / den) if w.size else 0.0
    assert err <= rtol, "max rel err %.3g > %.1g" % (err, rtol)


@pytest.mark.parametrize("name", golden_cases())
def test_golden_fixtures(name):
    g = load_golden(name)
    got = run(g["x"], g["method"], g["p"])
    assert_parity(got, g["expected"], g["method"])


def test_uniform_1000x64_euclidean():
    x = wl.uniform(1000, 64, seed=1)
    assert_parity(run(x, "euclidean"), orc.dist(x, "euclidean"), "euclidean")


@pytest.mark.parametrize("method", ["euclidean", "maximum", "manhattan", "canberra", "binary",
                                    "minkowski"])
def test_all_methods_8192x256(method):
    x = wl.method_input(method, 8192, 256, seed=2)
    p = 3.0 if method == "minkowski" else 2.0
    print("case all_methods_8192x256", method, flush=True)
    got = run(x, method, p)
    if method in EXACT:   # equal to the FP64 oracle rounded to FP32
        np.testing.assert_array_equal(got, orc.dist(x, method, p).astype(np.float32))
    else:
        err, where = orc.compare_packed(x, got, method, p)
        assert err <= RTOL, (err, where)


@pytest.mark.parametrize("p", [1.0, 2.0, 0.5, 1.5, 4.0, 7.25])
def test_minkowski_powers(p):
    x = wl.normal(300, 37, seed=11)
    assert_parity(run(x, "minkowski", p), orc.dist(x, "minkowski", p), "minkowski", rtol=2e-5)


@pytest.mark.parametrize("n,k", [(2, 1), (3, 2), (127, 5), (128, 16), (129, 17), (255, 33),
                                 (257, 64), (640, 100), (1031, 3)])
@pytest.mark.parametrize("method", ["euclidean", "maximum", "manhattan", "canberra", "binary",
                                    "minkowski"])
def test_ragged_sizes(n, k, method):
    x = wl.method_input(method, n, k, seed=n * 1000 + k)
    p = 3.0 if method == "minkowski" else 2.0
    assert_parity(run(x, method, p), orc.dist(x, method, p), method)


@pytest.mark.parametrize("n,k", [(300, 4096), (130, 16384), (64, 65536), (200, 3001)])
@pytest.mark.parametrize("method", ["euclidean", "maximum", "manhattan", "canberra", "binary",
                                    "minkowski"])
def test_many_features(n, k, method):
    """Thousands of features: the chunk partials are summed in two balanced
    levels (about sqrt(k) features each), which keeps the FP32 sum within
    the 1e-5 bar at k = 65536.  Every pair against the FP64 oracle."""
    x = wl.method_input(method, n, k, seed=n + k)
    if method not in ("canberra", "binary"):
        x = x + np.float32(0.25)          # a mean, so that sums do not cancel to zero
    p = 3.0 if method == "minkowski" else 2.0
    print("case many features", method, n, k, flush=True)
    assert_parity(run(x, method, p), orc.dist(x, method, p), method)


@pytest.mark.parametrize("k", [100, 600])
def test_binary_matrix_core_large_tiles(k):
    """12000 points: the packed binary path counts on the matrix cores in
    256 x 256 tiles, two workgroups of 128 columns each (fewer points use
    128 x 128 tiles); k = 100 takes the reciprocal table, k = 600 v_rcp_f32.
    Every element written (the output starts as NaN: the tiles' line
    ownership covers each value once); whole rows across the tile edges and
    random pairs bit-exact."""
    n = 12000
    x = wl.bernoulli(n, k, seed=25)
    x[300] = 0.0
    print("case binary matrix-core 12000 x", k, flush=True)
    got = torch.full((n * (n - 1) // 2,), float("nan"), device="cuda")
    gt.dist(dev(x), "binary", out=got)
    torch.cuda.synchronize()
    assert not torch.isnan(got).any()
    rng = np.random.default_rng(3)
    jj = np.concatenate([np.repeat([0, 255, 256, 300, 11775, 11998], n),
                         rng.integers(0, n, 200000)])
    ii = np.concatenate([np.tile(np.arange(n), 6), rng.integers(0, n, 200000)])
    keep = ii != jj
    ii, jj = ii[keep], jj[keep]
    hi, lo = np.maximum(ii, jj), np.minimum(ii, jj)
    idx = lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)
    vals = got[torch.from_numpy(idx).cuda()].cpu().numpy()
    np.testing.assert_array_equal(vals, orc.dist_pairs(x, hi, lo, "binary").astype(np.float32))


@pytest.mark.parametrize("n", [3000, 11700])
@pytest.mark.parametrize("k", [480, 481])
def test_binary_quotient_forms_at_their_boundary(n, k):
    """k = 480 is the widest row the matrix-core tiles serve with the LDS
    reciprocal table (2 x 480 padded features < 1024); k = 481 pads to 512
    and takes v_rcp_f32.  n = 3000 runs the 128 x 128 tiles (every pair
    compared), n = 11700 the 256 x 256 tiles with their staggered start and a
    ragged last tile row (the output starts as NaN: every element written;
    tile-edge columns and random pairs compared).  Bit-exact."""
    x = wl.bernoulli(n, k, seed=n + k)
    x[7] = 0.0
    print("case binary quotient boundary", n, k, flush=True)
    got = torch.full((n * (n - 1) // 2,), float("nan"), device="cuda")
    gt.dist(dev(x), "binary", out=got)
    torch.cuda.synchronize()
    assert not torch.isnan(got).any()
    if n <= 3000:
        np.testing.assert_array_equal(got.cpu().numpy(), orc.dist(x, "binary").astype(np.float32))
        return
    rng = np.random.default_rng(n + k)
    edge = [0, 7, 127, 128, 255, 256, 11263, 11264, 11519, 11520, n - 2]
    jj = np.concatenate([np.repeat(edge, n), rng.integers(0, n, 100000)])
    ii = np.concatenate([np.tile(np.arange(n), len(edge)), rng.integers(0, n, 100000)])
    keep = ii != jj
    ii, jj = ii[keep], jj[keep]
    hi, lo = np.maximum(ii, jj), np.minimum(ii, jj)
    idx = lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)
    vals = got[torch.from_numpy(idx).cuda()].cpu().numpy()
    np.testing.assert_array_equal(vals, orc.dist_pairs(x, hi, lo, "binary").astype(np.float32))


@pytest.mark.parametrize("layout", ["packed", "full"])
def test_binary_beyond_the_fast_quotient(layout):
    """Past 131072 features the binary ratio takes the IEEE divide instead of
    the reciprocal-and-correction form; both are bit-exact."""
    n, k = 70, 140001
    x = wl.bernoulli(n, k, seed=23)
    x[5] = 0.0                          # no bit on in either row of (5, 6): 0
    x[6] = 0.0
    print("case binary k beyond fast quotient", layout, flush=True)
    want = orc.dist(x, "binary")
    if layout == "packed":
        assert_parity(run(x, "binary"), want, "binary")
    else:
        full = gt.dist(dev(x), "binary", layout="full")
        torch.cuda.synchronize()
        m = full.cpu().numpy()
        ii, jj = np.tril_indices(n, -1)
        order = np.lexsort((ii, jj))
        np.testing.assert_array_equal(m[ii[order], jj[order]], want.astype(np.float32))
        assert m[6, 5] == 0.0 and np.all(np.diag(m) == 0.0)


@pytest.mark.parametrize("method", ["euclidean", "manhattan", "minkowski"])
def test_many_features_minkowski_powers_and_cross(method):
    x = wl.uniform(100, 20000, seed=19)
    y = wl.uniform(37, 20000, seed=20)
    p = 1.5 if method == "minkowski" else 2.0
    got = gt.dist_cross(dev(x), dev(y), method, p)
    torch.cuda.synchronize()
    want = orc.dist_cross(x, y, method, p)
    err = np.max(np.abs(got.cpu().numpy().astype(np.float64) - want) / want)
    assert err <= (2e-5 if method == "minkowski" else RTOL), err


@pytest.mark.parametrize("method", ["euclidean", "maximum", "manhattan", "canberra", "binary",
                                    "minkowski"])
def test_full_layout_matches_packed_with_pitch(method):
    n, k = 300, 45
    x = wl.method_input(method, n, k, seed=7)
    p = 3.0 if method == "minkowski" else 2.0
    # padded leading dimensions on both sides
    xp = torch.zeros((n, k + 3), dtype=torch.float32, device="cuda")
    xp[:, :k] = dev(x)
    outp = torch.full((n, n + 7), -1.0, dtype=torch.float32, device="cuda")
    gt.dist(xp[:, :k], method, p, layout="full", out=outp[:, :n])
    packed = gt.dist(xp[:, :k], method, p)
    torch.cuda.synchronize()
    full = outp.cpu().numpy()
    assert np.all(full[:, n:] == -1.0)           # padding untouched
    m = full[:, :n]
    assert np.array_equal(m, m.T)
    assert np.all(np.diag(m) == 0.0)
    ii, jj = np.tril_indices(n, -1)
    order = np.lexsort((ii, jj))
    np.testing.assert_array_equal(m[ii[order], jj[order]], packed.cpu().numpy())
    assert_parity(packed.cpu().numpy(), orc.dist(x, method, p), method)


@pytest.mark.parametrize("method", ["euclidean", "maximum", "manhattan", "canberra", "binary",
                                    "minkowski"])
def test_full_layout_whole_tiles(method):
    """384 points with a pitch of 384: the off-diagonal 128 x 128 tiles are
    whole and store 16 bytes at a time, directly and mirrored."""
    n, k = 384, 70
    x = wl.method_input(method, n, k, seed=17)
    p = 3.0 if method == "minkowski" else 2.0
    print("case full layout whole tiles", method, flush=True)
    xt = dev(x)
    full = gt.dist(xt, method, p, layout="full").cpu().numpy()
    packed = gt.dist(xt, method, p).cpu().numpy()
    assert np.array_equal(full, full.T)
    assert np.all(np.diag(full) == 0.0)
    ii, jj = np.tril_indices(n, -1)
    order = np.lexsort((ii, jj))
    np.testing.assert_array_equal(full[ii[order], jj[order]], packed)
    assert_parity(packed, orc.dist(x, method, p), method)


def test_zero_features_give_nan():
    x = np.zeros((5, 0), np.float32)
    xt = torch.zeros((5, 0), dtype=torch.float32, device="cuda")
    for m in ("euclidean", "binary", "canberra"):
        got = gt.dist(xt, m).cpu().numpy()
        assert got.shape == (10,) and np.all(np.isnan(got))
        full = gt.dist(xt, m, layout="full").cpu().numpy()
        assert np.all(np.diag(full) == 0) and np.all(np.isnan(full[~np.eye(5, dtype=bool)]))
    assert np.all(np.isnan(orc.dist(x, "euclidean")))


def test_tiny_n():
    for n in (0, 1):
        xt = torch.ones((n, 4), dtype=torch.float32, device="cuda")
        assert gt.dist(xt, "euclidean").numel() == 0
    full = gt.dist(torch.ones((1, 4), dtype=torch.float32, device="cuda"), "manhattan",
                   layout="full").cpu().numpy()
    assert full.shape == (1, 1) and full[0, 0] == 0.0


def test_canberra_r_edge_rules_on_gpu():
    x = np.array([[1, 0, 2, 0], [3, 0, 2, 0], [0, 0, 0, 0], [0, 0, 0, 0], [-1, 0, -2, 5]],
                 np.float32)
    got = run(x, "canberra")
    want = orc.dist(x, "canberra")
    assert_parity(got, want, "canberra")
    assert math.isnan(got[orc.packed_index(5, 3, 2)])       # all 0/0 -> NA
    assert got[orc.packed_index(5, 4, 0)] == math.inf        # a = -b != 0 -> Inf


@pytest.mark.parametrize("n,k", [(384, 112), (300, 100)])
def test_canberra_chunk_modes(n, k):
    """Each 16-feature chunk takes the reciprocal path when its values allow it
    (no 0 / non-finite value: clean; zeros: max(q, 0) and zero-mask counts;
    NaN: reciprocal + NaN test) and the IEEE divide otherwise (a tiny or huge
    magnitude).  One pair's sum mixes all four; 384 rows fill whole 128-row
    tiles, 300 leaves a zero-padded one."""
    rng = np.random.default_rng(n + k)
    x = rng.standard_normal((n, k)).astype(np.float32)
    x[:, 0:16] = rng.uniform(0.1, 10.0, (n, 16))                      # clean, positive
    zb = x[:, 16:32]                                                   # zeros, no NaN
    zb[rng.random(zb.shape) < 0.3] = 0.0
    blk = x[:, 32:48]                                                  # zeros and NaN
    blk[rng.random(blk.shape) < 0.3] = 0.0
    blk[rng.random(blk.shape) < 0.02] = np.nan
    x[:, 48:64] *= np.float32(1e-36)                                   # tiny
    x[:, 52] = rng.standard_normal(n).astype(np.float32) * np.float32(1e-42)  # denormals
    x[:, 64:80] *= np.float32(1e35)                                    # huge
    x[:, 80] = np.float32(2.0 ** -100)                                 # the range bounds
    x[::2, 81] = np.float32(-(2.0 ** 100))
    zt = x[:, 96:k]                                                    # zeros, and
    zt[rng.random(zt.shape) < 0.3] = 0.0
    x[3, 96:k] = -x[7, 96:k]                                           # a = -b: Inf term
    print("case canberra chunk modes n=%d k=%d" % (n, k), flush=True)
    got = run(x, "canberra")
    want = orc.dist(x, "canberra")
    assert_parity(got, want, "canberra")
    assert got[orc.packed_index(n, 7, 3)] == math.inf


def test_binary_nonbinary_values_and_zero_rows():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((200, 300)).astype(np.float32)
    x[rng.random(x.shape) < 0.8] = 0.0
    x[5] = 0.0
    x[17] = 0.0
    x[33] = -0.0
    got = run(x, "binary")
    want = orc.dist(x, "binary")
    np.testing.assert_array_equal(got, want.astype(np.float32))
    assert got[orc.packed_index(200, 17, 5)] == 0.0


def test_Rdistances_dotC_convention():
    x = wl.normal(200, 31, seed=21)
    for method, p in (("euclidean", 2.0), ("maximum", 2.0), ("minkowski", 4.0),
                      ("no-such-method", 2.0)):
        full = gt.Rdistances(x, method, p)
        eff = method if method != "no-such-method" else "euclidean"  # getDistEnum default
        want = orc.dist(x, eff, p)
        n = 200
        ii, jj = np.tril_indices(n, -1)
        order = np.lexsort((ii, jj))
        assert np.array_equal(full, full.T) and np.all(np.diag(full) == 0)
        assert_parity(full[ii[order], jj[order]], want, eff)


def test_Rdistances_packed_dotC_convention():
    x = wl.normal(300, 17, seed=22)
    for method, p in (("canberra", 2.0), ("binary", 2.0), ("minkowski", 3.0), ("manhattan", 2.0)):
        xin = np.abs(x) + 0.1 if method == "canberra" else x
        got = gt.Rdistances_packed(xin, method, p)
        assert_parity(got, orc.dist(xin, method, p), method)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3", "fp32"])
def test_Rdistances_packed_prec_dotC_convention(precision):
    """The precision-selecting .C export, called with R's pointer convention."""
    x = wl.mixture(1500, 128, seed=23) + np.float32(3.0)
    got = gt.Rdistances_packed_prec(x, "euclidean", 2.0, precision)
    assert_parity(got, orc.dist(x, "euclidean"), "euclidean")
    d = gt.gpuDist(x, "euclidean", precision=precision)
    np.testing.assert_array_equal(d.values, got.astype(np.float64))


def test_Rdistances_packed_prec_errors():
    x = wl.normal(20, 4, seed=24)
    with pytest.raises(gt.GputoolsError, match="euclidean only"):
        gt.Rdistances_packed_prec(x, "manhattan", 2.0, "bf16x3")
    with pytest.raises(gt.GputoolsError, match="unknown precision"):
        gt.Rdistances_packed_prec(x, "euclidean", 2.0, "fp8")
    # fp32 takes every method, as Rdistances does
    got = gt.Rdistances_packed_prec(np.abs(x) + 0.1, "canberra", 2.0, "fp32")
    assert_parity(got, orc.dist(np.abs(x) + 0.1, "canberra"), "canberra")


def test_device_selection_round_trip():
    """chooseGpu / getGpuId through rsetDevice / rgetDevice
    (src/correlation.cpp:163-175): a valid id becomes current, ids outside
    0 .. count-1 fail with the reference's message."""
    count = torch.cuda.device_count()
    assert gt.chooseGpu(0) == [0]
    assert gt.getGpuId() == 0
    assert gt.chooseGpu(0.9) == [0]          # as.integer truncates
    for bad in (-1, count, count + 5):
        with pytest.raises(gt.GputoolsError, match="The gpu id number is not valid."):
            gt.chooseGpu(bad)
    assert gt.getGpuId() == 0                # a rejected id leaves the device alone
    x = wl.normal(50, 6, seed=25)             # and the library still computes
    assert_parity(gt.gpuDist(x).values, orc.dist(x, "euclidean"), "euclidean")


def test_gpuDist_mirror():
    x = wl.uniform(50, 10, seed=4)
    d = gt.gpuDist(x, "minkowski", 4)
    assert d.size == 50 and d.method == "minkowski" and d.attrs["p"] == 4
    assert d.values.dtype == np.float64
    assert_parity(d.values, orc.dist(x, "minkowski", 4.0), "minkowski")
    d = gt.gpuDist(x, "max")
    assert d.method == "maximum" and "p" not in d.attrs
    np.testing.assert_array_equal(d.values, orc.dist(x, "maximum").astype(np.float32))
    d = gt.gpuDist(x, "euclidian")
    assert d.method == "euclidean"
    # a single column vector
    v = gt.gpuDist(np.arange(6, dtype=np.float64), "manhattan")
    assert v.size == 6 and v[5, 0] == 5.0


def test_device_full_distanceLeaveOnGpu():
    x = wl.normal(333, 20, seed=8)
    ptr, pitch = gt.dist_device_full(x, "manhattan")
    try:
        assert pitch >= 333 * 4 and pitch % 4 == 0
        host = np.empty((333, pitch // 4), np.float32)
        hip = ctypes.CDLL("libamdhip64.so.7")   # the runtime torch already loaded
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert hip.hipMemcpy(host.ctypes.data, ptr, host.nbytes, 2) == 0  # device -> host
    finally:
        gt.free_device(ptr)
    m = host[:, :333]
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
    ii, jj = np.tril_indices(333, -1)
    order = np.lexsort((ii, jj))
    assert_parity(m[ii[order], jj[order]], orc.dist(x, "manhattan"), "manhattan")


@pytest.mark.parametrize("method", ["euclidean", "maximum", "manhattan", "canberra", "binary",
                                    "minkowski"])
def test_cross_distances(method):
    a = wl.method_input(method, 150, 29, seed=12)
    b = wl.method_input(method, 77, 29, seed=13)
    p = 3.0 if method == "minkowski" else 2.0
    got = gt.dist_cross(dev(a), dev(b), method, p).cpu().numpy()
    want = orc.dist_cross(a, b, method, p)
    assert_parity(got.ravel(), want.ravel(), method)


def test_deterministic():
    x = dev(wl.normal(2000, 100, seed=14))
    a = gt.dist(x, "euclidean").cpu().numpy()
    b = gt.dist(x, "euclidean").cpu().numpy()
    assert np.array_equal(a, b)


def test_nonzero_stream():
    x = dev(wl.normal(700, 40, seed=15))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        got = gt.dist(x, "manhattan")
    s.synchronize()
    assert_parity(got.cpu().numpy(), orc.dist(x.cpu().numpy(), "manhattan"), "manhattan")


def test_mixture_65536x512_fp32_sampled():
    """Largest single-GPU shape: sampled pairs (random and same-centre) vs the oracle."""
    n, k = 65536, 512
    x = wl.mixture(n, k, seed=3)
    print("case mixture_65536x512 fp32", flush=True)
    got = gt.dist(dev(x), "euclidean")
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    ii = rng.integers(0, n, 100000)
    jj = rng.integers(0, n, 100000)
    same = rng.integers(0, n - 64, 20000)
    ii = np.concatenate([ii, same + 64])      # i and i+64 share a centre
    jj = np.concatenate([jj, same])
    keep = ii != jj
    ii, jj = ii[keep], jj[keep]
    hi, lo = np.maximum(ii, jj), np.minimum(ii, jj)
    idx = lo * n - lo * (lo + 1) // 2 + (hi - lo - 1)
    vals = got[torch.from_numpy(idx).cuda()].cpu().numpy().astype(np.float64)
    want = orc.dist_pairs(x, ii, jj, "euclidean")
    err = np.max(np.abs(vals - want) / want)
    assert err <= RTOL, err
