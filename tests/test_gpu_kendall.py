"""GPU parity of gpuCor(method = "kendall") (RgpuKendall, gt_kendall).

The count of concordant sample pairs is exact on both sides and the final
FP32 / FP64 steps are the same, so tau is compared bit for bit with the
oracle (oracle/cor_oracle.c, pinned in test_cor_cpu.py), NaN with NaN.
"""
import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402

pytestmark = pytest.mark.gpu


def assert_same(got, want):
    got = np.asarray(got, np.float64)
    nan = np.isnan(want)
    np.testing.assert_array_equal(np.isnan(got), nan)
    np.testing.assert_array_equal(got[~nan].view(np.uint64), want[~nan].view(np.uint64))


def check_gpucor(x, y=None):
    res = gt.gpuCor(x, y, method="kendall")
    a = np.ascontiguousarray(x.T, dtype=np.float32)
    b = a if y is None else np.ascontiguousarray(y.T, dtype=np.float32)
    tau, _ = orc.kendall(a, b, nthreads=8)
    assert_same(res["coefficients"], tau.T)        # R's nx x ny matrix
    assert (res["pairs"] == x.shape[0]).all()
    return res


@pytest.mark.parametrize("n,nx,ny", [(30, 5, 10), (200, 130, 70), (33, 3, 4), (65, 129, 1),
                                     (97, 2, 257), (2, 3, 3), (1, 2, 2)])
def test_kendall_matches_oracle_continuous(n, nx, ny):
    rng = np.random.default_rng(n + nx + ny)
    x = rng.standard_normal((n, nx)).astype(np.float32)
    y = rng.standard_normal((n, ny)).astype(np.float32)
    print("case kendall continuous n=%d nx=%d ny=%d" % (n, nx, ny), flush=True)
    check_gpucor(x, y)


def test_kendall_ties():
    rng = np.random.default_rng(3)
    x = rng.integers(0, 4, (150, 40)).astype(np.float32)
    y = rng.integers(0, 3, (150, 25)).astype(np.float32)
    y[:, 0] = 1.0                                   # a constant variable
    print("case kendall ties", flush=True)
    check_gpucor(x, y)


def test_kendall_nan_through_the_native_api():
    """R refuses NaN before calling RgpuKendall; the native API defines it:
    every comparison with NaN is false, so such sample pairs are discordant."""
    rng = np.random.default_rng(13)
    a = rng.integers(0, 4, (40, 150)).astype(np.float32)
    b = rng.integers(0, 3, (25, 150)).astype(np.float32)
    a[rng.random(a.shape) < 0.05] = np.nan
    a[1] = np.nan                                   # no comparable sample at all
    print("case kendall nan (device api)", flush=True)
    tau = gt.kendall(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    torch.cuda.synchronize()
    want, _ = orc.kendall(a, b, nthreads=8)
    assert_same(tau.cpu().numpy(), want)
    assert (tau.cpu().numpy()[:, 1] == -1.0).all()  # nothing concordant


def test_kendall_self_is_symmetric():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((300, 300)).astype(np.float32)
    x[:, :20] = np.round(x[:, :20])                 # some ties
    print("case kendall self 300x300", flush=True)
    res = check_gpucor(x)
    c = res["coefficients"]
    np.testing.assert_array_equal(c, c.T)
    assert (np.diag(c) == 1.0).all()


def test_kendall_count_above_2_24():
    """6000 samples: 17,997,000 sample pairs, counts above 2^24 whose FP32
    conversion rounds (the oracle follows the same steps)."""
    n = 6000
    rng = np.random.default_rng(5)
    base = np.arange(n, dtype=np.float32)
    cols = [base, base.copy(), rng.permutation(base), base + rng.standard_normal(n).astype(np.float32) * 50]
    cols[1][:1002] = cols[1][:1002][::-1]
    x = np.stack(cols, axis=1)
    print("case kendall n=6000", flush=True)
    check_gpucor(x)


def test_device_kendall_strided():
    rng = np.random.default_rng(6)
    base = torch.from_numpy(rng.standard_normal((140, 260)).astype(np.float32)).cuda()
    a = base[:, :250]
    b = base[:60, 5:255]
    out = torch.full((60, 144), 7.0, dtype=torch.float64, device="cuda")[:, :140]
    print("case kendall device strided", flush=True)
    gt.kendall(a, b, out=out)
    torch.cuda.synchronize()
    tau, _ = orc.kendall(a.cpu().numpy(), b.cpu().numpy(), nthreads=8)
    assert_same(out.cpu().numpy(), tau)


def test_kendall_nan_and_clean_variables_mixed():
    """Tiles whose variables have no tie and no NaN count with XOR on the P
    planes; a NaN (here also in the last sample, which is no row's x_i)
    sends its tiles to the P/N planes.  Both forms meet in one result."""
    rng = np.random.default_rng(21)
    a = rng.standard_normal((300, 200)).astype(np.float32)   # 3 tiles of variables
    b = rng.standard_normal((140, 200)).astype(np.float32)
    a[5, 17] = np.nan
    b[139, 199] = np.nan
    print("case kendall mixed clean / nan tiles", flush=True)
    tau = gt.kendall(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    torch.cuda.synchronize()
    want, _ = orc.kendall(a, b, nthreads=8)
    assert_same(tau.cpu().numpy(), want)
    self_tau = gt.kendall(torch.from_numpy(a).cuda())
    torch.cuda.synchronize()
    want_self, _ = orc.kendall(a, a, nthreads=8)
    assert_same(self_tau.cpu().numpy(), want_self)


def test_kendall_long_variables_global_rows():
    """8200 samples: above 8192 the planes pass reads each row's values
    through the vector memory path instead of an LDS copy of the variable.
    Continuous, tied and NaN-holding variables side by side."""
    rng = np.random.default_rng(31)
    n = 8200
    a = rng.standard_normal((4, n)).astype(np.float32)
    b = rng.standard_normal((3, n)).astype(np.float32)
    a[1] = rng.integers(0, 5, n)          # ties
    a[2, 4000] = np.nan
    b[0] = a[0] + rng.standard_normal(n).astype(np.float32) * 0.3
    b[2, n - 1] = np.nan                  # the last sample is no row's x_i
    print("case kendall n=8200 global rows", flush=True)
    tau = gt.kendall(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    torch.cuda.synchronize()
    want, _ = orc.kendall(a, b, nthreads=8)
    assert_same(tau.cpu().numpy(), want)


@pytest.mark.parametrize("same", [True, False])
def test_kendall_workspace_with_garbage_reused(same):
    """One workspace filled with 0xFF bytes, then used for two calls on
    different data: every word the count reads (the P and N planes up to the
    padded row length, the E planes, the flags, the counts) must be written
    or cleared by each call, whatever the workspace held."""
    rng = np.random.default_rng(71 if same else 72)
    cases = []
    # n = 100: 204 words of sample pairs, padded further for the chunks
    x1 = rng.standard_normal((150, 100)).astype(np.float32)            # tie-free
    x2 = rng.integers(0, 4, size=(90, 100)).astype(np.float32)         # ties everywhere
    x2[rng.random(x2.shape) < 0.05] = np.nan                           # and NaN
    for xa in (x1, x2):
        xb = None if same else rng.standard_normal((70, 100)).astype(np.float32)
        cases.append((xa, xb))
    need = max(gt_kendall_ws_bytes(xa, xb) for xa, xb in cases)
    ws = torch.full((need + 4096,), 0xFF, dtype=torch.uint8, device="cuda")
    for xa, xb in cases:
        a = torch.from_numpy(xa).cuda()
        b = None if xb is None else torch.from_numpy(xb).cuda()
        got = gt.kendall(a, b, workspace=ws).cpu().numpy()
        want, _ = orc.kendall(xa, xa if xb is None else xb, nthreads=8)
        assert_same(got, want)


def gt_kendall_ws_bytes(xa, xb):
    import ctypes
    nb = ctypes.c_size_t()
    na, n = xa.shape
    gt.lib().gt_kendall_workspace_size(na, na if xb is None else xb.shape[0], n,
                                       1 if xb is None else 0, ctypes.byref(nb))
    return nb.value


def test_kendall_invariant_under_monotone_maps():
    """Kendall's tau depends on the orderings alone: a strictly increasing
    map of a variable leaves it unchanged bit for bit, and negating a
    tie-free variable negates it exactly."""
    rng = np.random.default_rng(67)
    a = rng.standard_normal((40, 700)).astype(np.float32)
    b = rng.standard_normal((30, 700)).astype(np.float32)
    cube = a.astype(np.float64) ** 3
    a3 = cube.astype(np.float32)
    for r in range(a.shape[0]):     # the map keeps every sample distinct
        assert np.unique(a3[r]).size == np.unique(a[r]).size == a.shape[1]
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    t0 = gt.kendall(da, db).cpu().numpy()
    t1 = gt.kendall(torch.from_numpy(a3).cuda(), db).cpu().numpy()
    t2 = gt.kendall(-da, db).cpu().numpy()
    np.testing.assert_array_equal(t1, t0)
    np.testing.assert_array_equal(t2, -t0)
