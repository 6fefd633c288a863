"""Failure paths: every device allocation of an entry point made to fail in
turn (GT_FAIL_ALLOC=N fails the Nth allocation after the variable is set).

The reference raises an R error from inside distance() and friends with
longjmp, so a failed cudaMalloc leaks the buffers allocated before it
(src/distance.cpp:277-311, src/cuseful.cpp:95-102).  Here each failure must
come back as GT_ERR_ALLOC (or the message of one for the void .C exports),
with every buffer the call held released: the device's free memory after the
failed call equals what it was before.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402
from gputools_amd import _lib  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

FP = ctypes.POINTER(ctypes.c_float)
IP = ctypes.POINTER(ctypes.c_int)
SLACK = 1 << 20          # every buffer below is several MB


def free_bytes():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def fail_each_allocation(call):
    """call() -> (ok, message).  Returns the number of device allocations the
    call makes, having failed each of them once."""
    ok, msg = call()                      # warm-up: kernels loaded, pools settled
    assert ok, msg
    base = free_bytes()
    n = 1
    while True:
        os.environ["GT_FAIL_ALLOC"] = str(n)
        try:
            ok, msg = call()
        finally:
            del os.environ["GT_FAIL_ALLOC"]
        if ok:
            break
        assert "out of memory" in msg, msg
        after = free_bytes()
        # a leak lowers the free memory; more free memory than before is the
        # stream-ordered pool returning blocks it had cached
        assert base - after <= SLACK, "allocation %d: %d bytes not released" % (n, base - after)
        n += 1
        assert n < 20
    assert base - free_bytes() <= SLACK
    return n - 1


def c_call(fn, *args):
    fn(*args)
    msg = _lib.last_error()
    return msg == "", msg


def test_Rdistances_releases_on_allocation_failure():
    x = np.ascontiguousarray(wl.normal(1500, 700, seed=41))     # 4.2 MB of points, 9 MB out
    n, k = x.shape
    d = np.zeros(n * n, np.float32)
    cn, ck, cp = ctypes.c_int(n), ctypes.c_int(k), ctypes.c_float(2.0)
    for method in (b"euclidean", b"binary"):
        cm = (ctypes.c_char_p * 1)(method)
        count = fail_each_allocation(lambda: c_call(
            _lib.lib().Rdistances, x.ctypes.data_as(FP), ctypes.byref(cn), ctypes.byref(ck),
            d.ctypes.data_as(FP), cm, ctypes.byref(cp)))
        # points and distances; binary also its bit rows
        assert count == (3 if method == b"binary" else 2)
    want = orc.dist(x[:200], "binary")
    d2 = d.reshape(n, n)[:200, :200]
    ii, jj = np.tril_indices(200, -1)
    order = np.lexsort((ii, jj))
    np.testing.assert_array_equal(d2[ii[order], jj[order]], want.astype(np.float32))


def test_gt_dist_workspace_failure_returns_alloc():
    x = torch.from_numpy(wl.normal(2000, 512, seed=42)).cuda()
    out = torch.empty(2000 * 1999 // 2, device="cuda")
    L = _lib.lib()

    def call():
        s = L.gt_dist(x.data_ptr(), 2000, 512, 512, _lib.EUCLIDEAN, 2.0, _lib.PREC_BF16X3,
                      _lib.LAYOUT_PACKED, out.data_ptr(), 2000, None)
        if s != _lib.GT_OK:
            assert s == _lib.GT_ERR_ALLOC, (s, _lib.last_error())
        return s == _lib.GT_OK, _lib.last_error() if s else ""

    assert fail_each_allocation(call) == 1      # the split-precision workspace
    torch.cuda.synchronize()
    idx = np.arange(0, out.numel(), 997)
    want = orc.dist(x.cpu().numpy(), "euclidean")[idx]
    got = out.cpu().numpy()[idx].astype(np.float64)
    assert np.max(np.abs(got - want) / want) <= 1e-5


def test_Rdistclust_releases_on_allocation_failure():
    rng = np.random.default_rng(43)
    n, k = 1200, 900
    x = np.ascontiguousarray(rng.standard_normal((n, k)), dtype=np.float32)
    merge = np.zeros(2 * (n - 1), np.int32)
    order = np.zeros(n, np.int32)
    val = np.zeros(n - 1, np.float32)
    cn, ck = ctypes.c_int(n), ctypes.c_int(k)
    dm = (ctypes.c_char_p * 1)(b"euclidean")
    cm = (ctypes.c_char_p * 1)(b"average")
    count = fail_each_allocation(lambda: c_call(
        _lib.lib().Rdistclust, dm, cm, x.ctypes.data_as(FP), ctypes.byref(cn), ctypes.byref(ck),
        merge.ctypes.data_as(IP), order.ctypes.data_as(IP), val.ctypes.data_as(FP)))
    # points, the pitched matrix, the clustering workspace and its results
    assert count == 4
    sub, sup, h = orc.hclust(gt.dist(torch.from_numpy(x).cuda(), layout="full").cpu().numpy(),
                             "average")
    np.testing.assert_array_equal(val, h)


def test_rpmcc_releases_on_allocation_failure():
    a = np.ascontiguousarray(wl.normal(700, 1500, seed=44))
    b = np.ascontiguousarray(wl.normal(600, 1500, seed=45))
    na, nb, size = 700, 600, 1500
    pairs = np.zeros(na * nb, np.float32)
    corr = np.zeros(na * nb, np.float32)
    ts = np.zeros(na * nb, np.float32)
    for use in (0, 1):
        cu, cna, cnb, cs = ctypes.c_int(use), ctypes.c_int(na), ctypes.c_int(nb), ctypes.c_int(size)
        count = fail_each_allocation(lambda: c_call(
            _lib.lib().rpmcc, ctypes.byref(cu), a.ctypes.data_as(FP), ctypes.byref(cna),
            b.ctypes.data_as(FP), ctypes.byref(cnb), ctypes.byref(cs), pairs.ctypes.data_as(FP),
            corr.ctypes.data_as(FP), ts.ctypes.data_as(FP)))
        # both variable sets and the outputs; "everything" also its split terms
        assert count == (4 if use == 0 else 3)
    want = orc.pmcc(a, b, "pairwise.complete.obs")[1]
    assert np.max(np.abs(corr.reshape(na, nb) - want)) <= 1e-5


def test_gt_dist_device_full_and_kendall_release():
    x = np.ascontiguousarray(wl.normal(1400, 800, seed=46))
    L = _lib.lib()

    def full():
        ptr, pitch = ctypes.c_void_p(), ctypes.c_size_t()
        s = L.gt_dist_device_full(x.ctypes.data, 1400, 800, _lib.MANHATTAN, 2.0,
                                  ctypes.byref(ptr), ctypes.byref(pitch))
        if s == _lib.GT_OK:
            assert L.gt_free(ptr) == 0
        else:
            assert s == _lib.GT_ERR_ALLOC and not ptr.value
        return s == _lib.GT_OK, _lib.last_error() if s else ""

    assert fail_each_allocation(full) == 2
    v = np.ascontiguousarray(wl.normal(300, 2500, seed=47))
    tau = np.zeros(300 * 300, np.float64)

    def kendall():
        s = L.gt_kendall_host(v.ctypes.data, 300, None, 300, 2500, tau.ctypes.data)
        if s != _lib.GT_OK:
            assert s == _lib.GT_ERR_ALLOC
        return s == _lib.GT_OK, _lib.last_error() if s else ""

    assert fail_each_allocation(kendall) == 3
