"""gt_dist_noalloc (workspace in a static device array) against gt_dist.

Each case makes two calls in a row through the C-ABI, the second on a
smaller input, so the second reads a scratch array that still holds the
first call's workspace.  Both outputs must equal gt_dist's bit for bit (the
kernels are the same), and gt_dist's output is itself held to the oracle's
bar here (max relative error 1e-5; binary bit-exact).
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

torch = pytest.importorskip("torch")

import gputools_amd as gt  # noqa: E402
from gputools_amd import _lib  # noqa: E402
from gputools_amd import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu


def _noalloc(x, method, precision, layout):
    n, k = x.shape
    L = _lib.lib()
    if layout == "packed":
        out = torch.full((n * (n - 1) // 2,), float("nan"), dtype=torch.float32, device="cuda")
        ldd = n
    else:
        out = torch.full((n, n + 3), float("nan"), dtype=torch.float32, device="cuda")
        ldd = n + 3
    sp = torch.cuda.current_stream().cuda_stream or None
    s = L.gt_dist_noalloc(x.data_ptr(), n, k, x.stride(0), _lib.METHOD_NAMES.index(method), 2.0,
                          _lib.PRECISIONS[precision],
                          _lib.LAYOUT_PACKED if layout == "packed" else _lib.LAYOUT_FULL,
                          out.data_ptr(), ldd, sp)
    _lib.check(s)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    return got if layout == "packed" else got[:, :n]


def _reference(x, method, precision, layout):
    out = gt.dist(x, method, precision=precision, layout=layout)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _inputs(method, n, k, seed):
    if method == "binary":
        return wl.method_input("binary", n, k, seed=seed)
    return wl.mixture(n, k, seed=seed) if seed % 2 else wl.normal(n, k, seed=seed)


CASES = [
    # (method, precision, layout, first (n, k), second (n, k))
    ("euclidean", "bf16x3", "packed", (2051, 515), (700, 96)),
    ("euclidean", "fp16x3", "packed", (1500, 300), (333, 17)),
    ("euclidean", "bf16x3", "full", (900, 130), (257, 64)),
    ("binary", "fp32", "packed", (3000, 1000), (1000, 77)),
    ("binary", "fp32", "full", (700, 300), (300, 65)),
    ("euclidean", "fp32", "packed", (1200, 64), (300, 33)),
    # near the scratch limit: 3 x 8192 x 1024 BF16 terms are 48 MiB
    ("euclidean", "bf16x3", "packed", (8192, 1024), (4097, 512)),
]


@pytest.mark.parametrize("method,precision,layout,first,second", CASES)
def test_noalloc_matches_gt_dist_on_two_calls(method, precision, layout, first, second):
    print("noalloc case %s %s %s %s -> %s" % (method, precision, layout, first, second),
          flush=True)
    for seed, (n, k) in enumerate((first, second), start=3):
        xh = _inputs(method, n, k, seed)
        x = torch.from_numpy(xh).cuda()
        got = _noalloc(x, method, precision, layout)
        want = _reference(x, method, precision, layout)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, k)
        if layout == "packed" and n <= 3000:
            ref = orc.dist(xh.astype(np.float64), method)
            if method == "binary":
                np.testing.assert_array_equal(got, ref.astype(np.float32))
            else:
                err = np.max(np.abs(got.astype(np.float64) - ref) / np.maximum(ref, 1e-30))
                assert err <= 1e-5, err


def test_noalloc_rejects_beyond_the_limit_and_stays_usable():
    L = _lib.lib()
    n, k = 16384, 1024          # BF16 terms: 96 MiB
    x = torch.zeros((n, k), dtype=torch.float32, device="cuda")
    out = torch.empty(1, dtype=torch.float32, device="cuda")
    s = L.gt_dist_noalloc(x.data_ptr(), n, k, k, 0, 2.0, _lib.PREC_BF16X3, _lib.LAYOUT_PACKED,
                          out.data_ptr(), n, None)
    assert s == _lib.GT_ERR_WORKSPACE and "gt_dist_ws" in _lib.last_error()
    xs = torch.from_numpy(wl.normal(64, 8, seed=1)).cuda()
    got = _noalloc(xs, "euclidean", "bf16x3", "packed")
    np.testing.assert_array_equal(got, _reference(xs, "euclidean", "bf16x3", "packed"))
